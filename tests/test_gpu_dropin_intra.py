"""The drop-in molfit on the reference's own EPSR-test box (200 DMSO + 400 water read from its
.ato, .pcof and .EPSR.inp, DMSO with two rotational groups, stepri 0.3, rotfreq 5) restarted
from the .ato's ran1 state and step sizes: a whole-molecule cycle, the cycle with internal
rotations (nisum = 5) and the shake (nisum = intra), each bit for bit against the correctly
rounded oracle's complete molfit from the same state -- coordinates, ran1 state, the four
step sizes, nstepreject and the ledger."""

import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as po
from tests import helpers as H

F32 = np.float32
pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, F32).view(np.int32)


def test_epsr_test_box_every_cycle_kind():
    from epsr_python_translation_amd import tables as T
    from epsr_python_translation_amd.common import clustercom, epsr16, gpu, molecom, potential_par, ran1com, setup
    from epsr_python_translation_amd.library import ato_manipulation_rev4 as A
    from epsr_python_translation_amd.library import molfit as M
    from epsr_python_translation_amd.programs import epsr_standalone_rev3 as P
    print("drop-in molfit on the EPSR test box: set-up", flush=True)
    a, _, kf = H.epsr_test_files()
    s = H.system("epsr_test")
    g = setup.load(s)
    ran1com.set_state(np.asarray(a.ran1, np.int32))
    P.initpt_truncation()
    P.inita_foqols()
    P.generate_rp()
    P.generate_ep(False)
    A.fmolec1(0, 0, F32(0), F32(0), F32(0), F32(0))
    # the oracle on the same tables, coordinates, cells and intramolecular energies
    ld = g.ld
    cols = lambda x: np.ascontiguousarray(x[:ld, 1:s.npairs + 1].T).reshape(-1)
    tab = {k: cols(potential_par.__dict__[k]) for k in ("xavs", "xavs1", "ecoul", "ecoul1")}
    tt = po.Tables(ld, g.npt2, po.Pf(g.xstar), po.Pf(tab["xavs"]), po.Pf(tab["xavs1"]), po.Pf(tab["ecoul"]),
                   po.Pf(tab["ecoul1"]))
    L = po.lib("cr")
    ob = po.OracleBox(s)
    grid = T.Grid(x=epsr16.x, xstar=epsr16.xstar, ycoh=epsr16.ycoh, r2=None, q=epsr16.q, npt=epsr16.npt,
                  npt1=epsr16.npt1, npt2=epsr16.npt2, nptmin=epsr16.nptmin, nptmax=epsr16.nptmax, nq=epsr16.nq)
    cl = po.Cells()
    L.ora_cells_init(C.byref(cl), s.nmolcell, po.Pf(ob.xadd), po.Pf(ob.yadd), po.Pf(ob.zadd), s.nmol,
                     A._sepmax(grid))
    it = H.oracle_intra(s, ld, *H.fmolec1_weights(s))
    ei = np.array([L.ora_en_mole_intra(C.byref(ob.box), C.byref(tt), C.byref(it), im, None, None, None)
                   for im in range(s.nmol)], F32)
    led = po.Ledger()
    ft, fb, f64 = C.c_float(), C.c_float(), C.c_double()
    L.ora_fmolec1(C.byref(ob.box), C.byref(ob.centres), C.byref(tt), C.byref(cl), po.Pf(ei), C.byref(led),
                  C.byref(ft), C.byref(fb), C.byref(f64))
    nmt = s.nmoltype
    r = s.mc_run
    arr = {k: np.ascontiguousarray(r[k], F32).copy() for k in ("wholerot", "wholetrans", "internrot", "interntrans")}
    cnt = {k: np.zeros(2 * nmt, np.int32) for k in ("nwhole", "nmovem", "ninter", "nmovei")}
    nsr = np.ascontiguousarray(r["nstepreject"], np.int32).copy()
    mcs = po.MCFull(po.Pf(arr["wholerot"]), po.Pf(arr["wholetrans"]), po.Pf(arr["internrot"]),
                    po.Pf(arr["interntrans"]), po.Pi(cnt["nwhole"]), po.Pi(cnt["nmovem"]), po.Pi(cnt["ninter"]),
                    po.Pi(cnt["nmovei"]), po.Pi(nsr), F32(s.simt), F32(r["stepr"]), F32(r["stepm"]), F32(r["stepri"]),
                    F32(r["stepmi"]), F32(epsr16.rejrate), F32(r["tempangreq"]), F32(r["tempdihreq"]), 0)
    rng = po.Ran1.from_array(np.asarray(a.ran1, np.int32))
    intra, rotfreq = int(r["intra"]), int(r["rotfreq"])
    fadds, faddbs = F32(0.0), F32(0.0)
    ofa, ofb = C.c_float(0.0), C.c_float(0.0)
    for nisum in (4, 5, intra):
        kind = "shake" if nisum >= intra else ("rotation" if nisum % rotfreq == 0 else "whole-molecule")
        print(f"cycle nisum={nisum} ({kind})", flush=True)
        fadds, faddbs, ngot = M.molfit(fadds, faddbs, nisum, intra, rotfreq, 0)
        onis = C.c_int(nisum)
        L.ora_molfit_full(C.byref(ob.box), po.Pf(ob.xa), po.Pf(ob.ya), po.Pf(ob.za), C.byref(ob.centres),
                          C.byref(tt), C.byref(it), C.byref(cl), C.byref(led), C.byref(mcs), C.byref(rng),
                          C.byref(ofa), C.byref(ofb), C.byref(onis), intra, rotfreq)
        n, m = s.natom, s.nmol
        assert np.array_equal(ran1com.state(), rng.as_array()), f"{kind}: ran1"
        for key, com in (("xa", molecom.xa[1:n + 1]), ("ya", molecom.ya[1:n + 1]), ("za", molecom.za[1:n + 1]),
                         ("xadd", clustercom.xadd[1:m + 1]), ("yadd", clustercom.yadd[1:m + 1]),
                         ("zadd", clustercom.zadd[1:m + 1])):
            assert np.array_equal(_bits(com), _bits(getattr(ob, key))), f"{kind}: {key}"
        for key in arr:
            assert np.array_equal(_bits(clustercom.__dict__[key][1:nmt + 1]), _bits(arr[key])), f"{kind}: {key}"
        assert np.array_equal(clustercom.nstepreject[1:m + 1], nsr), f"{kind}: nstepreject"
        erg = np.ctypeslib.as_array(led.ergpair, shape=(led.nergpair,))
        ex = gpu.ledger().export()
        assert np.array_equal(_bits(ex["ergpair"]), _bits(erg)), f"{kind}: ledger"
        assert F32(fadds) == F32(ofa.value) and F32(faddbs) == F32(ofb.value), f"{kind}: fadds"
        assert ngot == onis.value
        if kind == "rotation":
            assert cnt["ninter"][0] > 0 and cnt["ninter"][1] > 0, "DMSO rotations attempted and rejected"
        if kind == "shake":
            assert cnt["nmovei"][0] > 0
