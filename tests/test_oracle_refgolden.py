"""The oracle (glibc flavour: the host libm, as the reference calls it) against vectors the
reference's own routines wrote (tests/golden/ref_tinybox.npz, made by
tests/golden/make_ref_vectors.py from the reference's `epsr` program on a reduced copy of its
EPSR-test case: 4 DMSO + 20 water in a 16 A box).  Every comparison is bit for bit.

What each test pins that the mathematical checks leave open:
  cells          the cell lattice and the order of molecules in cells (the MC chain order)
  fmolec1        the ledger: the slot-0 quirk (store loop from the second neighbour), the
                 energy != 0 filter, partner order, entry numbering, ftest / ftestb
  molfit         ran1 consumption order, the z/y/x rotation and translation order, the
                 Metropolis operands, step adaptation, ledger write-back, fadds
  atogen         random-origin selection (ran1 use, forced under-represented types) and
                 the normalisation of the sampled histogram
  fit_expansion  the Metropolis fit's ran1 use, smoop cadence and clamp
  generate_ep    trapezoid integration, the pressure factor and the gradients
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as po
from tests import refgolden as RG

F32 = np.float32


@pytest.fixture(scope="module")
def L():
    return po.lib("glibc")


def test_fixture_records_this_hosts_libm():
    d = RG.load()
    import platform
    assert str(d["meta.glibc"]) == "-".join(platform.libc_ver()), \
        "vectors were written against another glibc; transcendentals may differ"


def _cells(L, cap, st):
    cl = po.Cells()
    L.ora_cells_init(C.byref(cl), int(cap["nmolcellin"]), po.Pf(st.xadd), po.Pf(st.yadd), po.Pf(st.zadd),
                     int(cap["nmol"]), F32(cap["rmax"]))
    return cl


def test_cells_match_the_reference(L):
    cap = RG.group("cells.0")
    st = RG.State(RG.group("fmolec1.0"), RG.group("fmolec1.0.tables"))
    cl = _cells(L, cap, st)
    try:
        assert (cl.ncelldiva, cl.ncelldivb, cl.ncelldivc) == (int(cap["ncelldiva"]), int(cap["ncelldivb"]),
                                                              int(cap["ncelldivc"]))
        assert cl.ncell == int(cap["ncell"]) and cl.maxcellno == int(cap["maxcellno"])
        assert cl.ncellneighb == int(cap["ncellneighb"])
        nmic = np.ctypeslib.as_array(cl.nmolsincell, shape=(cl.ncell,))
        assert np.array_equal(nmic, cap["nmolsincell"])
        mic = np.ctypeslib.as_array(cl.molsincell, shape=(cl.ncell * cl.maxcellno,)).reshape(cl.ncell, -1)
        for ic in range(cl.ncell):
            n = int(nmic[ic])
            assert np.array_equal(mic[ic, :n] + 1, cap["molsincell"][1:n + 1, ic])
        assert np.array_equal(np.ctypeslib.as_array(cl.molcellno, shape=(st.nmol,)) + 1, cap["molcellno"])
    finally:
        L.ora_cells_free(C.byref(cl))


@pytest.mark.parametrize("k", [0, 1])
def test_fmolec1_ledger_matches_the_reference(L, k):
    cap = RG.group(f"fmolec1.{k}")
    st = RG.State(cap, RG.group(f"fmolec1.{k}.tables"))
    cl = _cells(L, RG.group("cells.0"), RG.State(RG.group("fmolec1.0"), RG.group("fmolec1.0.tables")))
    eintra = RG.self_energies(cap, st.nmol)
    led = po.Ledger()
    ft, fb, f64 = C.c_float(), C.c_float(), C.c_double()
    L.ora_fmolec1(C.byref(st.box), C.byref(st.centres), C.byref(st.tables), C.byref(cl), po.Pf(eintra),
                  C.byref(led), C.byref(ft), C.byref(fb), C.byref(f64))
    try:
        assert led.nnbrefmax == int(cap["nnbrefmax"])
        assert led.nergpair == int(cap["nergpair"])
        got = RG.ledger_arrays(led, st.nmol)
        want = RG.record_rows(cap, st.nmol)
        assert np.array_equal(got[0], want[0]), "entries per molecule"
        for im in range(st.nmol):
            assert np.array_equal(got[1][im], want[1][im]), f"partners of molecule {im + 1}"
            assert np.array_equal(got[2][im], want[2][im]), f"entry numbers of molecule {im + 1}"
        assert np.array_equal(RG.f32_bits(got[4]), RG.f32_bits(want[4])), "pair energies"
        assert np.float32(ft.value) == F32(cap["ftest"]) and np.float32(fb.value) == F32(cap["ftestb"])
    finally:
        L.ora_ledger_free(C.byref(led))


def _rigid_cycles():
    out = []
    for k in range(RG.count("molfit")):
        cap = RG.group(f"molfit.{k}")
        nisum, intra, rotfreq = int(cap["nisum"]), int(cap["intra"]), int(cap["rotfreq"])
        if nisum < intra and nisum % rotfreq != 0:
            out.append(k)
    return out


def test_fixture_has_rigid_rotation_and_shake_cycles():
    rigid = _rigid_cycles()
    assert len(rigid) >= 3
    kinds = set()
    for k in range(RG.count("molfit")):
        cap = RG.group(f"molfit.{k}")
        n, i, r = int(cap["nisum"]), int(cap["intra"]), int(cap["rotfreq"])
        kinds.add("shake" if n >= i else ("rot" if n % r == 0 else "rigid"))
    assert kinds == {"shake", "rot", "rigid"}


@pytest.mark.parametrize("k", _rigid_cycles())
def test_molfit_rigid_cycle_matches_the_reference(L, k):
    cap = RG.group(f"molfit.{k}")
    cin, cout = RG.group(f"molfit.{k}.in"), RG.group(f"molfit.{k}.out")
    st = RG.State(cin, RG.group("fmolec1.0.tables"))
    cl = _cells(L, RG.group("cells.0"), RG.State(RG.group("fmolec1.0"), RG.group("fmolec1.0.tables")))
    led = RG.ledger_from_record(cin, st.nmol)
    nt = st.nmoltype
    wr = np.ascontiguousarray(cin["wholerot"], F32).copy()
    wt = np.ascontiguousarray(cin["wholetrans"], F32).copy()
    nwh, nmv = np.zeros(2 * nt, np.int32), np.zeros(2 * nt, np.int32)
    nsr = np.ascontiguousarray(cin["nstepreject"], np.int32).copy()
    mc = po.MCState(po.Pf(wr), po.Pf(wt), po.Pi(nwh), po.Pi(nmv), po.Pi(nsr), F32(cap["simt"]), F32(cap["stepr"]),
                    F32(cap["stepm"]), F32(cap["rejrate"]), int(cap["ireset"]))
    rng = RG.ran1(cin["ran1"])
    fadds, nacc = C.c_float(F32(cap["fadds"])), C.c_int()
    L.ora_molfit(C.byref(st.box), po.Pf(st.xa), po.Pf(st.ya), po.Pf(st.za), C.byref(st.centres),
                 C.byref(st.tables), C.byref(cl), C.byref(led), C.byref(mc), C.byref(rng), C.byref(fadds),
                 C.byref(nacc))
    assert np.array_equal(rng.as_array(), cout["ran1"]), "ran1 state after the cycle"
    for key in ("xa", "ya", "za", "xadd", "yadd", "zadd"):
        assert np.array_equal(RG.f32_bits(getattr(st, key)), RG.f32_bits(cout[key])), key
    assert np.array_equal(RG.f32_bits(wr), RG.f32_bits(cout["wholerot"]))
    assert np.array_equal(RG.f32_bits(wt), RG.f32_bits(cout["wholetrans"]))
    assert np.array_equal(nsr, cout["nstepreject"])
    assert np.array_equal(nwh.reshape(nt, 2).T, cout["nwhole"][1:3])
    assert np.array_equal(nmv.reshape(nt, 2).T, cout["nmovem"][1:3])
    erg = np.ctypeslib.as_array(led.ergpair, shape=(led.nergpair,))
    assert np.array_equal(RG.f32_bits(erg), RG.f32_bits(cout["ergpair"])), "ledger after the cycle"
    assert F32(fadds.value) == F32(cout["fadds"])
    assert 0 < nacc.value < st.nmol, "a cycle with both outcomes"


def test_atogen_random_origins_match_the_reference(L):
    cin, cout = RG.group("atogen.0.in"), RG.group("atogen.0.out")
    st = RG.State(cin, RG.group("fmolec1.0.tables"))
    ycoh = np.ascontiguousarray(cin["ycoh"], F32)
    npt1 = ycoh.shape[0] - 1
    ldw = npt1 + 1
    assert 0 < int(cin["ngrsamples"]) < st.nmol, "the random-origin branch"
    wav = np.zeros(2 * st.npairs * ldw, F32)
    px, pm = np.zeros(st.npairs, F32), np.zeros(st.npairs, F32)
    rng = RG.ran1(cin["ran1"])
    ir = np.ascontiguousarray(cin["irefpcof"] - 1, np.int32)
    rminex = np.ascontiguousarray(cin["rminex"], F32)
    ngr = L.ora_atogen(C.byref(st.box), C.byref(st.centres), npt1, ldw, po.Pf(ycoh), po.Pi(ir), po.Pf(rminex),
                       int(cin["ngrsamples"]), C.byref(rng), po.Pf(wav), po.Pf(px), po.Pf(pm), None, None)
    assert ngr == int(cout["ngrsamples"])
    assert np.array_equal(rng.as_array(), cout["ran1"])
    got = wav.reshape(2, st.npairs, ldw)[:, :, 1:]
    want = np.transpose(cout["wav"], (0, 2, 1))
    assert np.array_equal(RG.f32_bits(got), RG.f32_bits(want))
    assert np.array_equal(RG.f32_bits(px), RG.f32_bits(cout["potex"]))


@pytest.mark.parametrize("k", range(3))
def test_fit_expansion_matches_the_reference(L, k):
    cin, cout = RG.group(f"fit.{k}.in"), RG.group(f"fit.{k}.out")
    npt, nterms = int(cin["npt"]), int(cin["nterms"])
    yin = np.ascontiguousarray(cin["yin"], F32)
    expr = np.ascontiguousarray(cin["expr"], F32)
    yout = np.ascontiguousarray(cin["yout"], F32).copy()
    wt = np.ascontiguousarray(cin["wt"], F32).copy()
    wo = np.ascontiguousarray(cin["wtold"], F32).copy()
    st, ac, rj, cs = (np.array([F32(cin[key])], F32) for key in ("stepr", "acceptr", "rejrat", "chisqnew"))
    rng = RG.ran1(cin["ran1"])
    ntries = C.c_int()
    L.ora_fit_expansion(po.Pf(yin), po.Pf(expr), po.Pf(yout), po.Pf(wt), po.Pf(wo), po.Pf(st), po.Pf(ac),
                        po.Pf(rj), po.Pf(cs), int(cin["nmin"]), npt, int(cin["ncfirst"]), nterms, int(cin["nits"]),
                        int(cin["nfreq"]), C.byref(rng), C.byref(ntries))
    assert np.array_equal(rng.as_array(), cout["ran1"]), "ran1 state after the fit"
    for key, arr in (("wt", wt), ("wtold", wo), ("yout", yout)):
        assert np.array_equal(RG.f32_bits(arr[1:]), RG.f32_bits(cout[key][1:])), key
    for key, arr in (("stepr", st), ("acceptr", ac), ("rejrat", rj), ("chisqnew", cs)):
        assert RG.f32_bits(arr[0]) == RG.f32_bits(F32(cout[key])), key


@pytest.mark.parametrize("k", range(2))
def test_generate_ep_matches_the_reference(L, k):
    cin, cout = RG.group(f"genep.{k}.in"), RG.group(f"genep.{k}.out")
    assert not bool(cin["gaussian"]) and not np.any(cin["npgauss"])
    npairs = int(RG.group("topo")["npairs"])
    npt, npt2, nterms = int(cin["npt"]), int(cin["npt2"]), int(cin["ncoeffp"])
    ld = npt + 1
    pc = np.ascontiguousarray(cin["pcoefp"], F32).copy()
    ldp = pc.shape[1]
    ir = np.ascontiguousarray(cin["irefpcof"] - 1, np.int32)
    out = {key: np.zeros(npairs * ld, F32) for key in ("xavs", "xavs1", "xavtot", "xavtot1")}
    ea = C.c_float()
    L.ora_generate_ep(npt, npt2, po.Pf(np.ascontiguousarray(cin["x"], F32)),
                      po.Pf(np.ascontiguousarray(cin["xstar"], F32)), po.Pf(np.ascontiguousarray(cin["wfac"], F32)),
                      npairs, po.Pi(ir), nterms, po.Pf(pc.reshape(-1)), ldp, F32(cin["pdmin"]), F32(cin["pdstep"]),
                      F32(cin["pdmax"]), F32(cin["psigma2"]), F32(cin["rho"]), F32(cin["rbroad"]), F32(cin["dmax"]),
                      F32(cin["absolute_energy"]), F32(cin["erequnit"]), F32(cin["ereqstep"]),
                      po.Pf(np.ascontiguousarray(cin["ecoul"], F32).reshape(-1)), po.Pf(out["xavs"]),
                      po.Pf(out["xavs1"]), po.Pf(out["xavtot"]), po.Pf(out["xavtot1"]), ld, C.byref(ea))
    assert np.array_equal(RG.f32_bits(pc[:, 1:nterms + 1]), RG.f32_bits(cout["pcoefp"][:, 1:nterms + 1])), \
        "coefficients after the pressure factor"
    for key in out:
        got = out[key].reshape(npairs, ld)[:, 1:npt + 1]
        want = cout[key][:, 1:npt + 1]
        assert np.array_equal(RG.f32_bits(got), RG.f32_bits(want)), key


@pytest.mark.parametrize("k", [0, 1])
def test_en_mole_intra_gives_the_references_self_energies(L, k):
    # fmolec1 stores en_mole_intra of every molecule as its self entry (weights
    # tempang = tempangreq, tempdih = tempdihreq * tempb, ato_manipulation_rev4.py:2416-2419)
    cap = RG.group(f"fmolec1.{k}")
    tab = RG.group(f"fmolec1.{k}.tables")
    st = RG.State(cap, tab)
    it = RG.intra_from_record(st, tab)
    want = RG.self_energies(cap, st.nmol)
    got = np.array([L.ora_en_mole_intra(C.byref(st.box), C.byref(st.tables), C.byref(it), im, None, None, None)
                    for im in range(st.nmol)], F32)
    assert np.array_equal(RG.f32_bits(got), RG.f32_bits(want))
    assert np.all(want != 0.0)
    fb = F32(0.0)
    for e in got:
        fb = F32(fb + e)
    assert F32(fb / F32(st.nmol)) == F32(cap["ftestb"])


def _run_full(L, k):
    cap = RG.group(f"molfit.{k}")
    cin = RG.group(f"molfit.{k}.in")
    tab = RG.group("fmolec1.0.tables")
    st = RG.State(cin, tab)
    it = RG.intra_from_record(st, tab)
    cl = _cells(L, RG.group("cells.0"), RG.State(RG.group("fmolec1.0"), tab))
    led = RG.ledger_from_record(cin, st.nmol)
    nt = st.nmoltype
    arr = {key: np.ascontiguousarray(cin[key], F32).copy() for key in ("wholerot", "wholetrans", "internrot",
                                                                     "interntrans")}
    cnt = {key: np.zeros(2 * nt, np.int32) for key in ("nwhole", "nmovem", "ninter", "nmovei")}
    nsr = np.ascontiguousarray(cin["nstepreject"], np.int32).copy()
    mc = po.MCFull(po.Pf(arr["wholerot"]), po.Pf(arr["wholetrans"]), po.Pf(arr["internrot"]),
                   po.Pf(arr["interntrans"]), po.Pi(cnt["nwhole"]), po.Pi(cnt["nmovem"]), po.Pi(cnt["ninter"]),
                   po.Pi(cnt["nmovei"]), po.Pi(nsr), F32(cap["simt"]), F32(cap["stepr"]), F32(cap["stepm"]),
                   F32(cap["stepri"]), F32(cap["stepmi"]), F32(cap["rejrate"]), F32(cap["tempangwt"]),
                   F32(cap["tempdihwt"]), int(cap["ireset"]))
    rng = RG.ran1(cin["ran1"])
    fadds, faddbs = C.c_float(F32(cap["fadds"])), C.c_float(F32(cap["faddbs"]))
    nisum = C.c_int(int(cap["nisum"]))
    L.ora_molfit_full(C.byref(st.box), po.Pf(st.xa), po.Pf(st.ya), po.Pf(st.za), C.byref(st.centres),
                      C.byref(st.tables), C.byref(it), C.byref(cl), C.byref(led), C.byref(mc), C.byref(rng),
                      C.byref(fadds), C.byref(faddbs), C.byref(nisum), int(cap["intra"]), int(cap["rotfreq"]))
    return st, led, arr, cnt, nsr, rng, fadds, faddbs, nisum


@pytest.mark.parametrize("k", range(6))
def test_molfit_full_cycle_matches_the_reference(L, k):
    # every recorded cycle: the shake (nisum >= intra), whole-molecule cycles and the cycle
    # with internal rotations (nisum a multiple of rotfreq)
    st, led, arr, cnt, nsr, rng, fadds, faddbs, nisum = _run_full(L, k)
    cout = RG.group(f"molfit.{k}.out")
    nt = st.nmoltype
    assert np.array_equal(rng.as_array(), cout["ran1"]), "ran1 state after the cycle"
    for key in ("xa", "ya", "za", "xadd", "yadd", "zadd"):
        assert np.array_equal(RG.f32_bits(getattr(st, key)), RG.f32_bits(cout[key])), key
    for key, a in arr.items():
        assert np.array_equal(RG.f32_bits(a), RG.f32_bits(cout[key])), key
    for key, a in cnt.items():
        assert np.array_equal(a.reshape(nt, 2).T, cout[key][1:3]), key
    assert np.array_equal(nsr, cout["nstepreject"])
    erg = np.ctypeslib.as_array(led.ergpair, shape=(led.nergpair,))
    assert np.array_equal(RG.f32_bits(erg), RG.f32_bits(cout["ergpair"])), "ledger after the cycle"
    assert F32(fadds.value) == F32(cout["fadds"]) and F32(faddbs.value) == F32(cout["faddbs"])
    assert nisum.value == int(cout["nisum"])
