"""atogen oracle pinned by brute force and by exact integer invariants.

The brute force restates the pair loop with numpy float32 element-wise operations
(IEEE single, one rounding per operation, as the reference's scalar code), over every
ordered pair of atoms of distinct molecules, and bins with NINT(d/stepcl)+2
(src/epsr26/library/ato_manipulation_rev4.py:3780-3837)."""

import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as po
from tests import helpers as H

F32 = np.float32


def _nint(x):
    x = x.astype(np.float64)
    t = np.trunc(x)
    f = x - t
    return (t + (f >= 0.5) - (f <= -0.5)).astype(np.int64)


def _brute_counts(s, g, intra=False):
    amol = np.repeat(np.arange(s.nmol), s.nal - s.naf + 1)
    px = (s.xa + s.xadd[amol]).astype(F32)
    py = (s.ya + s.yadd[amol]).astype(F32)
    pz = (s.za + s.zadd[amol]).astype(F32)
    stepcl = F32(F32(s.cellst) / F32(s.dmax))
    cnt = np.zeros((s.npairs, g.npt1 + 1), np.int64)
    for ia in range(s.natom):
        if intra:
            sel = (amol == amol[ia]) & (np.arange(s.natom) > ia)
            dx = ((s.xa[ia] - s.xa[sel]) * F32(1.0)).astype(F32)
            dy = ((s.ya[ia] - s.ya[sel]) * F32(1.0)).astype(F32)
            dz = ((s.za[ia] - s.za[sel]) * F32(1.0)).astype(F32)
        else:
            sel = amol > amol[ia]
            dx = (px[ia] - px[sel]).astype(F32)
            dy = (py[ia] - py[sel]).astype(F32)
            dz = (pz[ia] - pz[sel]).astype(F32)
        dx = (dx - _nint(dx).astype(F32)).astype(F32)
        dy = (dy - _nint(dy).astype(F32)).astype(F32)
        dz = (dz - _nint(dz).astype(F32)).astype(F32)
        d = np.sqrt(((dx * dx + dy * dy).astype(F32) + (dz * dz).astype(F32)).astype(F32)).astype(F32)
        iz = _nint((d / stepcl).astype(F32)) + 2
        ic = s.npairref[s.itype[ia], s.itype[sel]]
        ok = (iz > 1) & (iz <= g.npt1)
        np.add.at(cnt, (ic[ok], iz[ok]), 1)
    return cnt


@pytest.mark.parametrize("name", ["tiny_11_10_0", "tiny_12_6_2"])
def test_oracle_counts_equal_brute_force(name):
    s, g = H.system(name), H.grid(name)
    ldw = g.npt1 + 1
    cnt = np.zeros(2 * s.npairs * ldw, np.int64)
    pc = np.zeros(s.npairs, np.int64)
    ob = po.OracleBox(s)
    rng = po.Ran1.seeded(-7)
    rm = np.zeros(s.npairs, F32)
    ngr = po.lib().ora_atogen_counts(C.byref(ob.box), C.byref(ob.centres), g.npt1, ldw,
                                     po.Pi(np.arange(s.npairs, dtype=np.int32)), po.Pf(rm), 0, C.byref(rng),
                                     cnt.ctypes.data_as(po.i64p), pc.ctypes.data_as(po.i64p))
    assert ngr == s.nmol                               # ngrsamples<=0 -> 1000 -> nmol
    cnt = cnt.reshape(2, s.npairs, ldw)
    np.testing.assert_array_equal(cnt[1], _brute_counts(s, g))
    np.testing.assert_array_equal(cnt[0], _brute_counts(s, g, intra=True))
    # one ran1 per candidate origin molecule (im = 1..nmol-1 when every molecule is used)
    assert rng.mcall == s.nmol - 1


def test_oracle_atogen_invariants_and_normalisation():
    name = "tiny_13_12_3"
    s, g = H.system(name), H.grid(name)
    r = H.oracle_atogen(name)
    wav = r["wav"].reshape(2, s.npairs, r["ldw"])
    # full mode: every molecule is an origin, csampled = atoms of each type
    np.testing.assert_array_equal(r["csampled"], np.bincount(s.itype, minlength=s.ntype).astype(F32))
    # counts recovered from the normalised histogram are integers summing to the
    # number of in-range pairs
    for ic in range(s.npairs):
        it, jt = s.n1extr[ic], s.n2extr[ic]
        cprod2 = F32(r["csampled"][it] * s.cfrac[jt])
        if it != jt:
            cprod2 = F32(cprod2 + F32(r["csampled"][jt] * s.cfrac[it]))
        k = wav[1, ic, 1:g.npt1 + 1].astype(np.float64) * float(cprod2) / 2.0
        assert np.allclose(k, np.round(k), atol=1e-3)


def test_sampled_origins_consume_one_ran1_per_candidate():
    name = "tiny_14_40_0"
    s = H.system(name)
    r = H.oracle_atogen(name, ngrsamples=7, idum=-99)
    assert r["ngr"] == 7
    assert r["ran1"][0] == s.nmol        # mcall: one draw per molecule 1..nmol


@pytest.mark.parametrize("name,ngr", [("c0", 0), ("c0", 150), ("tiny_14_40_0", 7)])
def test_oracle_omp_counts_equal_sequential(name, ngr):
    """The all-cores oracle (origin molecules across OpenMP threads, integer histograms
    summed at the end) counts exactly what the sequential one does, ran1 included."""
    s, g = H.system(name), H.grid(name)
    ldw = g.npt1 + 1
    rm = np.full(s.npairs, 2.2, F32)
    out = []
    for flavour in ("glibc", "omp"):
        cnt = np.zeros(2 * s.npairs * ldw, np.int64)
        pc = np.zeros(s.npairs, np.int64)
        ob = po.OracleBox(s)
        rng = po.Ran1.seeded(-11)
        n = po.lib(flavour).ora_atogen_counts(C.byref(ob.box), C.byref(ob.centres), g.npt1, ldw,
                                              po.Pi(np.arange(s.npairs, dtype=np.int32)), po.Pf(rm), ngr,
                                              C.byref(rng), cnt.ctypes.data_as(po.i64p), pc.ctypes.data_as(po.i64p))
        out.append((n, cnt, pc, rng.as_array()))
    assert out[0][0] == out[1][0]
    for a, b in zip(out[0][1:], out[1][1:]):
        np.testing.assert_array_equal(a, b)
    assert out[0][2].sum() > 0
