"""The oracle's reference potential beyond n-6 LJ + point charges, pinned by what the
mathematics fixes (generate_rp, src/epsr26/programs/epsr_standalone_rev3.py:5675-5872):

  chsph     the Coulomb energy of two uniformly charged spheres (ato_manipulation_rev4.py:
            589-611) against a direct numerical integration of the overlap (all three
            branches: apart, one inside the other, partial overlap)
  pcoefex   the exponential repulsive core pcoefex exp((rminex - r) / (expecf sqrt(rminex)))
            and, with a gradient limit G, its linear continuation pcoefex G (1 + d - ln G)
            past d = ln G, added to the Coulomb table times wfac
  refpotfac a factor on the LJ and Coulomb terms
"""

import numpy as np
import pytest
from scipy import integrate

from oracle import pyoracle as po

F32 = np.float32


def _phi(ra, s):
    """potential of a unit charge spread uniformly over a sphere of radius ra, at distance s"""
    return 1.0 / s if s >= ra else (3.0 * ra * ra - s * s) / (2.0 * ra ** 3)


def _sphere_energy(ra, rb, r):
    """E = int over sphere b (radius rb, centre at distance r) of phi_a, by shells: a shell of
    radius u averages a radial f over s in [|r - u|, r + u] with weight s / (2 u r)."""
    def shell(u):
        lo, hi = abs(r - u), r + u
        pts = [ra] if lo < ra < hi else None
        v, _ = integrate.quad(lambda s: _phi(ra, s) * s, lo, hi, points=pts, epsabs=1e-13, epsrel=1e-11)
        return v / (2.0 * u * r)
    v, _ = integrate.quad(lambda u: 3.0 * u * u / rb ** 3 * shell(u), 0.0, rb, epsabs=1e-13, epsrel=1e-11,
                          limit=200)
    return v


@pytest.mark.parametrize("ra,rb", [(1.0, 0.5), (0.8, 0.8), (1.2, 0.3)])
def test_charged_sphere_energy_against_integration(ra, rb):
    L = po.lib("cr")
    rs = np.concatenate([np.linspace(0.05, abs(ra - rb) * 0.98, 4) if ra != rb else [],
                         np.linspace(abs(ra - rb) + 0.02, ra + rb - 0.02, 6), [ra + rb + 0.1, 3.0]])
    for r in rs:
        got = float(L.ora_chsph(F32(ra), F32(rb), F32(r)))
        want = _sphere_energy(float(F32(ra)), float(F32(rb)), float(F32(r)))
        assert abs(got - want) <= 2e-5 * abs(want), (ra, rb, r, got, want)
        # symmetric in the two radii
        assert got == float(L.ora_chsph(F32(rb), F32(ra), F32(r)))


def _rp(L, x, xstar, wfac, wfacc, ecor, dcor, qcor, qr, pcoefex=None, rminex=None, refpotfac=1.0, expecf=0.3,
        expecfpwr=0.0):
    npt = len(x) - 1
    npairs = len(ecor)
    ld = npt + 1
    out = {k: np.zeros(npairs * ld, F32) for k in ("epot", "ecoul", "epot1", "ecoul1")}
    a = lambda v: np.ascontiguousarray(v, F32)
    L.ora_generate_rp(npt, npt, po.Pf(a(x)), po.Pf(a(xstar)), po.Pf(a(wfac)), po.Pf(a(wfacc)), npairs,
                      po.Pf(a(ecor)), po.Pf(a(dcor)), po.Pf(a(qcor)), po.Pf(a(qr)), po.Pf(a(qr)), F32(12.0),
                      F32(refpotfac), None if pcoefex is None else po.Pf(a(pcoefex)),
                      None if rminex is None else po.Pf(a(rminex)), F32(expecf), F32(expecfpwr),
                      po.Pf(out["epot"]), po.Pf(out["ecoul"]), po.Pf(out["epot1"]), po.Pf(out["ecoul1"]), ld)
    return {k: v.reshape(npairs, ld) for k, v in out.items()}


def _grid():
    npt = 400
    x = np.zeros(npt + 1, F32)
    x[1:] = (np.arange(npt, dtype=np.float64) * 0.03 + 0.015).astype(F32)
    xstar = x.copy()
    L = po.lib("cr")
    wfac = np.array([0.0] + [L.ora_truncate(F32(v), F32(9.0), F32(11.0)) for v in x[1:]], F32)
    wfacc = np.array([0.0] + [L.ora_truncatec(F32(v), F32(11.0)) for v in x[1:]], F32)
    return x, xstar, wfac, wfacc


@pytest.mark.parametrize("gradlimit", [0.0, 40.0])
def test_repulsive_core_closed_form(gradlimit):
    L = po.lib("cr")
    x, xstar, wfac, wfacc = _grid()
    ecor, dcor, qcor = np.array([0.65, 0.3], F32), np.array([3.1, 2.0], F32), np.array([-300.0, 120.0], F32)
    qr = np.zeros(2, F32)
    pcoefex, rminex, expecf = np.array([1.7, 0.0], F32), np.array([2.6, 2.0], F32), 0.35
    base = _rp(L, x, xstar, wfac, wfacc, ecor, dcor, qcor, qr, expecf=expecf, expecfpwr=gradlimit)
    core = _rp(L, x, xstar, wfac, wfacc, ecor, dcor, qcor, qr, pcoefex, rminex, expecf=expecf, expecfpwr=gradlimit)
    # the core goes to the Coulomb table only, and only for the pair with pcoefex != 0
    assert np.array_equal(core["epot"], base["epot"])
    assert np.array_equal(core["ecoul"][1], base["ecoul"][1])
    xd = x[2:].astype(np.float64)
    d = (2.6 - xd) / (expecf * np.sqrt(2.6))
    add = 1.7 * np.exp(d)
    if gradlimit > 0.0:
        lim = d >= np.log(gradlimit)
        assert lim.any() and (~lim).any()
        add = np.where(lim, 1.7 * gradlimit * (1.0 + d - np.log(gradlimit)), add)
    got = core["ecoul"][0, 2:].astype(np.float64) - base["ecoul"][0, 2:]
    want = add * wfac[2:]
    scale = np.abs(core["ecoul"][0, 2:]).astype(np.float64) + np.abs(want)
    assert np.all(np.abs(got - want) <= 4e-7 * scale + 1e-30)
    # the gradients of the Coulomb table follow the table on the r* grid
    g = (core["ecoul"][0, 2:].astype(np.float64) - core["ecoul"][0, 1:-1]) / (xstar[2:] - xstar[1:-1])
    assert np.allclose(core["ecoul1"][0, 1:-1], g, rtol=1e-5, atol=1e-6 * np.abs(g).max())


def test_refpotfac_scales_lj_and_coulomb():
    L = po.lib("cr")
    x, xstar, wfac, wfacc = _grid()
    ecor, dcor, qcor = np.array([0.65], F32), np.array([3.1], F32), np.array([-300.0], F32)
    qr = np.zeros(1, F32)
    one = _rp(L, x, xstar, wfac, wfacc, ecor, dcor, qcor, qr, refpotfac=1.0)
    half = _rp(L, x, xstar, wfac, wfacc, ecor, dcor, qcor, qr, refpotfac=0.5)
    # multiplying by 0.5 is exact in float32: each term halves, and so does their sum
    assert np.array_equal(half["ecoul"], (one["ecoul"] * F32(0.5)).astype(F32))
    assert np.array_equal(half["epot"], (one["epot"] * F32(0.5)).astype(F32))
