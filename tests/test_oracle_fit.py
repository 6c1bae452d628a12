"""The oracle's EP-fit functions pinned by what the mathematics fixes:
smoop is a [1/4, 1/2, 1/4] convolution with zero ends; expand clamps after every term
(not once at the end); fit_expansion at an exact fit rejects every trial, runs all 150
tries and consumes exactly 150 x 100 x nd ran1 values; on a noisy problem it never beats
the FP64 least-squares optimum and never ends above its start; the ptgnsq difference is
data minus the weighted partials inside [nqmin, nqmax], the linear extrapolation to
szeros below qmin, zero elsewhere."""

import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as po

F32 = np.float32
SUMMAX = F32(np.exp(25.0))


def _lib(flavour="cr"):
    return po.lib(flavour)


def test_smoop_is_a_three_point_convolution():
    rng = np.random.default_rng(3)
    n = 40
    d = rng.normal(size=n + 2).astype(F32)
    for n1, n2, ns in ((1, n, 1), (5, 30, 1), (2, 20, 3)):
        got = d.copy()
        _lib().ora_smoop(n1, n2, ns, po.Pf(got))
        want = d.astype(np.float64).copy()
        for _ in range(ns):
            seg = np.concatenate([[0.0], want[n1:n2 + 1], [0.0]])
            want[n1:n2 + 1] = 0.25 * seg[:-2] + 0.5 * seg[1:-1] + 0.25 * seg[2:]
        np.testing.assert_allclose(got[n1:n2 + 1], want[n1:n2 + 1], rtol=0, atol=4e-7 * ns * np.abs(d).max())
        assert np.array_equal(got[:n1], d[:n1]) and np.array_equal(got[n2 + 1:], d[n2 + 1:])
    # fewer than three points: untouched
    got = d.copy()
    _lib().ora_smoop(4, 5, 1, po.Pf(got))
    assert np.array_equal(got, d)


def test_expand_dot_product_and_clamp_after_every_term():
    rng = np.random.default_rng(4)
    nterms, npt = 7, 5
    expr = np.zeros(npt * nterms + 1, F32)
    expr[1:] = rng.normal(size=npt * nterms).astype(F32)
    coeff = np.zeros(nterms + 1, F32)
    coeff[1:] = rng.normal(size=nterms).astype(F32)
    for j in range(1, npt + 1):
        row = expr[(j - 1) * nterms + 1:j * nterms + 1].astype(np.float64)
        want = float(row @ coeff[1:].astype(np.float64))
        got = _lib().ora_expand(nterms, j, po.Pf(expr), po.Pf(coeff), SUMMAX)
        assert abs(got - want) <= 1e-6 * np.sum(np.abs(row * coeff[1:]))
    # a partial sum beyond summax is clamped before the next term is added:
    # (2S) -> S, then S - S = 0, where a single clamp at the end would give S
    e = np.array([0.0, 2.0, -1.0], F32)
    c = np.array([0.0, SUMMAX, SUMMAX], F32)
    assert _lib().ora_expand(2, 1, po.Pf(e), po.Pf(c), SUMMAX) == 0.0
    # clamped to -S after the first term, and again after the second (-S - 0.5)
    c2 = np.array([0.0, -SUMMAX, 0.5], F32)
    assert _lib().ora_expand(2, 1, po.Pf(e), po.Pf(c2), SUMMAX) == -SUMMAX


def _problem(seed, npt, nterms, noise):
    rng = np.random.default_rng(seed)
    x = np.linspace(0.05, 1.0, npt)
    basis = np.stack([np.exp(-((x - c) / 0.15) ** 2) for c in np.linspace(0.0, 1.0, nterms)], axis=1)
    expr = np.zeros(npt * nterms + 1, F32)
    expr[1:] = basis.astype(F32).reshape(-1)
    wtrue = rng.normal(size=nterms)
    y = basis @ wtrue + noise * rng.normal(size=npt)
    yin = np.zeros(npt + 1, F32)
    yin[1:] = y.astype(F32)
    return expr, yin, wtrue


def _fit(flavour, yin, expr, wtold, nmin, npt, ncfirst, nterms, nits, nfreq, stepr, seed):
    nt = nterms + 1
    wt = np.zeros(nt, F32)
    wo = wtold.astype(F32).copy()
    yout = np.zeros(npt + 1, F32)
    st, ac, rj, cs = (np.array([v], F32) for v in (stepr, 0.0, 0.0, 0.0))
    rng = po.Ran1.seeded(seed)
    ntries = C.c_int()
    _lib(flavour).ora_fit_expansion(po.Pf(yin), po.Pf(expr), po.Pf(yout), po.Pf(wt), po.Pf(wo), po.Pf(st),
                                    po.Pf(ac), po.Pf(rj), po.Pf(cs), nmin, npt, ncfirst, nterms, nits, nfreq,
                                    C.byref(rng), C.byref(ntries))
    return dict(wt=wt, wtold=wo, yout=yout, stepr=st[0], acceptr=ac[0], rejrat=rj[0], chisq=cs[0], rng=rng,
                ntries=ntries.value)


@pytest.mark.parametrize("flavour", ["cr", "glibc"])
def test_fit_expansion_at_an_exact_fit_rejects_every_trial(flavour):
    # yin = 0 and w = 0: chi^2 = 0, so accept = 0 and every trial (additive, |w| < 1, so
    # never equal to w) has exponent = +inf.  (A trial equal to w would give 0/0 and be
    # accepted, as in the reference; additive steps on zero coefficients never are.)
    npt, nterms, ncfirst = 30, 6, 2
    expr, _, _ = _problem(5, npt, nterms, 0.0)
    w = np.zeros(nterms + 1, F32)
    yin = np.zeros(npt + 1, F32)
    r = _fit(flavour, yin, expr, w, 1, npt, ncfirst, nterms, 1, 10, 0.01, -77)
    assert np.array_equal(r["wt"], w) and np.array_equal(r["wtold"], w)
    assert r["chisq"] == 0.0 and r["rejrat"] == 1.0 and r["ntries"] == 150
    assert np.array_equal(r["yout"], yin)
    # every trial draws nd values and no test value (the exponent is +inf)
    nd = nterms - ncfirst + 1
    ref = po.Ran1.seeded(-77)
    for _ in range(150 * 100 * nd):
        _lib(flavour).ora_ran1(C.byref(ref))
    assert np.array_equal(r["rng"].as_array(), ref.as_array())
    # the step shrinks by 0.75 per try (rejrat = 1), from max(stepr, 2e-6 max|w|)
    s = F32(0.01)
    for _ in range(150):
        s = F32(s * F32(0.75))
    assert r["stepr"] == s


def test_fit_expansion_descends_towards_the_least_squares_optimum():
    npt, nterms = 60, 6
    expr, yin, _ = _problem(7, npt, nterms, 0.05)
    A = expr[1:].astype(np.float64).reshape(npt, nterms)
    wls, *_ = np.linalg.lstsq(A, yin[1:].astype(np.float64), rcond=None)
    chi_ls = float(np.sum((yin[1:] - A @ wls) ** 2))
    w0 = np.zeros(nterms + 1, F32)
    w0[1:] = (0.5 * wls).astype(F32)
    chi0 = float(np.sum((yin[1:] - A @ w0[1:].astype(np.float64)) ** 2))
    r = _fit("cr", yin, expr, w0, 1, npt, 1, nterms, 1, 10, 0.05, -123)
    assert chi_ls * (1 - 1e-4) <= r["chisq"] <= chi0
    assert r["chisq"] < 0.5 * chi0
    assert 10 <= r["ntries"] <= 150


def test_ptgnsq_difference_regions():
    nfile, nq, npairs = 2, 50, 3
    ldq = nq + 1
    rng = np.random.default_rng(8)
    q = np.zeros(nq + 1, F32)
    q[1:] = (0.1 * np.arange(nq)).astype(F32)
    qwind = np.zeros(npairs * ldq, F32)
    qwind.reshape(npairs, ldq)[:, 1:] = rng.normal(size=(npairs, nq)).astype(F32)
    cwt = rng.uniform(0.1, 1.0, size=nfile * npairs).astype(F32)
    data = np.zeros(nfile * ldq, F32)
    data.reshape(nfile, ldq)[:, 1:] = rng.normal(size=(nfile, nq)).astype(F32)
    szeros = np.array([0.0, 0.3], F32)
    nqmin = np.array([5, 8], np.int32)
    nqmax = np.array([40, 45], np.int32)
    qmin = F32(0.35)
    out = np.full(nfile * ldq, F32(np.nan))
    _lib().ora_ptgnsq_diff(nfile, nq, npairs, po.Pf(q), po.Pf(data), ldq, po.Pf(qwind), po.Pf(cwt), po.Pf(szeros),
                           po.Pi(nqmin), po.Pi(nqmax), qmin, po.Pf(out))
    for id in range(nfile):
        qw = qwind.reshape(npairs, ldq)[:, 1:].astype(np.float64)
        s = cwt[id * npairs:(id + 1) * npairs].astype(np.float64) @ qw
        dat = data.reshape(nfile, ldq)[id, 1:].astype(np.float64)
        got = out.reshape(nfile, ldq)[id, 1:]
        qlim = max(float(qmin), float(q[nqmin[id]]))
        i = np.arange(1, nq + 1)
        low = q[1:] < qlim
        mid = ~low & (i >= nqmin[id]) & (i <= nqmax[id])
        high = ~low & ~mid
        np.testing.assert_allclose(got[mid], (dat - s)[mid], atol=1e-5)
        assert np.all(got[high] == 0.0)
        if szeros[id] == 0.0:
            assert np.all(got[low] == 0.0)
        else:
            # a straight line through (q1, szeros) and (q[nqmin], data[nqmin]), minus the partials
            g = (dat[nqmin[id] - 1] - szeros[id]) / (float(q[nqmin[id]]) - float(q[1]))
            line = szeros[id] + g * (q[1:].astype(np.float64) - float(q[1]))
            np.testing.assert_allclose(got[low], (line - s)[low], atol=1e-5)
    # rfactor: sum |fit + qt| / |qt| (or |fit + qt| where qt = 0) over the range, averaged
    fit = rng.normal(size=nq + 1).astype(F32)
    qt = out[:ldq].copy()
    r = _lib().ora_ptgnsq_rfactor(5, 40, po.Pf(fit), po.Pf(qt))
    num = np.abs(fit[5:41].astype(np.float64) + qt[5:41])
    den = np.where(np.abs(qt[5:41]) > 0, np.abs(qt[5:41]).astype(np.float64), 1.0)
    assert abs(r - float(np.mean(num / den))) <= 1e-5 * float(np.mean(num / den))
