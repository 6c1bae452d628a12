"""The EP fit on the GPU against the oracle: expand and smoop bit-identical; the
Metropolis fit_expansion replayed bit for bit against the correctly rounded oracle
(coefficients, fitted curve, step, acceptance, chi^2, tries and the ran1 state after it),
for one chain and for several at once; the FP64 normal-equation mode against numpy's
least squares; the ptgnsq difference, r-factor and coefficient update."""

import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as po

F32 = np.float32
SUMMAX = F32(np.exp(25.0))
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from epsr_python_translation_amd import device as D
    return D.Context(0)


def _gauss_problem(seed, npt, nterms, noise, scale=1.0):
    rng = np.random.default_rng(seed)
    x = np.linspace(0.05, 1.0, npt)
    basis = np.stack([np.exp(-((x - c) / 0.15) ** 2) for c in np.linspace(0.0, 1.0, nterms)], axis=1)
    expr = np.zeros(npt * nterms + 1, F32)
    expr[1:] = basis.astype(F32).reshape(-1)
    w = rng.normal(size=nterms) * scale
    yin = np.zeros(npt + 1, F32)
    yin[1:] = (basis @ w + noise * rng.normal(size=npt)).astype(F32)
    return expr, yin


def _poisson_problem(seed, nq, nterms):
    """The Q-space Poisson basis of the oracle and a smooth difference to fit."""
    rng = np.random.default_rng(seed)
    q = np.zeros(nq + 1, F32)
    q[1:] = (0.05 * np.arange(nq)).astype(F32)
    expr = np.zeros(nq * nterms + 1, F32)
    pdstep = F32(F32(12.0) / F32(nterms))
    po.lib("cr").ora_calc_pwrexpq(nq, nterms, F32(0.0), pdstep, F32(12.0), F32(0.3), F32(0.1), po.Pf(q), po.Pf(expr),
                                  F32(0.0))
    w = rng.normal(size=nterms) * 0.05
    A = expr[1:].astype(np.float64).reshape(nq, nterms)
    yin = np.zeros(nq + 1, F32)
    yin[1:] = (A @ w + 0.002 * rng.normal(size=nq)).astype(F32)
    return expr, yin, q


def _oracle_fit(yin, expr, wtold, nmin, npt, ncfirst, nterms, nits, nfreq, stepr, idum):
    wt = np.zeros(nterms + 1, F32)
    wo = np.asarray(wtold, F32).copy()
    yout = np.zeros(npt + 1, F32)
    st, ac, rj, cs = (np.array([v], F32) for v in (stepr, 0.0, 0.0, 0.0))
    rng = po.Ran1.seeded(idum)
    nt = C.c_int()
    po.lib("cr").ora_fit_expansion(po.Pf(yin), po.Pf(expr), po.Pf(yout), po.Pf(wt), po.Pf(wo), po.Pf(st), po.Pf(ac),
                                   po.Pf(rj), po.Pf(cs), nmin, npt, ncfirst, nterms, nits, nfreq, C.byref(rng),
                                   C.byref(nt))
    return dict(wt=wt, wtold=wo, yout=yout, stepr=st[0], acceptr=ac[0], rejrat=rj[0], chisq=cs[0],
                ran1=rng.as_array(), ntries=nt.value)


def _bits(a):
    return np.asarray(a, F32).view(np.int32)


def test_expand_and_smoop_bit_identical(ctx):
    from epsr_python_translation_amd import device as D
    print("expand/smoop", flush=True)
    expr, yin, q = _poisson_problem(1, 300, 20)
    coeff = np.zeros(21, F32)
    coeff[1:] = np.random.default_rng(2).normal(size=20).astype(F32)
    coeff[3] = F32(1e12)   # drives partial sums past summax: the clamp path
    got = D.expand_all(ctx, 20, 300, ctx.tensor(expr), ctx.tensor(coeff)).cpu().numpy()
    L = po.lib("cr")
    want = np.array([0.0] + [L.ora_expand(20, j, po.Pf(expr), po.Pf(coeff), SUMMAX) for j in range(1, 301)], F32)
    assert np.array_equal(_bits(got[1:]), _bits(want[1:]))
    d = np.random.default_rng(3).normal(size=64).astype(F32)
    for n1, n2, ns in ((1, 60, 1), (4, 40, 3)):
        t = ctx.tensor(d)
        D.smoop(ctx, n1, n2, ns, t)
        ref = d.copy()
        L.ora_smoop(n1, n2, ns, po.Pf(ref))
        assert np.array_equal(_bits(t.cpu().numpy()), _bits(ref))


CASES = [
    # name, problem, nmin, ncfirst, nfreq, stepr, idum
    ("gauss_noisy", ("gauss", 11, 120, 8, 0.05, 1.0), 1, 1, 10, 0.05, -5),
    ("gauss_ncfirst3_nmin5", ("gauss", 12, 90, 10, 0.02, 3.0), 5, 3, 4, 0.2, -19),
    ("poisson_c0_like", ("poisson", 13, 300, 20), 1, 1, 10, 0.01, -101),
    ("exact_fit_nan_path", ("exact", 14, 40, 6), 1, 2, 10, 0.01, -77),
    # the EPSR test case's size: 116 Poisson terms (FIT_NTMAX is 128) over 600 Q points
    ("poisson_epsr_test_116", ("poisson", 15, 600, 116), 1, 1, 10, 0.01, -131),
]


def _make(problem):
    kind = problem[0]
    if kind == "gauss":
        _, seed, npt, nterms, noise, scale = problem
        expr, yin = _gauss_problem(seed, npt, nterms, noise, scale)
        w0 = np.zeros(nterms + 1, F32)
        w0[1:] = np.random.default_rng(seed + 100).normal(size=nterms).astype(F32) * F32(scale)
        return expr, yin, w0, npt, nterms
    if kind == "poisson":
        _, seed, nq, nterms = problem
        expr, yin, _ = _poisson_problem(seed, nq, nterms)
        w0 = np.zeros(nterms + 1, F32)
        return expr, yin, w0, nq, nterms
    _, seed, npt, nterms = problem
    expr, _ = _gauss_problem(seed, npt, nterms, 0.0)
    w0 = np.zeros(nterms + 1, F32)
    w0[1:] = np.random.default_rng(seed).normal(size=nterms).astype(F32)
    L = po.lib("cr")
    yin = np.array([0.0] + [L.ora_expand(nterms, j, po.Pf(expr), po.Pf(w0), SUMMAX) for j in range(1, npt + 1)], F32)
    return expr, yin, w0, npt, nterms


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_fit_expansion_replay_bit_exact(ctx, case):
    from epsr_python_translation_amd import device as D
    name, problem, nmin, ncfirst, nfreq, stepr, idum = case
    print(f"fit_expansion {name}", flush=True)
    expr, yin, w0, npt, nterms = _make(problem)
    ref = _oracle_fit(yin, expr, w0, nmin, npt, ncfirst, nterms, 1, nfreq, stepr, idum)
    st = np.zeros(35, np.int32)
    st[1] = idum
    wt = ctx.zeros(nterms + 1)
    wo = ctx.tensor(w0)
    yout = ctx.zeros(npt + 1)
    r = D.fit_expansion(ctx, ctx.tensor(yin), ctx.tensor(expr), yout, wt, wo, stepr, 0.0, 0.0, 0.0, nmin, npt, ncfirst,
                        nterms, 1, nfreq, st)
    print(f"  tries {r.ntries[0]} (oracle {ref['ntries']}), chi2 {r.chisqnew[0]:.6g}, rejrat {r.rejrat[0]}", flush=True)
    assert r.ntries[0] == ref["ntries"]
    assert np.array_equal(_bits(wt.cpu().numpy()), _bits(ref["wt"]))
    assert np.array_equal(_bits(wo.cpu().numpy()), _bits(ref["wtold"]))
    assert np.array_equal(_bits(yout.cpu().numpy()[1:]), _bits(ref["yout"][1:]))
    for k, v in (("stepr", r.stepr), ("acceptr", r.acceptr), ("rejrat", r.rejrat), ("chisq", r.chisqnew)):
        assert _bits(v[0]) == _bits(ref[k]), k
    assert np.array_equal(st, ref["ran1"])


def test_fit_expansion_several_chains_at_once(ctx):
    from epsr_python_translation_amd import device as D
    print("fit_expansion 3 chains", flush=True)
    npt, nterms = 100, 8
    refs, ys, es, ws = [], [], [], []
    for c in range(3):
        expr, yin = _gauss_problem(30 + c, npt, nterms, 0.03)
        w0 = np.zeros(nterms + 1, F32)
        refs.append(_oracle_fit(yin, expr, w0, 1, npt, 1, nterms, 1, 10, 0.05, -(40 + c)))
        ys.append(yin)
        es.append(expr)
        ws.append(w0)
    st = np.zeros(3 * 35, np.int32)
    for c in range(3):
        st[35 * c + 1] = -(40 + c)
    lde = npt * nterms + 1
    wt = ctx.zeros(3 * (nterms + 1))
    wo = ctx.tensor(np.concatenate(ws))
    yout = ctx.zeros(3 * (npt + 1))
    r = D.fit_expansion(ctx, ctx.tensor(np.concatenate(ys)), ctx.tensor(np.concatenate(es)), yout, wt, wo, 0.05, 0.0,
                        0.0, 0.0, 1, npt, 1, nterms, 1, 10, st, nchain=3, lde=lde)
    wt_h, yo_h = wt.cpu().numpy().reshape(3, -1), yout.cpu().numpy().reshape(3, -1)
    for c in range(3):
        assert np.array_equal(_bits(wt_h[c]), _bits(refs[c]["wt"]))
        assert np.array_equal(_bits(yo_h[c, 1:]), _bits(refs[c]["yout"][1:]))
        assert _bits(r.chisqnew[c]) == _bits(refs[c]["chisq"]) and r.ntries[c] == refs[c]["ntries"]
        assert np.array_equal(st[35 * c:35 * (c + 1)], refs[c]["ran1"])


def _ne_fit(ctx, yin, expr, w0, npt, ncfirst, nterms):
    from epsr_python_translation_amd import _native
    from epsr_python_translation_amd import device as D
    wt, wo, yout = ctx.zeros(nterms + 1), ctx.tensor(w0), ctx.zeros(npt + 1)
    r = D.fit_expansion(ctx, ctx.tensor(yin), ctx.tensor(expr), yout, wt, wo, 0.01, 0.0, 0.0, 0.0, 1, npt, ncfirst,
                        nterms, 1, 10, None, mode=_native.FIT_NORMAL)
    return wt.cpu().numpy(), r


def test_fit_expansion_normal_equations_fp64(ctx):
    from epsr_python_translation_amd import device as D
    print("fit_expansion normal equations", flush=True)
    # a well-conditioned basis: the coefficients themselves match numpy's least squares
    npt, nterms, ncfirst = 200, 8, 2
    expr, yin = _gauss_problem(41, npt, nterms, 0.01)
    w0 = np.zeros(nterms + 1, F32)
    w0[1] = F32(0.3)
    got, r = _ne_fit(ctx, yin, expr, w0, npt, ncfirst, nterms)
    A = expr[1:].astype(np.float64).reshape(npt, nterms)
    y = yin[1:].astype(np.float64) - A[:, 0] * float(w0[1])
    wls, *_ = np.linalg.lstsq(A[:, 1:], y, rcond=None)
    np.testing.assert_allclose(got[ncfirst:].astype(np.float64), wls, rtol=1e-5, atol=1e-6 * np.abs(wls).max())
    assert got[1] == w0[1]
    # the Poisson basis is ill-conditioned (alternating coefficients): the solve reaches
    # the least-squares chi^2 and no Metropolis fit from the same start does better
    expr, yin, _ = _poisson_problem(21, 400, 20)
    npt, nterms, ncfirst = 400, 20, 1
    w0 = np.zeros(nterms + 1, F32)
    got, r = _ne_fit(ctx, yin, expr, w0, npt, ncfirst, nterms)
    A = expr[1:].astype(np.float64).reshape(npt, nterms)
    wls, *_ = np.linalg.lstsq(A, yin[1:].astype(np.float64), rcond=None)
    chi_ls = float(np.sum((yin[1:] - A @ wls) ** 2))
    print(f"  chi2 normal equations {float(r.chisqnew[0]):.6g}, numpy lstsq {chi_ls:.6g}", flush=True)
    assert float(r.chisqnew[0]) <= chi_ls * 1.01 + 1e-9
    st = np.zeros(35, np.int32)
    st[1] = -3
    wt2, wo2 = ctx.zeros(nterms + 1), ctx.tensor(w0)
    r2 = D.fit_expansion(ctx, ctx.tensor(yin), ctx.tensor(expr), ctx.zeros(npt + 1), wt2, wo2, 0.01, 0.0, 0.0, 0.0, 1,
                         npt, ncfirst, nterms, 1, 10, st)
    assert float(r.chisqnew[0]) <= float(r2.chisqnew[0]) * (1 + 1e-5)


def test_ptgnsq_pieces(ctx):
    from epsr_python_translation_amd import device as D
    print("ptgnsq diff/rfactor/update", flush=True)
    rng = np.random.default_rng(9)
    nfile, nq, npairs = 3, 200, 3
    ldq = nq + 1
    q = np.zeros(ldq, F32)
    q[1:] = (0.05 * np.arange(nq)).astype(F32)
    qwind = np.zeros(npairs * ldq, F32)
    qwind.reshape(npairs, ldq)[:, 1:] = rng.normal(size=(npairs, nq)).astype(F32)
    data = np.zeros(nfile * ldq, F32)
    data.reshape(nfile, ldq)[:, 1:] = rng.normal(size=(nfile, nq)).astype(F32)
    cwt = rng.uniform(0.1, 1.0, size=nfile * npairs).astype(F32)
    szeros = np.array([0.0, 0.4, -0.2], F32)
    nqmin = np.array([3, 10, 6], np.int32)
    nqmax = np.array([180, 200, 150], np.int32)
    qmin = F32(0.3)
    got = D.ptgnsq_diff(ctx, nfile, nq, npairs, ctx.tensor(q), ctx.tensor(data), ldq, ctx.tensor(qwind), cwt, szeros,
                        nqmin, nqmax, qmin)
    want = np.zeros(nfile * ldq, F32)
    po.lib("cr").ora_ptgnsq_diff(nfile, nq, npairs, po.Pf(q), po.Pf(data), ldq, po.Pf(qwind), po.Pf(cwt),
                                 po.Pf(szeros), po.Pi(nqmin), po.Pi(nqmax), qmin, po.Pf(want))
    g = got.cpu().numpy().reshape(nfile, ldq)[:, 1:]
    assert np.array_equal(_bits(g), _bits(want.reshape(nfile, ldq)[:, 1:]))
    fit = rng.normal(size=ldq).astype(F32)
    rf = D.ptgnsq_rfactor(ctx, 3, 180, ctx.tensor(fit), got[:ldq])
    assert _bits(rf) == _bits(po.lib("cr").ora_ptgnsq_rfactor(3, 180, po.Pf(fit), po.Pf(want[:ldq])))
    # coefficient update: pcoefp(ic, irefp) += ((pfact * efacp) * cwtpot) * pwtold(ic)
    nterms, ldp = 10, 11
    pw = rng.normal(size=nterms + 1).astype(F32)
    irefp = np.array([0, 1, 1], np.int32)
    efacp = np.array([1.0, 0.5], F32)
    cwtpot = rng.uniform(size=npairs * nfile).astype(F32)
    pcoefd = np.zeros(nfile * ldp, F32)
    pcoefp = rng.normal(size=2 * ldp).astype(F32)
    before = pcoefp.copy()
    D.ptgnsq_update(ctx, 1, nterms, npairs, pw, irefp, efacp, cwtpot, nfile, F32(0.7), pcoefd, pcoefp, ldp)
    assert np.array_equal(pcoefd[ldp + 2:ldp + nterms + 1], pw[2:])
    exp = before.copy()
    for ic in range(2, nterms + 1):
        for ip in range(npairs):
            r_ = irefp[ip]
            exp[r_ * ldp + ic] = F32(exp[r_ * ldp + ic] + F32(F32(F32(F32(0.7) * efacp[r_]) * cwtpot[ip * nfile + 1])
                                                              * pw[ic]))
    assert np.array_equal(_bits(pcoefp), _bits(exp))
