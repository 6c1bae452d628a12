"""Checks that fix oracle functions no reference golden covers, by what the mathematics
decides (closed forms, identities, limits), chosen so that a dropped term, a wrong sign
or a shifted index fails:

* gensq: a histogram with one non-zero bin gives that bin's foqols column, exactly;
  an empty intermolecular histogram gives zero; gensq is linear.
* stog: one non-zero S(Q) point gives the single term of the transform (FP64 numpy),
  the r = 0 branch, and the Lorch window's zero at its edge.
* j1qr is the textbook integral (sin qr - qr cos qr)/q^3 (r^3/3 at q = 0); foqols'
  bin differences telescope to j1qr at the outer edge; the resolution factor is a
  Gaussian in q r.
* the Poisson basis: the Q-space functions (calc_pwrexpq) are the 3-D Fourier transforms
  of the r-space ones, which calc_pwrexpr_grad gives as derivatives: integrated back and
  transformed by quadrature, they agree.
* molfit: the ledger equals freshly computed pair energies after sweeps; the energy
  change booked in fadds equals the change of the ledger's total; at a vanishing
  temperature no move raises the energy; at a huge temperature every move is accepted."""

import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as po
from tests import helpers as H

F32 = np.float32


def _gensq(L, nq, npairs, npt1, wav, ycoh, fo, ldf):
    ldw = npt1 + 1
    ldq = nq + 1
    sofq = np.zeros(2 * npairs * ldq, F32)
    n2 = np.zeros(npairs, np.int32)
    L.ora_gensq(nq, npairs, npt1, ldw, po.Pf(wav), po.Pf(ycoh), po.Pf(fo), ldf, po.Pi(n2), None, po.Pf(sofq), ldq)
    return sofq.reshape(2, npairs, ldq)


def test_gensq_single_bin_closed_form_and_linearity():
    L = po.lib("cr")
    rng = np.random.default_rng(1)
    nq, npairs, npt1 = 40, 3, 60
    ldf = npt1 + 1
    fo = rng.normal(0, 1, (nq + 1) * ldf).astype(F32)
    ycoh = np.abs(rng.normal(1, 0.2, npt1 + 1)).astype(F32)
    ldw = npt1 + 1
    j, w = 17, F32(3.0)
    wav = np.zeros(2 * npairs * ldw, F32)
    wav.reshape(2, npairs, ldw)[0, 1, j] = w                  # intra, pair 1: one bin
    inter = wav.reshape(2, npairs, ldw)[1]
    inter[:, 2:] = ycoh[2:]                                  # inter = uniform gas ...
    inter[2, j] = ycoh[j] + w                                # ... plus w in one bin of pair 2
    s = _gensq(L, nq, npairs, npt1, wav, ycoh, fo, ldf)
    col = fo.reshape(nq + 1, ldf)[1:, j]
    assert np.array_equal(s[0, 1, 1:], (w * col).astype(F32))
    assert not s[0, 0, 1:].any() and not s[0, 2, 1:].any()
    assert np.array_equal(s[1, 2, 1:], ((inter[2, j] - ycoh[j]) * col).astype(F32))
    assert not s[1, 0, 1:].any() and not s[1, 1, 1:].any()   # wav2 = ycoh: nothing
    # an empty intermolecular histogram is zeroed (sum <= 0)
    wav0 = np.zeros_like(wav)
    assert not _gensq(L, nq, npairs, npt1, wav0, ycoh, fo, ldf)[1].any()
    # linearity of the intramolecular part against FP64
    a = np.zeros_like(wav)
    a.reshape(2, npairs, ldw)[0, :, 2:] = rng.integers(0, 5, (npairs, npt1 - 1)).astype(F32)
    s = _gensq(L, nq, npairs, npt1, a, ycoh, fo, ldf)[0, :, 1:]
    want = a.reshape(2, npairs, ldw)[0, :, 2:].astype(np.float64) @ fo.reshape(nq + 1, ldf)[1:, 2:npt1 + 1].T
    np.testing.assert_allclose(s, want, rtol=1e-5, atol=1e-5 * np.abs(want).max())


def test_stog_single_point_closed_form():
    L = po.lib("cr")
    nq, nr, ncols = 50, 40, 2
    q = np.zeros(nq + 1, F32)
    q[1:] = (0.1 * np.arange(nq)).astype(F32)
    r = np.zeros(nr + 1, F32)
    r[1:] = (0.25 * np.arange(nr)).astype(F32)   # r[1] = 0
    ldq, ldg = nq + 1, nr + 1
    rho = F32(0.1)
    i0 = 20
    sofq = np.zeros(ncols * ldq, F32)
    sofq[i0] = F32(2.0)                  # column 0: one point
    sofq[ldq + nq] = F32(5.0)            # column 1: only at nqmax, where lorch = sin(pi)/pi -> 0 branch
    nqmin = np.array([2, 2], np.int32)
    nqmax = np.array([nq, nq], np.int32)
    xwt = np.array([0.5, -1.0], F32)
    gofr = np.zeros(ncols * ldg, F32)
    L.ora_stog(rho, po.Pf(q), po.Pf(sofq), ldq, po.Pf(r), po.Pf(gofr), ldg, po.Pi(nqmin), po.Pi(nqmax), nr, ncols,
               po.Pf(xwt))
    g = gofr.reshape(ncols, ldg)
    qstep = float(q[4]) - float(q[3])
    sfac = 0.5 * qstep / np.pi ** 2 / float(rho)
    qq = float(q[i0])
    x = np.pi * qq / float(q[nq])
    lorch = np.sin(x) / x
    rr = r[2:].astype(np.float64)
    want = 0.5 + sfac * 2.0 * qq * np.sin(qq * rr) * lorch / rr
    np.testing.assert_allclose(g[0, 2:], want, rtol=2e-5, atol=2e-6)
    assert g[0, 1] == pytest.approx(0.5 + sfac * 2.0 * qq * qq * lorch, rel=2e-5)   # r = 0: q^2
    assert np.all(g[1, 1:] == F32(-1.0))                                          # lorch(pi) = 0


def test_j1qr_textbook_integral_and_foqols_telescoping():
    L = po.lib("cr")
    rng = np.random.default_rng(2)
    # the integral itself, by quadrature, where qr is small enough that the float32 product
    # q*r the reference forms does not matter at 1e-6
    for qq, rr in zip(rng.uniform(0.1, 3, 10), rng.uniform(0.05, 3, 10)):
        q, r = float(F32(qq)), float(F32(rr))
        t = np.linspace(0.0, r, 20001)
        quad = np.trapezoid(t * np.sin(q * t) / q, t)
        assert L.ora_j1qr(F32(q), F32(r)) == pytest.approx(quad, rel=1e-6, abs=1e-12)
    # over the whole range, the closed form at the reference's float32 qr and q^3
    for qq, rr in zip(rng.uniform(0.1, 30, 20), rng.uniform(0.05, 20, 20)):
        q, r = F32(qq), F32(rr)
        x, q3 = float(F32(q * r)), float(F32(F32(q * q) * q))
        want = (np.sin(x) - x * np.cos(x)) / q3
        assert L.ora_j1qr(q, r) == pytest.approx(want, rel=1e-12, abs=1e-15)
    assert L.ora_j1qr(F32(0.0), F32(2.0)) == pytest.approx(8.0 / 3.0, rel=1e-7)
    # foqols: with no broadening, sum_j fo[j] ycoh[j] = 4 pi rho (j1qr(q, r_outer) - j1qr(q, 0))
    nq, npt1 = 12, 80
    q = np.zeros(nq + 1, F32)
    q[1:] = (0.5 + 0.7 * np.arange(nq)).astype(F32)
    x = np.zeros(npt1 + 1, F32)
    x[1:] = (0.05 * np.arange(npt1)).astype(F32)
    ycoh = np.zeros(npt1 + 1, F32)
    ycoh[1:] = np.abs(rng.normal(1, 0.1, npt1)).astype(F32)
    ldf = npt1 + 1
    rho = F32(0.1)
    fo = np.zeros((nq + 1) * ldf, F32)
    L.ora_foqols(nq, po.Pf(q), npt1, po.Pf(x), po.Pf(ycoh), rho, F32(0), F32(0), F32(0), F32(0), F32(4.0),
                 po.Pf(fo), ldf)
    fo = fo.reshape(nq + 1, ldf)
    r1 = np.float32(0.0)
    edges = []
    for j in range(1, npt1 + 1):
        r2 = F32(x[j] + F32(x[j] - r1))
        edges.append(r2)
        r1 = r2
    for i in range(1, nq + 1):
        tot = float(np.sum(fo[i, 1:].astype(np.float64) * ycoh[1:]))
        want = 4 * np.pi * float(rho) * (L.ora_j1qr(q[i], edges[-1]) - L.ora_j1qr(q[i], F32(0)))
        assert tot == pytest.approx(want, rel=2e-5, abs=1e-6 * abs(want) + 1e-6)
    # the resolution factor with a Q-proportional width: exp(-(sigma q r)^2 / 2),
    # sigma = fwhmq / (2 sqrt(2 ln 2))
    fo2 = np.zeros((nq + 1) * ldf, F32)
    fwhmq = F32(0.02)
    L.ora_foqols(nq, po.Pf(q), npt1, po.Pf(x), po.Pf(ycoh), rho, F32(0), fwhmq, F32(0), F32(0), F32(4.0),
                 po.Pf(fo2), ldf)
    fo2 = fo2.reshape(nq + 1, ldf)
    sig = 0.5 * float(fwhmq) / np.sqrt(2 * np.log(2))
    i, j = 7, 60
    fac = np.exp(-0.5 * (sig * float(q[i]) * float(x[j])) ** 2)
    assert fo2[i, j] == pytest.approx(fo[i, j] * fac, rel=1e-5)


def test_poisson_basis_q_space_is_the_fourier_transform_of_r_space():
    L = po.lib("cr")
    nterms, dmin, dstep, dmax, sigmab, rho = 8, F32(0.0), F32(1.2), F32(12.0), F32(0.35), F32(0.1)
    # r-space on a fine grid (the reference's log form), Q-space at a few Q
    nr = 6000
    rgrid = np.zeros(nr + 1, F32)
    rgrid[1:] = (0.005 * np.arange(nr)).astype(F32)
    er = np.zeros(nr * (nterms + 1) + 1, F32)
    rc = L.ora_calc_pwrexpr_grad(nr, nterms, dmin, dstep, dmax, sigmab, rho, po.Pf(rgrid), po.Pf(er), F32(0), F32(1),
                                 F32(0))
    assert rc >= 0
    qs = np.zeros(6, F32)
    qs[1:] = np.array([0.3, 1.0, 2.5, 5.0, 9.0], F32)
    eq = np.zeros(5 * (nterms + 1) + 1, F32)
    L.ora_calc_pwrexpq(5, nterms, dmin, dstep, dmax, sigmab, rho, po.Pf(qs), po.Pf(eq), F32(0))
    rr = rgrid[1:].astype(np.float64)
    h = np.diff(rr)
    for m in range(1, nterms + 1):
        # calc_pwrexpr_grad stores the r-derivative of the basis function; the function
        # itself is minus its integral from r to the (vanishing) far end
        grad = np.array([er[(i - 1) * nterms + m] for i in range(1, nr + 1)], np.float64)
        seg = 0.5 * (grad[1:] + grad[:-1]) * h
        pr = np.concatenate([-np.cumsum(seg[::-1])[::-1], [0.0]])
        for k in range(1, 6):
            qk = float(qs[k])
            integrand = 4 * np.pi * float(rho) * rr * rr * pr * np.sinc(qk * rr / np.pi)
            ft = np.trapezoid(integrand, rr)
            assert ft == pytest.approx(float(eq[(k - 1) * nterms + m]), rel=2e-3, abs=2e-4), (m, qk)


def _oracle_box(name, flavour="cr"):
    s = H.system(name)
    ref = H.oracle_fmolec1(name, flavour)
    return s, ref


def _sweep(L, s, ref, simt, n, idum=-21):
    ob = ref["ob"]
    wr, wt = H.initial_steps(s)
    nwh = np.zeros(2 * s.nmoltype, np.int32)
    nmv = np.zeros(2 * s.nmoltype, np.int32)
    nsr = np.zeros(s.nmol, np.int32)
    mc = po.MCState(po.Pf(wr), po.Pf(wt), po.Pi(nwh), po.Pi(nmv), po.Pi(nsr), F32(simt), F32(1), F32(1), F32(0.75),
                    1)
    rng = po.Ran1.seeded(idum)
    fadds, nacc = C.c_float(0.0), C.c_int()
    accs = []
    for _ in range(n):
        L.ora_molfit(C.byref(ob.box), po.Pf(ob.xa), po.Pf(ob.ya), po.Pf(ob.za), C.byref(ob.centres),
                     C.byref(ref["ctab"]), C.byref(ref["cells"]), C.byref(ref["ledger"]), C.byref(mc), C.byref(rng),
                     C.byref(fadds), C.byref(nacc))
        accs.append(nacc.value)
    return float(fadds.value), accs


def _ledger_total(ref, s):
    d = H.ledger_dict(ref["ledger"], s.nmol)
    nself = int(np.count_nonzero(s.eintra))
    return float(np.sum(d["ergpair"][nself:].astype(np.float64))), d


def test_molfit_ledger_matches_fresh_pair_energies_and_books_the_energy_change():
    L = po.lib("cr")
    name = "tiny_13_12_3"
    s, ref = _oracle_box(name)
    e0, _ = _ledger_total(ref, s)
    fadds, accs = _sweep(L, s, ref, s.simt, 3)
    assert sum(accs) > 0
    e1, d = _ledger_total(ref, s)
    # every stored pair equals en_mole_pair at the current coordinates, bit for bit, in
    # the summation order of the molecule that wrote it last (the mover is the outer loop)
    ob = ref["ob"]
    for im in range(s.nmol):
        for k in range(1, d["nergpairm"][im]):
            jm = int(d["molergpair"][im, k])
            ea = F32(L.ora_en_mole_pair(C.byref(ob.box), C.byref(ob.centres), C.byref(ref["ctab"]), im, jm))
            eb = F32(L.ora_en_mole_pair(C.byref(ob.box), C.byref(ob.centres), C.byref(ref["ctab"]), jm, im))
            got = d["ergpair"][d["ergpairref"][im, k]].view(np.int32)
            assert got in (ea.view(np.int32), eb.view(np.int32)), (im, jm)
    # fadds (accepted fdiff / nmol per sweep) books the change of the pair-energy total
    booked = fadds * s.nmol
    assert booked == pytest.approx(e1 - e0, rel=1e-4, abs=1e-3 * (abs(e0) + abs(e1)) / max(1, len(d["ergpair"])))


def test_molfit_temperature_limits():
    L = po.lib("cr")
    name = "tiny_13_12_3"
    s, ref = _oracle_box(name)
    e0, _ = _ledger_total(ref, s)
    _, accs = _sweep(L, s, ref, 1e-12, 2)
    e1, _ = _ledger_total(ref, s)
    assert e1 <= e0 + 1e-6 * abs(e0)                  # only downhill moves at T -> 0
    s, ref = _oracle_box(name)
    _, accs = _sweep(L, s, ref, 1e30, 2, idum=-22)
    assert accs == [s.nmol, s.nmol]                   # every move at T -> infinity
