"""The oracle of the learned surrogate (oracle/ora_surrogate.c) and of delayed acceptance
(ora_molfit_da), checked on the CPU against what fixes them independently:

* descriptors against a vectorised numpy evaluation of the same definition and against
  closed forms (one pair at a Gaussian centre, pairs beyond the cutoff, channel symmetry);
* the network's analytic gradient against central finite differences of its loss;
* delayed acceptance with dEs = 0 against plain molfit, bit for bit (stage 1 always
  passes and stage 2 is then the plain test), and with dEs = +inf (nothing accepted);
* the stage-1 uniforms: range and moments."""

import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as po
from tests import helpers as H

F32 = np.float32


def _desc_numpy(xm, ym, zm, tm, xn, yn, zn, tn, npr, npairs, dmax, ng, rc, r0, dmu, width, xscale):
    dx = xn[None, :] - xm[:, None]
    dy = yn[None, :] - ym[:, None]
    dz = zn[None, :] - zm[:, None]
    mic = lambda v: (v - np.where(v >= 0, np.floor(v + F32(0.5)), np.ceil(v - F32(0.5))).astype(F32)).astype(F32)
    dx, dy, dz = mic(dx.astype(F32)), mic(dy.astype(F32)), mic(dz.astype(F32))
    r = (np.sqrt((dx * dx + dy * dy + dz * dz).astype(F32)).astype(F32) * F32(dmax)).astype(F32)
    out = np.zeros(npairs * ng)
    # the oracle takes its parameters as float32
    rc, r0, dmu, width, xscale = (float(F32(v)) for v in (rc, r0, dmu, width, xscale))
    a_idx, b_idx = np.nonzero(r < F32(rc))
    for a, b in zip(a_idx, b_idx):
        rr = float(r[a, b])
        ch = npr[tm[a] * int(np.sqrt(len(npr))) + tn[b]]
        mu = r0 + np.arange(ng) * dmu
        out[ch * ng:(ch + 1) * ng] += np.exp(-((rr - mu) / width) ** 2) * 0.5 * (np.cos(np.pi * rr / rc) + 1.0)
    return out * xscale


def _call_desc(L, xm, ym, zm, tm, xn, yn, zn, tn, npr, ntype, npairs, dmax, ng, rc, r0, dmu, width, xscale):
    out = np.zeros(npairs * ng)
    L.ora_descriptor(len(xm), po.Pf(xm), po.Pf(ym), po.Pf(zm), po.Pi(tm), len(xn), po.Pf(xn), po.Pf(yn),
                     po.Pf(zn), po.Pi(tn), po.Pi(npr), ntype, npairs, F32(dmax), ng, F32(rc), F32(r0), F32(dmu),
                     F32(width), F32(xscale), out.ctypes.data_as(po.dp))
    return out


def test_descriptor_against_numpy_and_closed_forms():
    L = po.lib("cr")
    rng = np.random.default_rng(3)
    ntype, npairs = 3, 6
    npr = np.zeros((ntype, ntype), np.int32)
    ip = 0
    for i in range(ntype):
        for j in range(i, ntype):
            npr[i, j] = npr[j, i] = ip
            ip += 1
    npr = npr.ravel()
    dmax, ng, rc, r0, width, xs = 20.0, 16, 8.0, 1.0, 0.4, 0.5
    dmu = (rc - r0) / (ng - 1)
    xm, ym, zm = (rng.uniform(-0.5, 0.5, 4).astype(F32) for _ in range(3))
    xn, yn, zn = (rng.uniform(-0.5, 0.5, 40).astype(F32) for _ in range(3))
    tm = rng.integers(0, ntype, 4).astype(np.int32)
    tn = rng.integers(0, ntype, 40).astype(np.int32)
    got = _call_desc(L, xm, ym, zm, tm, xn, yn, zn, tn, npr, ntype, npairs, dmax, ng, rc, r0, dmu, width, xs)
    want = _desc_numpy(xm, ym, zm, tm, xn, yn, zn, tn, npr, npairs, dmax, ng, rc, r0, dmu, width, xs)
    assert np.count_nonzero(want) > 0
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14)
    # one pair at the centre of Gaussian 5, along x: that Gaussian is the cutoff factor
    r = F32(r0 + 5 * dmu)
    xm1, z1 = np.array([0.0], F32), np.array([0.0], F32)
    xn1 = np.array([r / F32(dmax)], F32)
    one = _call_desc(L, xm1, z1, z1, np.array([1], np.int32), xn1, z1, z1, np.array([2], np.int32), npr, ntype,
                     npairs, dmax, ng, rc, r0, dmu, width, 1.0)
    ch = npr[1 * ntype + 2]
    rr = float(F32(F32(np.sqrt(F32(xn1[0] * xn1[0]))) * F32(dmax)))
    mu5 = float(F32(r0)) + 5 * float(F32(dmu))
    assert one[ch * ng + 5] == pytest.approx(np.exp(-((rr - mu5) / float(F32(width))) ** 2) *
                                            0.5 * (np.cos(np.pi * rr / rc) + 1), rel=1e-12)
    assert np.count_nonzero(one[:ch * ng]) == 0 and np.count_nonzero(one[(ch + 1) * ng:]) == 0
    # swapping the two atoms' types keeps the channel (npairref is symmetric)
    sw = _call_desc(L, xm1, z1, z1, np.array([2], np.int32), xn1, z1, z1, np.array([1], np.int32), npr, ntype,
                    npairs, dmax, ng, rc, r0, dmu, width, 1.0)
    np.testing.assert_array_equal(one, sw)
    # beyond the cutoff nothing is added
    far = np.array([F32(rc + 0.01) / F32(dmax)], F32)
    zero = _call_desc(L, xm1, z1, z1, np.array([1], np.int32), far, z1, z1, np.array([2], np.int32), npr, ntype,
                      npairs, dmax, ng, rc, r0, dmu, width, 1.0)
    assert not zero.any()


def _mlp_loss(L, d, h, w, bf, x, t):
    grad = np.zeros(len(w))
    loss = L.ora_mlp_grad(d, h, po.Pf(w), bf, len(t), po.Pf(x), d, po.Pf(t), grad.ctypes.data_as(po.dp))
    return loss, grad


def test_mlp_gradient_against_finite_differences():
    L = po.lib("cr")
    d, h, npair = 32, 128, 6
    rng = np.random.default_rng(5)
    n = h * d + h + h * h + h + h + 1
    w = (rng.normal(0, 0.3, n)).astype(F32)
    x = rng.normal(0, 1, (2 * npair, d)).astype(F32).ravel()
    t = rng.normal(0, 1, npair).astype(F32)
    loss, grad = _mlp_loss(L, d, h, w, 0, x, t)
    assert loss > 0
    # finite differences in double through a float64 copy of the loss
    blocks = {"W1": (0, h * d), "b1": (h * d, h), "W2": (h * d + h, h * h), "b2": (h * d + h + h * h, h),
              "w3": (h * d + 2 * h + h * h, h)}
    idx = [o + int(k) for o, sz in blocks.values() for k in rng.choice(sz, 6, replace=False)]
    for i in idx:
        eps = 1e-3 * max(1.0, abs(float(w[i])))
        wp, wm = w.copy(), w.copy()
        wp[i] = F32(w[i] + eps)
        wm[i] = F32(w[i] - eps)
        step = float(wp[i]) - float(wm[i])
        lp, _ = _mlp_loss(L, d, h, wp, 0, x, t)
        lm, _ = _mlp_loss(L, d, h, wm, 0, x, t)
        fd = (lp - lm) / step
        assert abs(fd - grad[i]) <= 1e-4 * max(1e-3, abs(grad[i])) + 1e-7, (i, fd, grad[i])
    # b3 cancels in f(new) - f(old)
    assert grad[-1] == 0.0
    # the forward pass agrees with the gradient routine's loss
    y = np.zeros(2 * npair)
    L.ora_mlp_forward(d, h, po.Pf(w), 0, 2 * npair, po.Pf(x), d, y.ctypes.data_as(po.dp))
    assert np.mean((y[1::2] - y[0::2] - t) ** 2) == pytest.approx(loss, rel=1e-12)


def test_stage1_uniforms():
    L = po.lib("cr")
    u = np.array([L.ora_da_uniform(12345, m) for m in range(20000)])
    assert u.min() > 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.01 and abs(u.var() - 1 / 12) < 0.003
    v = np.array([L.ora_da_uniform(12346, m) for m in range(20000)])
    assert abs(np.corrcoef(u, v)[0, 1]) < 0.03


def _oracle_state(name, flavour="cr"):
    s = H.system(name)
    ref = H.oracle_fmolec1(name, flavour)
    wr, wt = H.initial_steps(s)
    st = dict(ref=ref, ob=ref["ob"], wr=wr, wt=wt, nwh=np.zeros(2 * s.nmoltype, np.int32),
              nmv=np.zeros(2 * s.nmoltype, np.int32), nsr=np.zeros(s.nmol, np.int32))
    st["mc"] = po.MCState(po.Pf(st["wr"]), po.Pf(st["wt"]), po.Pi(st["nwh"]), po.Pi(st["nmv"]), po.Pi(st["nsr"]),
                          s.simt, F32(1), F32(1), F32(0.75), 0)
    return s, st


def _sweep(L, st, rng, des=None, seed=0, target=None, fdiff=None):
    ob, ref = st["ob"], st["ref"]
    fadds, nacc, nrej = C.c_float(0.0), C.c_int(), C.c_int()
    L.ora_molfit_da(C.byref(ob.box), po.Pf(ob.xa), po.Pf(ob.ya), po.Pf(ob.za), C.byref(ob.centres),
                    C.byref(ref["ctab"]), C.byref(ref["cells"]), C.byref(ref["ledger"]), C.byref(st["mc"]),
                    C.byref(rng), C.byref(fadds), C.byref(nacc), po.Pf(des) if des is not None else None,
                    C.c_ulonglong(seed), C.byref(nrej), target.ctypes.data_as(po.dp) if target is not None else None,
                    po.Pf(fdiff) if fdiff is not None else None)
    return nacc.value, nrej.value


def test_delayed_acceptance_with_zero_surrogate_is_plain_molfit():
    name = "tiny_13_12_3"
    L = po.lib("cr")
    s, a = _oracle_state(name)
    _, b = _oracle_state(name)
    ra, rb = po.Ran1.seeded(-99), po.Ran1.seeded(-99)
    zero = np.zeros(s.nmol, F32)
    for _ in range(3):
        na, _ = _sweep(L, a, ra)
        nb, rej = _sweep(L, b, rb, des=zero, seed=7)
        assert na == nb and rej == 0
    for k in ("xa", "ya", "za", "xadd", "yadd", "zadd"):
        assert np.array_equal(getattr(a["ob"], k).view(np.int32), getattr(b["ob"], k).view(np.int32)), k
    np.testing.assert_array_equal(ra.as_array(), rb.as_array())
    la, lb = H.ledger_dict(a["ref"]["ledger"], s.nmol), H.ledger_dict(b["ref"]["ledger"], s.nmol)
    assert np.array_equal(la["ergpair"].view(np.int32), lb["ergpair"].view(np.int32))


def test_delayed_acceptance_rejecting_surrogate_accepts_nothing():
    name = "tiny_13_12_3"
    L = po.lib("cr")
    s, a = _oracle_state(name)
    x0 = a["ob"].xadd.copy()
    rng = po.Ran1.seeded(-5)
    nacc, nrej = _sweep(L, a, rng, des=np.full(s.nmol, np.inf, F32), seed=3)
    assert nacc == 0 and nrej == s.nmol
    assert np.array_equal(a["ob"].xadd, x0)


def test_later_partner_target_against_en_change():
    """target = the part of en_change's sum over partners later in the chain.  With every
    move rejected (dEs = +inf) no partner has moved when a move is made, so the first move
    of the chain (all partners later) has target = fdiff, and the last (none later) has
    target = 0; every target lies within sum |terms| of fdiff."""
    name = "c0"
    L = po.lib("cr")
    s, a = _oracle_state(name)
    cl = a["ref"]["cells"]
    cnt = np.ctypeslib.as_array(cl.nmolsincell, shape=(cl.ncell,))
    mic = np.ctypeslib.as_array(cl.molsincell, shape=(cl.ncell * cl.maxcellno,)).reshape(cl.ncell, cl.maxcellno)
    order = np.concatenate([mic[c, :cnt[c]] for c in range(cl.ncell)])
    target = np.zeros(s.nmol)
    fdiff = np.zeros(s.nmol, F32)
    _sweep(L, a, po.Ran1.seeded(-11), des=np.full(s.nmol, np.inf, F32), seed=1, target=target, fdiff=fdiff)
    first, last = int(order[0]), int(order[-1])
    n = a["ref"]["ledger"].nnbrefmax
    bound = 2 * n * 2.0 ** -24 * max(1.0, abs(float(fdiff[first]))) * 50
    assert abs(target[first] - float(fdiff[first])) <= bound, (target[first], fdiff[first])
    assert target[last] == 0.0
    assert np.all(np.isfinite(target)) and np.count_nonzero(target) > s.nmol // 2
