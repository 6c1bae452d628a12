"""Potential tables on the GPU (truncation, generate_rp, generate_ep, calc_pwrexpq)
against the oracle: bit-identical to the correctly rounded build, within 1e-5 of the
glibc build (relative to the table's scale; the forward-difference gradients relative to
the value scale over one grid step)."""

import ctypes as C

import numpy as np
import pytest
import torch

from oracle import pyoracle as po
from tests import helpers as H

F32 = np.float32
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from epsr_python_translation_amd import device as D
    return D.Context(0)


def _bits_equal_frac(a, b):
    return float(np.mean(np.asarray(a, F32).view(np.int32) == np.asarray(b, F32).view(np.int32)))


@pytest.mark.parametrize("name", ["c0", "c1"])
def test_truncation_and_generate_rp(ctx, name):
    from epsr_python_translation_amd import device as D
    from epsr_python_translation_amd import pipeline as PL
    print(f"generate_rp {name}", flush=True)
    s, g = H.system(name), H.grid(name)
    for flavour in ("cr", "glibc"):
        wfac_r, wfacc_r = H.truncation(name, flavour)
        x = ctx.tensor(g.x)
        wf, wfc = ctx.zeros(g.ld), ctx.zeros(g.ld)
        ctx.check(ctx.lib.epsr_truncation(ctx.h, g.npt, D._ptr(x), F32(s.rminpt), F32(s.rmaxpt), D._ptr(wf),
                                          D._ptr(wfc)))
        wf_h, wfc_h = wf.cpu().numpy(), wfc.cpu().numpy()
        if flavour == "cr":
            assert _bits_equal_frac(wf_h[1:], wfac_r[1:]) == 1.0
            assert _bits_equal_frac(wfc_h[1:], wfacc_r[1:]) == 1.0
        assert np.abs(wf_h[1:] - wfac_r[1:]).max() <= 2.4e-7
        # generate_rp with the oracle's truncation functions on both sides
        ld, npairs = g.ld, s.npairs
        ref = {k: np.zeros(npairs * ld, F32) for k in ("epot", "ecoul", "epot1", "ecoul1")}
        qr = np.zeros(npairs, F32)
        po.lib(flavour).ora_generate_rp(g.npt, g.npt2, po.Pf(g.x), po.Pf(g.xstar), po.Pf(wfac_r), po.Pf(wfacc_r),
                                        npairs, po.Pf(s.ecor), po.Pf(s.dcor), po.Pf(s.qcor), po.Pf(qr), po.Pf(qr),
                                        F32(s.power), F32(1.0), None, None, F32(0.3), F32(0.0), po.Pf(ref["epot"]),
                                        po.Pf(ref["ecoul"]), po.Pf(ref["epot1"]), po.Pf(ref["ecoul1"]), ld)
        ins = [ctx.tensor(a.astype(F32)) for a in (s.ecor, s.dcor, s.qcor)]
        p = PL._native_rp_params(s, ins)
        out = {k: ctx.zeros(npairs * ld) for k in ref}
        wft, wfct, xs = ctx.tensor(wfac_r), ctx.tensor(wfacc_r), ctx.tensor(g.xstar)
        ctx.check(ctx.lib.epsr_generate_rp(ctx.h, g.npt, g.npt2, D._ptr(x), D._ptr(xs), D._ptr(wft), D._ptr(wfct),
                                           C.byref(p), D._ptr(out["epot"]), D._ptr(out["ecoul"]),
                                           D._ptr(out["epot1"]), D._ptr(out["ecoul1"]), ld))
        got = {k: v.cpu().numpy().reshape(npairs, ld)[:, 1:g.npt + 1] for k, v in out.items()}
        want = {k: v.reshape(npairs, ld)[:, 1:g.npt + 1] for k, v in ref.items()}
        for k in ("epot", "ecoul"):
            frac = _bits_equal_frac(got[k], want[k])
            scale = np.abs(want[k]).max(axis=1, keepdims=True) + 1e-30
            rel = (np.abs(got[k].astype(np.float64) - want[k]) / scale).max()
            print(f"  {flavour} {k}: bit-identical {frac:.5f}, max err/scale {rel:.2e}")
            if flavour == "cr":
                assert frac >= 0.9999
            assert rel <= 1e-5
        dx = float(g.xstar[3] - g.xstar[2])
        for k, base in (("epot1", "epot"), ("ecoul1", "ecoul")):
            scale = np.abs(want[base]).max(axis=1, keepdims=True) / dx + 1e-30
            rel = (np.abs(got[k].astype(np.float64) - want[k]) / scale).max()
            assert rel <= 1e-5, (k, rel)


@pytest.mark.parametrize("name", ["c0", "c4"])
def test_generate_ep(ctx, name):
    from epsr_python_translation_amd import device as D
    from epsr_python_translation_amd import pipeline as PL
    print(f"generate_ep {name}", flush=True)
    s, g = H.system(name), H.grid(name)
    ld, npairs = g.ld, s.npairs
    for flavour in ("cr", "glibc"):
        t = H.oracle_tables(name, flavour)
        wfac = t["wfac"]
        ref = {k: np.zeros(npairs * ld, F32) for k in ("xavs", "xavs1", "xavtot", "xavtot1")}
        pc_ref = np.ascontiguousarray(s.pcoefp.astype(F32).copy())
        ir = np.arange(npairs, dtype=np.int32)
        ea = C.c_float()
        pf_ref = po.lib(flavour).ora_generate_ep(
            g.npt, g.npt2, po.Pf(g.x), po.Pf(g.xstar), po.Pf(wfac), npairs, po.Pi(ir), s.nterms,
            po.Pf(pc_ref.reshape(-1)), s.nterms + 1, F32(s.pdmin), F32(s.pdstep), F32(s.rmaxpt), F32(s.psigma2),
            F32(s.rho), F32(0.0), F32(s.dmax), F32(10.0), F32(0.0), F32(0.0), po.Pf(t["ecoul"]), po.Pf(ref["xavs"]),
            po.Pf(ref["xavs1"]), po.Pf(ref["xavtot"]), po.Pf(ref["xavtot1"]), ld, C.byref(ea))
        out = {k: ctx.zeros(npairs * ld) for k in ref}
        pc = np.ascontiguousarray(s.pcoefp.astype(F32).copy())
        p = PL.EPParams(F32(s.pdmin), F32(s.pdstep), F32(s.rmaxpt), F32(s.psigma2), F32(s.rho), F32(0.0),
                        F32(s.dmax), F32(10.0), F32(0.0), F32(0.0))
        pf, ea2 = C.c_float(), C.c_float()
        x, xs, wf, ec = ctx.tensor(g.x), ctx.tensor(g.xstar), ctx.tensor(wfac), ctx.tensor(t["ecoul"])
        ctx.check(ctx.lib.epsr_generate_ep(ctx.h, g.npt, g.npt2, D._ptr(x), D._ptr(xs), D._ptr(wf), npairs,
                                           ir.ctypes.data_as(C.POINTER(C.c_int)), s.nterms,
                                           pc.ctypes.data_as(C.POINTER(C.c_float)), s.nterms + 1, C.byref(p),
                                           D._ptr(ec), D._ptr(out["xavs"]), D._ptr(out["xavs1"]),
                                           D._ptr(out["xavtot"]), D._ptr(out["xavtot1"]), ld, C.byref(pf),
                                           C.byref(ea2)))
        assert F32(pf.value) == F32(pf_ref) and F32(ea2.value) == F32(ea.value)
        assert np.array_equal(pc, pc_ref)
        for k in ref:
            a = out[k].cpu().numpy().reshape(npairs, ld)[:, 1:g.npt + 1]
            b = ref[k].reshape(npairs, ld)[:, 1:g.npt + 1]
            frac = _bits_equal_frac(a, b)
            base = k.rstrip("1")
            scale = np.abs(ref[base].reshape(npairs, ld)[:, 1:g.npt + 1]).max(axis=1, keepdims=True) + 1e-30
            if k.endswith("1"):
                scale = scale / float(g.xstar[3] - g.xstar[2])
            rel = (np.abs(a.astype(np.float64) - b) / scale).max()
            print(f"  {flavour} {k}: bit-identical {frac:.5f}, max err/scale {rel:.2e}")
            if flavour == "cr":
                assert frac >= 0.999, k
            assert rel <= 1e-5, k


def test_calc_pwrexpq(ctx):
    print("calc_pwrexpq c0", flush=True)
    from epsr_python_translation_amd import device as D
    s, g = H.system("c0"), H.grid("c0")
    lpt, nterms = g.nq, s.nterms
    for flavour in ("cr", "glibc"):
        ref = np.zeros(lpt * nterms + 1, F32)
        nc = po.lib(flavour).ora_calc_pwrexpq(lpt, nterms, F32(s.pdmin), F32(s.pdstep), F32(s.rmaxpt),
                                              F32(s.psigma2), F32(s.rho), po.Pf(g.q), po.Pf(ref), F32(0.0))
        out = ctx.zeros(lpt * nterms + 1)
        q = ctx.tensor(g.q)
        ncf = C.c_int()
        ctx.check(ctx.lib.epsr_calc_pwrexpq(ctx.h, lpt, nterms, F32(s.pdmin), F32(s.pdstep), F32(s.rmaxpt),
                                            F32(s.psigma2), F32(s.rho), D._ptr(q), D._ptr(out), F32(0.0),
                                            C.byref(ncf)))
        got = out.cpu().numpy()[1:]
        frac = _bits_equal_frac(got, ref[1:])
        rel = np.abs(got.astype(np.float64) - ref[1:]).max() / (np.abs(ref[1:]).max() + 1e-30)
        print(f"  {flavour}: bit-identical {frac:.5f}, max err/scale {rel:.2e}")
        assert ncf.value == nc
        if flavour == "cr":
            assert frac >= 0.999
        assert rel <= 1e-5
