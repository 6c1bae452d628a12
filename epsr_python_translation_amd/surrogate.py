"""Learned Delta-E surrogate and delayed-acceptance MC, on the GPU.

The reference has no learned part; this module adds one beside molfit
(src/epsr26/library/molfit.py:51-541) without changing what molfit samples:

* descriptors: for the move of the molecule at chain position k, Gaussian radial
  functions of every atom-pair distance below ``rc`` between the molecule (at its current
  or trial coordinates) and its ledger partners that come later in the chain, one channel
  per type pair (kernel ``surr_descriptors``);
* the network f: two hidden layers of 128 SiLU units, BF16 weights and activations with
  FP32 accumulation on MFMA for the forward pass, FP32 MFMA for the gradient, Adam;
  dEs = f(d(trial)) - f(d(current)) is trained on the exact energy change over the same
  later partners, which the plain GPU cycle records while it runs;
* delayed acceptance (``epsr_molfit_da``): stage 1 accepts with min(1, exp(-dEs/simt));
  stage 2 accepts with min(1, exp(-(dE - dEs)/simt)) on the exact dE.  Because the later
  partners have not moved yet when move k is made, dEs is a function of the current state
  and the trial alone and changes sign when they are swapped, so the two stages keep
  detailed balance with respect to the exact energy; moves that stage 1 rejects skip the
  exact pair energies and receive no pushes in the chain resolver.

Everything runs through libepsr_mi355x.so; torch only holds device memory.
"""

from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np
import torch

from . import _native
from . import device as D

F32 = np.float32
H = 128


@dataclasses.dataclass
class DescriptorSpec:
    """Gaussians r0 + g*dmu (g < ng), width ``width``, cutoff ``rc`` (Angstrom)."""
    ng: int = 32
    rc: float = 8.0
    r0: float = 1.0
    width: float = 0.35
    xscale: float = 0.25

    @property
    def dmu(self):
        return (self.rc - self.r0) / (self.ng - 1)

    def params(self):
        return _native.DescParams(int(self.ng), F32(self.rc), F32(self.r0), F32(self.dmu), F32(self.width),
                                  F32(self.xscale))

    def width_for(self, npairs):
        """row width: npairs * ng rounded up to a multiple of 32 (the network's d)"""
        return (npairs * self.ng + 31) // 32 * 32


class Network:
    """f(x) = w3 . silu(W2 silu(W1 x + b1) + b2) + b3 with its fp32 master parameters,
    their bf16 copy and Adam moments in device memory."""

    def __init__(self, ctx: D.Context, d: int, seed: int = 0, w: np.ndarray | None = None):
        self.ctx = ctx
        self.d = int(d)
        self.n = int(ctx.lib.epsr_mlp_nparams(self.d, H))
        if w is None:
            w = self.initial(self.d, seed)
        assert w.shape == (self.n,)
        self.w = ctx.tensor(np.ascontiguousarray(w, F32))
        self.wb = ctx.zeros(H * self.d + H * H, torch.int16)
        self.m = ctx.zeros(self.n)
        self.v = ctx.zeros(self.n)
        self.grad_buf = ctx.zeros(self.n)
        self.struct = _native.MLP(self.d, H, D._ptr(self.w), D._ptr(self.wb), D._ptr(self.m), D._ptr(self.v), 0)
        ctx.check(ctx.lib.epsr_mlp_sync(ctx.h, C.byref(self.struct)))

    @staticmethod
    def initial(d, seed=0):
        """uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights, zero biases"""
        rng = np.random.default_rng(seed)
        W1 = rng.uniform(-1, 1, (H, d)) / np.sqrt(d)
        W2 = rng.uniform(-1, 1, (H, H)) / np.sqrt(H)
        w3 = rng.uniform(-1, 1, H) / np.sqrt(H)
        return np.concatenate([W1.ravel(), np.zeros(H), W2.ravel(), np.zeros(H), w3, [0.0]]).astype(F32)

    def parameters(self):
        return self.w.cpu().numpy()

    def forward(self, x: torch.Tensor, ldx: int | None = None):
        """f of every row of x (bf16 bits, int16 [rows, ldx])"""
        ldx = int(ldx or x.shape[-1])
        rows = x.numel() // ldx
        y = self.ctx.empty(rows)
        self.ctx.check(self.ctx.lib.epsr_mlp_forward(self.ctx.h, C.byref(self.struct), rows, D._ptr(x), ldx,
                                                     D._ptr(y)))
        return y

    def grad(self, x: torch.Tensor, target: torch.Tensor, ldx: int | None = None, tclip: float = 0.0, out=None):
        """(gradient, loss) of mean (f(x[2i+1]) - f(x[2i]) - t_i)^2, t_i the target clamped to
        [-tclip, tclip] when tclip > 0"""
        ldx = int(ldx or x.shape[-1])
        npair = int(target.numel())
        g = self.grad_buf if out is None else out
        loss = C.c_float()
        self.ctx.check(self.ctx.lib.epsr_mlp_grad(self.ctx.h, C.byref(self.struct), npair, D._ptr(x), ldx,
                                                  D._ptr(target), F32(tclip), D._ptr(g), C.byref(loss)))
        return g, float(loss.value)

    def adam(self, grad, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.ctx.check(self.ctx.lib.epsr_mlp_adam(self.ctx.h, C.byref(self.struct), D._ptr(grad), F32(lr),
                                                  F32(beta1), F32(beta2), F32(eps)))

    def train(self, X, T, ldx, epochs=4, batch=4096, lr=1e-3, tclip=0.0, seed=0, log=None):
        """Adam over contiguous batches of pairs (rows 2i, 2i+1 of X; T[i]), shuffled per
        epoch; returns the mean loss of each epoch.  Targets beyond +-tclip (moves into
        overlaps, energies of 1e10 and more) are clamped: stage 1 only has to rank moves
        whose acceptance is in doubt."""
        npair = int(T.numel())
        rng = np.random.default_rng(seed)
        starts = np.arange(0, npair - batch + 1, batch)
        hist = []
        for ep in range(epochs):
            tot = 0.0
            for s0 in rng.permutation(starts):
                s0 = int(s0)
                g, loss = self.grad(X[2 * s0:2 * (s0 + batch)], T[s0:s0 + batch], ldx, tclip=tclip)
                self.adam(g, lr=lr)
                tot += loss
            hist.append(tot / max(1, len(starts)))
            if log:
                log(f"epoch {ep}: loss {hist[-1]:.5g}")
        return hist


class Recorder:
    """Device buffers that one MC cycle fills by chain position: descriptors, the exact
    later-partner energy change, dEs and the stage-1 outcome."""

    def __init__(self, ctx: D.Context, nmol: int, spec: DescriptorSpec, ldx: int, x=None, target=None,
                 des=True, passed=True):
        self.spec = spec
        self.dp = spec.params()
        self.ldx = ldx
        self.x = x
        self.target = target
        self.des = ctx.zeros(nmol) if des else None
        self.passed = ctx.zeros(nmol, torch.int32) if passed else None

    def struct(self, x=None, target=None):
        x = self.x if x is None else x
        target = self.target if target is None else target
        return _native.MCRecord(C.pointer(self.dp), int(self.ldx), D._ptr(x), D._ptr(target), D._ptr(self.des),
                                D._ptr(self.passed))


class DelayedAcceptance:
    """The two-stage MC cycle driven by a trained Network; ``seed`` advances every cycle
    so that stage 1 draws fresh uniforms."""

    def __init__(self, net: Network | None, spec: DescriptorSpec, seed: int = 1, des=None):
        self.net = net
        self.spec = spec
        self.seed = int(seed)
        self.des = des
        self._da = _native.DA(C.pointer(net.struct) if net is not None else None, spec.params(), 0,
                              D._ptr(des) if des is not None else None)

    def struct(self):
        self._da.seed = C.c_ulonglong(self.seed & 0xFFFFFFFFFFFFFFFF).value
        self.seed += 1
        return self._da

    def sweep(self, ctx, box, tables, ledger, mc, ran1_state, record: Recorder | None = None):
        rec = record.struct() if record is not None else None
        return D.molfit_da(ctx, box, tables, ledger, mc, ran1_state, da=self.struct(), record=rec)


def collect(ctx: D.Context, box, tables, ledger, mc, ran1_state, spec: DescriptorSpec, nsweeps: int, ldx=None):
    """Training pairs from plain GPU MC: nsweeps cycles, each recording the descriptors
    and the exact later-partner energy change of every move.  Returns (X int16 [2 N, ldx],
    T [N]) with N = nsweeps * nmol."""
    ldx = int(ldx or spec.width_for(box.npairs))
    nmol = box.nmol
    X = ctx.zeros(2 * nsweeps * nmol * ldx, torch.int16).view(2 * nsweeps * nmol, ldx)
    T = ctx.zeros(nsweeps * nmol)
    rec = Recorder(ctx, nmol, spec, ldx, des=False, passed=False)
    for i in range(nsweeps):
        st = rec.struct(x=X[2 * i * nmol:2 * (i + 1) * nmol], target=T[i * nmol:(i + 1) * nmol])
        D.molfit_da(ctx, box, tables, ledger, mc, ran1_state, da=None, record=st)
    return X, T
