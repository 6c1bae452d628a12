"""Device-resident objects and thin wrappers over the C ABI (libepsr_mi355x.so).

torch is used for HBM allocation and streams only; every computation runs in the HIP
kernels of the library.  Nothing here falls back to a CPU path: a missing library or
GPU raises.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _native

F32 = np.float32


class EpsrError(RuntimeError):
    pass


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


class Context:
    """One epsr_ctx on one GPU, running on torch's current stream of that device."""

    def __init__(self, device: int = 0, timing: bool = False):
        if not torch.cuda.is_available():
            raise EpsrError("no GPU visible: the MI355X path has no CPU fallback")
        self.lib = _native.lib()
        self.device = device
        self.tdev = torch.device("cuda", device)
        h = C.c_void_p()
        rc = self.lib.epsr_ctx_create(device, C.byref(h))
        if rc != 0:
            raise EpsrError(f"epsr_ctx_create({device}) failed with {rc}")
        self.h = h
        self.set_stream(torch.cuda.current_stream(self.tdev))
        if timing:
            self.lib.epsr_ctx_enable_timing(self.h, 1)

    def set_stream(self, stream):
        self.stream = stream
        self.check(self.lib.epsr_ctx_set_stream(self.h, C.c_void_p(stream.cuda_stream)))

    def check(self, rc):
        if rc != 0:
            msg = self.lib.epsr_last_error(self.h)
            raise EpsrError(f"libepsr_mi355x error {rc}: {msg.decode() if msg else ''}")

    def kernel_ms(self, family: str) -> float:
        return float(self.lib.epsr_ctx_last_kernel_ms(self.h, family.encode()))

    def enable_timing(self, on=True):
        self.check(self.lib.epsr_ctx_enable_timing(self.h, 1 if on else 0))

    def close(self):
        if getattr(self, "h", None):
            self.lib.epsr_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def tensor(self, a, dtype=None):
        t = torch.as_tensor(np.ascontiguousarray(a))
        if dtype is not None:
            t = t.to(dtype)
        return t.to(self.tdev).contiguous()

    def empty(self, n, dtype=torch.float32):
        return torch.empty(int(n), dtype=dtype, device=self.tdev)

    def zeros(self, n, dtype=torch.float32):
        return torch.zeros(int(n), dtype=dtype, device=self.tdev)


def _ip(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(C.POINTER(C.c_int))


def _fp(a):
    a = np.ascontiguousarray(a, dtype=F32)
    return a, a.ctypes.data_as(C.POINTER(C.c_float))


class DeviceBox:
    """A box resident in HBM: static topology (epsr_box) plus mutable coordinates as
    torch tensors (atom offsets xa/ya/za, centres xadd/yadd/zadd)."""

    def __init__(self, ctx: Context, *, itype, naf, nal, mtype, npairref, n1extr, n2extr, cfrac, natomr,
                 sizefactor, dmax, cellst, xa, ya, za, xadd, yadd, zadd, nmoltype=None):
        self.ctx = ctx
        self.nmol = int(len(naf))
        self.natom = int(len(itype))
        npairref = np.asarray(npairref, dtype=np.int32)
        self.ntype = int(npairref.shape[0]) if npairref.ndim == 2 else int(round(np.sqrt(npairref.size)))
        self.npairs = int(len(n1extr))
        self.nmoltype = int(nmoltype if nmoltype is not None else int(np.max(mtype)) + 1)
        keep = []
        arrs = {}
        for name, a in (("itype", itype), ("naf", naf), ("nal", nal), ("mtype", mtype),
                        ("npairref", npairref.reshape(-1)), ("n1extr", n1extr), ("n2extr", n2extr)):
            a, p = _ip(a)
            keep.append(a)
            arrs[name] = p
        cf, cfp = _fp(cfrac)
        keep.append(cf)
        self.cfrac = cf.copy()
        self.natomr = F32(natomr)
        self.dmax = F32(dmax)
        self.cellst = F32(cellst)
        self.sizefactor = F32(sizefactor)
        desc = _native.BoxDesc(self.nmol, self.natom, self.ntype, self.npairs, self.nmoltype, arrs["itype"],
                               arrs["naf"], arrs["nal"], arrs["mtype"], arrs["npairref"], arrs["n1extr"],
                               arrs["n2extr"], cfp, F32(natomr), F32(sizefactor), F32(dmax), F32(cellst))
        h = C.c_void_p()
        ctx.check(ctx.lib.epsr_box_create(ctx.h, C.byref(desc), C.byref(h)))
        self.h = h
        self.host = {k: np.array(v) for k, v in (("itype", itype), ("naf", naf), ("nal", nal), ("mtype", mtype),
                                                  ("npairref", npairref), ("n1extr", n1extr), ("n2extr", n2extr))}
        self.xa = ctx.tensor(np.asarray(xa, dtype=F32))
        self.ya = ctx.tensor(np.asarray(ya, dtype=F32))
        self.za = ctx.tensor(np.asarray(za, dtype=F32))
        self.xadd = ctx.tensor(np.asarray(xadd, dtype=F32))
        self.yadd = ctx.tensor(np.asarray(yadd, dtype=F32))
        self.zadd = ctx.tensor(np.asarray(zadd, dtype=F32))

    @classmethod
    def from_system(cls, ctx, s):
        return cls(ctx, itype=s.itype, naf=s.naf, nal=s.nal, mtype=s.mtype, npairref=s.npairref, n1extr=s.n1extr,
                   n2extr=s.n2extr, cfrac=s.cfrac, natomr=s.natomr, sizefactor=1.0, dmax=s.dmax, cellst=s.cellst,
                   xa=s.xa, ya=s.ya, za=s.za, xadd=s.xadd, yadd=s.yadd, zadd=s.zadd, nmoltype=s.nmoltype)

    def coords(self):
        return _native.Coords(_ptr(self.xa), _ptr(self.ya), _ptr(self.za), _ptr(self.xadd), _ptr(self.yadd),
                              _ptr(self.zadd))

    def get_coords(self):
        return {k: getattr(self, k).cpu().numpy() for k in ("xa", "ya", "za", "xadd", "yadd", "zadd")}

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.epsr_box_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ran1_state_array(idum):
    a = np.zeros(35, dtype=np.int32)
    a[1] = idum
    return a


def atogen(ctx: Context, box: DeviceBox, npt1: int, ldw: int, ycoh, irefpcof, rminex, ngrsamples: int,
           ran1_state: np.ndarray, wav=None, potex=None, potexmax=None, counts=None):
    """GPU atogen.  Returns (wav, potex, potexmax, ngrsamples); ran1_state is updated in place."""
    assert ran1_state.dtype == np.int32 and ran1_state.size == 35 and ran1_state.flags["C_CONTIGUOUS"]
    if wav is None:
        wav = ctx.empty(2 * box.npairs * ldw)
    if potex is None:
        potex = ctx.empty(box.npairs)
    if potexmax is None:
        potexmax = ctx.empty(box.npairs)
    ngr = C.c_int(int(ngrsamples))
    c = box.coords()
    ctx.check(ctx.lib.epsr_atogen(ctx.h, box.h, C.byref(c), int(npt1), int(ldw), _ptr(ycoh), _ptr(irefpcof),
                                  _ptr(rminex), C.byref(ngr),
                                  ran1_state.ctypes.data_as(C.POINTER(C.c_int32)), _ptr(wav), _ptr(potex),
                                  _ptr(potexmax), _ptr(counts)))
    return wav, potex, potexmax, ngr.value


def gensq(ctx: Context, nq, npairs, npt1, ldw, wav, ycoh, foqols, ldf, sofq, ldq, natomsphere=None, mode=0):
    ctx.check(ctx.lib.epsr_gensq(ctx.h, int(nq), int(npairs), int(npt1), int(ldw), _ptr(wav), _ptr(ycoh),
                                 _ptr(foqols), int(ldf), _ptr(natomsphere), _ptr(sofq), int(ldq), int(mode)))
    return sofq


def stog(ctx: Context, rho, q, sofq, ldq, r, gofr, ldg, nqmin, nqmax, nr, ncols, xwt, mode=0):
    ctx.check(ctx.lib.epsr_stog(ctx.h, F32(rho), _ptr(q), _ptr(sofq), int(ldq), _ptr(r), _ptr(gofr), int(ldg),
                                _ptr(nqmin), _ptr(nqmax), int(nr), int(ncols), _ptr(xwt), int(mode)))
    return gofr


def foqols(ctx: Context, grid, rho, rmaxgr, fwhm=0.0, fwhmq=0.0, lorchqwidth=0.0, lorchqwidthq=0.0):
    """GPU foqols for a tables.Grid; returns (tensor [(nq+1)*ldf], ldf)."""
    ldf = grid.npt1 + 1
    out = ctx.zeros((grid.nq + 1) * ldf)
    ctx.check(ctx.lib.epsr_foqols(ctx.h, int(grid.nq), _ptr(ctx.tensor(grid.q)), int(grid.npt1),
                                  _ptr(ctx.tensor(grid.x)), _ptr(ctx.tensor(grid.r2)), _ptr(ctx.tensor(grid.ycoh)),
                                  F32(rho), F32(fwhm), F32(fwhmq), F32(lorchqwidth), F32(lorchqwidthq), F32(rmaxgr),
                                  _ptr(out), int(ldf)))
    return out, ldf
