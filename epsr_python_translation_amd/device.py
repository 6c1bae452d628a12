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


HIP_STREAM_PER_THREAD = 2   # hipStreamPerThread


class EpsrError(RuntimeError):
    pass


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _addr(t):
    return t.data_ptr() if t is not None else None


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
        # torch names its default stream 0, and its work on it is ordered with HIP's
        # per-thread default stream (handle 2), not with the legacy null stream that
        # handle 0 means to the library
        self.stream = stream
        handle = stream.cuda_stream if stream.cuda_stream != 0 else HIP_STREAM_PER_THREAD
        self.check(self.lib.epsr_ctx_set_stream(self.h, C.c_void_p(handle)))

    def check(self, rc):
        if rc != 0:
            msg = self.lib.epsr_last_error(self.h)
            raise EpsrError(f"libepsr_mi355x error {rc}: {msg.decode() if msg else ''}")

    def kernel_ms(self, family: str) -> float:
        return float(self.lib.epsr_ctx_last_kernel_ms(self.h, family.encode()))

    def kernel_ms_sum(self, family: str):
        """(summed ms, launches) of the family since the previous call; (-1.0, 0) for none."""
        n = C.c_longlong(0)
        t = float(self.lib.epsr_ctx_kernel_ms_sum(self.h, family.encode(), C.byref(n)))
        return t, int(n.value)

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

    def set_intra(self, topo):
        """upload the intramolecular topology (intra.IntraTopology) the intramolecular
        energies and moves read"""
        d, keep = topo.desc()
        self.ctx.check(self.ctx.lib.epsr_box_set_intra(self.h, C.byref(d)))
        self.intra = topo

    def set_launch(self, knob, value):
        """launch shape of this box's kernels (_native.LAUNCH_*); results do not depend on it"""
        self.ctx.check(self.ctx.lib.epsr_box_set_launch(self.h, int(knob), int(value)))

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


def atogen_bin_check(ctx: Context, stepcl, npt1, dmax, rabs, lo_bits, hi_bits):
    """(disagreements, values checked, values near a bin edge) of atogen's pair-kernel binning
    and close-contact test against the correctly rounded computation, over every float32
    squared distance with bit pattern in [lo_bits, hi_bits)."""
    out = (C.c_ulonglong * 3)()
    ctx.check(ctx.lib.epsr_atogen_bin_check(ctx.h, F32(stepcl), int(npt1), F32(dmax), F32(rabs), int(lo_bits),
                                            int(hi_bits), out))
    return int(out[0]), int(out[1]), int(out[2])


def gensq(ctx: Context, nq, npairs, npt1, ldw, wav, ycoh, foqols, ldf, sofq, ldq, natomsphere=None, mode=0):
    ctx.check(ctx.lib.epsr_gensq(ctx.h, int(nq), int(npairs), int(npt1), int(ldw), _ptr(wav), _ptr(ycoh),
                                 _ptr(foqols), int(ldf), _ptr(natomsphere), _ptr(sofq), int(ldq), int(mode)))
    return sofq


def stog(ctx: Context, rho, q, sofq, ldq, r, gofr, ldg, nqmin, nqmax, nr, ncols, xwt, mode=0):
    ctx.check(ctx.lib.epsr_stog(ctx.h, F32(rho), _ptr(q), _ptr(sofq), int(ldq), _ptr(r), _ptr(gofr), int(ldg),
                                _ptr(nqmin), _ptr(nqmax), int(nr), int(ncols), _ptr(xwt), int(mode)))
    return gofr


def gensq_batch(ctx: Context, nq, npt1, ycoh, foqols, ldf, boxes):
    """gensq mode 1 of many boxes on one grid as one GEMM (epsr_gensq_batch): boxes is a
    list of (npairs, wav, ldw, sofq, ldq[, natomsphere]); each box's sofq equals its
    gensq(..., mode=1) bit for bit."""
    arr = (_native.GensqBox * len(boxes))()
    for a, b in zip(arr, boxes):
        npairs, wav, ldw, sofq, ldq = b[:5]
        nas = b[5] if len(b) > 5 else None
        a.npairs, a.ldw, a.ldq = int(npairs), int(ldw), int(ldq)
        a.wav, a.natomsphere, a.sofq = _addr(wav), _addr(nas), _addr(sofq)
    ctx.check(ctx.lib.epsr_gensq_batch(ctx.h, int(nq), int(npt1), _ptr(ycoh), _ptr(foqols), int(ldf), len(boxes),
                                       arr))


def stog_batch(ctx: Context, rho, q, ldq, r, nr, boxes):
    """stog mode 1 of many boxes on one grid as one GEMM (epsr_stog_batch): boxes is a list
    of (sofq, gofr, ldg, nqmin, nqmax, ncols, xwt); each box's gofr equals its
    stog(..., mode=1) bit for bit."""
    arr = (_native.StogBox * len(boxes))()
    for a, (sofq, gofr, ldg, nqmin, nqmax, ncols, xwt) in zip(arr, boxes):
        a.sofq, a.gofr, a.nqmin, a.nqmax, a.xwt = (_addr(t) for t in (sofq, gofr, nqmin, nqmax, xwt))
        a.ncols, a.ldg = int(ncols), int(ldg)
    ctx.check(ctx.lib.epsr_stog_batch(ctx.h, F32(rho), _ptr(q), int(ldq), _ptr(r), int(nr), len(boxes), arr))


def foqols(ctx: Context, grid, rho, rmaxgr, fwhm=0.0, fwhmq=0.0, lorchqwidth=0.0, lorchqwidthq=0.0):
    """GPU foqols for a tables.Grid; returns (tensor [(nq+1)*ldf], ldf)."""
    ldf = grid.npt1 + 1
    out = ctx.zeros((grid.nq + 1) * ldf)
    # the inputs stay referenced until the call returns (the caching allocator would
    # otherwise hand a freed temporary to the next one)
    q, x, r2, ycoh = (ctx.tensor(a) for a in (grid.q, grid.x, grid.r2, grid.ycoh))
    ctx.check(ctx.lib.epsr_foqols(ctx.h, int(grid.nq), _ptr(q), int(grid.npt1), _ptr(x), _ptr(r2), _ptr(ycoh),
                                  F32(rho), F32(fwhm), F32(fwhmq), F32(lorchqwidth), F32(lorchqwidthq), F32(rmaxgr),
                                  _ptr(out), int(ldf)))
    return out, ldf


class DeviceTables:
    """Potential tables resident in HBM (float32 [npairs*ld] each, 1-based bins)."""

    def __init__(self, ctx: Context, ld, npt2, xstar, xavs, xavs1, ecoul, ecoul1):
        self.ctx = ctx
        self.ld = int(ld)
        self.npt2 = int(npt2)
        self.xstar = xstar if torch.is_tensor(xstar) else ctx.tensor(np.asarray(xstar, F32))
        self.xavs = xavs if torch.is_tensor(xavs) else ctx.tensor(np.asarray(xavs, F32))
        self.xavs1 = xavs1 if torch.is_tensor(xavs1) else ctx.tensor(np.asarray(xavs1, F32))
        self.ecoul = ecoul if torch.is_tensor(ecoul) else ctx.tensor(np.asarray(ecoul, F32))
        self.ecoul1 = ecoul1 if torch.is_tensor(ecoul1) else ctx.tensor(np.asarray(ecoul1, F32))

    def struct(self):
        return _native.Tables(self.ld, self.npt2, _ptr(self.xstar), _ptr(self.xavs), _ptr(self.xavs1),
                              _ptr(self.ecoul), _ptr(self.ecoul1))


class Ledger:
    """Cell lists + pair-energy ledger resident on the GPU (epsr_ledger)."""

    def __init__(self, ctx: Context, box: DeviceBox):
        self.ctx = ctx
        self.box = box
        h = C.c_void_p()
        ctx.check(ctx.lib.epsr_ledger_create(ctx.h, box.h, C.byref(h)))
        self.h = h
        self.nnbrefmax = 0
        self.nergpair = 0
        self.built = False   # a ledger from fmolec1 on the current cells

    def cells_init(self, nmolcellin, rmax):
        c = self.box.coords()
        self.ctx.check(self.ctx.lib.epsr_cells_init(self.ctx.h, self.box.h, self.h, C.byref(c), int(nmolcellin),
                                                    F32(rmax)))
        self.built = False

    def fmolec1(self, tables: DeviceTables, eintra):
        """Returns (ftest, ftestb) per molecule in FP64."""
        c = self.box.coords()
        t = tables.struct()
        ei = eintra if torch.is_tensor(eintra) else self.ctx.tensor(np.asarray(eintra, F32))
        f64 = C.c_double()
        fb64 = C.c_double()
        nn = C.c_int()
        ne = C.c_int()
        self.ctx.check(self.ctx.lib.epsr_fmolec1(self.ctx.h, self.box.h, C.byref(c), C.byref(t), self.h, _ptr(ei),
                                                 C.byref(f64), C.byref(fb64), None, None, C.byref(nn), C.byref(ne)))
        self.nnbrefmax = nn.value
        self.nergpair = ne.value
        self.built = True
        return f64.value, fb64.value

    def export(self):
        nmol = self.box.nmol
        ld = max(1, self.nnbrefmax)
        cnt = np.zeros(nmol, np.int32)
        mol = np.zeros(nmol * ld, np.int32)
        ref = np.zeros(nmol * ld, np.int32)
        erg = np.zeros(max(1, self.nergpair), F32)
        order = np.zeros(nmol, np.int32)
        P = lambda a, t=C.c_int: a.ctypes.data_as(C.POINTER(t))
        self.ctx.check(self.ctx.lib.epsr_ledger_export(self.ctx.h, self.h, ld, P(cnt), P(mol), P(ref),
                                                       P(erg, C.c_float), P(order)))
        return dict(nergpairm=cnt, molergpair=mol.reshape(nmol, ld), ergpairref=ref.reshape(nmol, ld),
                    ergpair=erg[:self.nergpair], order=order)

    def en_change(self, tables: DeviceTables, im, ergnew=None):
        c = self.box.coords()
        t = tables.struct()
        fd = C.c_float()
        self.ctx.check(self.ctx.lib.epsr_en_change(self.ctx.h, self.box.h, C.byref(c), C.byref(t), self.h, int(im),
                                                   _ptr(ergnew), C.byref(fd)))
        return F32(fd.value)

    def press(self, tables: DeviceTables, npt1: int, simt, volume, slots: bool = False):
        """press (ato_manipulation_rev4.py:3901-4063) at the box's current coordinates over
        this ledger's partner lists: (energy per molecule, pressure in kbar, virial sum) and,
        with slots=True, each ledger slot's energy and virial as [nmol, nnbrefmax] arrays in
        the reference's layout."""
        c = self.box.coords()
        t = tables.struct()
        e, p, w = C.c_float(), C.c_float(), C.c_double()
        ldl = max(1, self.nnbrefmax)
        se = sw = None
        if slots:
            se = np.zeros((self.box.nmol, ldl), F32)
            sw = np.zeros((self.box.nmol, ldl), F32)
        fp_ = C.POINTER(C.c_float)
        self.ctx.check(self.ctx.lib.epsr_press(
            self.ctx.h, self.box.h, C.byref(c), C.byref(t), self.h, int(npt1), F32(simt), F32(volume), C.byref(e),
            C.byref(p), C.byref(w), ldl, se.ctypes.data_as(fp_) if slots else None,
            sw.ctypes.data_as(fp_) if slots else None))
        out = (F32(e.value), F32(p.value), float(w.value))
        return out + (se, sw) if slots else out

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.epsr_ledger_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sumsq(ctx: Context, npairs: int, nq: int, ldq: int, qwind, fc1: float, fc2: float, soq, sq2):
    """sumsq (epsr_standalone_rev3.py:3642-3685): soq/sq2 [2*npairs*ldq] updated in place."""
    ctx.check(ctx.lib.epsr_sumsq(ctx.h, int(npairs), int(nq), int(ldq), _ptr(qwind), float(fc1), float(fc2),
                                 _ptr(soq), _ptr(sq2)))


def sumgr(ctx: Context, npairs: int, nz: int, ldw: int, wav, ycoh, fc1: float, fc2: float, gor, gr2, entsig):
    """sumgr (epsr_standalone_rev3.py:3578-3637): gor/gr2 [2*npairs*ldw] updated in place,
    entsig [npairs*ldw] the zero limits."""
    ctx.check(ctx.lib.epsr_sumgr(ctx.h, int(npairs), int(nz), int(ldw), _ptr(wav), _ptr(ycoh), float(fc1),
                                 float(fc2), _ptr(gor), _ptr(gr2), _ptr(entsig)))


class Graph:
    """Work issued on a context's stream inside ``with Graph.capture(ctx) as g:`` is
    recorded into a HIP graph; ``g.launch()`` replays it on the same stream."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.h = None

    @classmethod
    def capture(cls, ctx: Context):
        return _Capture(cls(ctx))

    def launch(self):
        self.ctx.check(self.ctx.lib.epsr_graph_launch(self.ctx.h, self.h))

    def close(self):
        if self.h:
            self.ctx.lib.epsr_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _Capture:
    def __init__(self, g: Graph):
        self.g = g

    def __enter__(self):
        self.g.ctx.check(self.g.ctx.lib.epsr_graph_begin(self.g.ctx.h))
        return self.g

    def __exit__(self, et, ev, tb):
        h = C.c_void_p()
        rc = self.g.ctx.lib.epsr_graph_end(self.g.ctx.h, C.byref(h))
        if et is None:
            self.g.ctx.check(rc)
            self.g.h = h
        return False


def en_mole_pair(ctx: Context, box: DeviceBox, tables: DeviceTables, im, jm):
    im_t = ctx.tensor(np.asarray(im, np.int32))
    jm_t = ctx.tensor(np.asarray(jm, np.int32))
    out = ctx.empty(im_t.numel())
    c = box.coords()
    t = tables.struct()
    ctx.check(ctx.lib.epsr_en_mole_pair(ctx.h, box.h, C.byref(c), C.byref(t), int(im_t.numel()), _ptr(im_t),
                                        _ptr(jm_t), _ptr(out)))
    return out


class MCState:
    """Step sizes and counters of the MC cycle (clustercom.wholerot/wholetrans,
    nwhole/nmovem, nstepreject) plus the Metropolis parameters."""

    def __init__(self, ctx: Context, nmol, nmoltype, simt, stepr=1.0, stepm=1.0, rejrate=0.75, ireset=0,
                 wholerot=None, wholetrans=None):
        self.params = _native.MCParams(F32(simt), F32(stepr), F32(stepm), F32(rejrate), int(ireset))
        # copies: the cycles adapt the step sizes in place, and the caller's arrays stay as given
        self.wholerot = np.array(np.zeros(nmoltype, F32) if wholerot is None else wholerot, dtype=F32, copy=True)
        self.wholetrans = np.array(np.zeros(nmoltype, F32) if wholetrans is None else wholetrans, dtype=F32,
                                   copy=True)
        self.nwhole = np.zeros(2 * nmoltype, np.int32)
        self.nmovem = np.zeros(2 * nmoltype, np.int32)
        self.nstepreject = ctx.zeros(nmol, torch.int32)
        self.fadds = F32(0.0)
        self.stats = _native.MCStats()


def _mc_args(mc: MCState, ran1_state):
    """ctypes pointers to the host arrays of mc and ran1_state, built once per set of arrays
    (building them costs several microseconds per array and call)"""
    objs = (mc.wholerot, mc.wholetrans, mc.nwhole, mc.nmovem, mc.nstepreject, ran1_state)
    cached = getattr(mc, "_args", None)
    if cached is None or any(a is not b for a, b in zip(cached[0], objs)):
        P = lambda a, ty=C.c_int: a.ctypes.data_as(C.POINTER(ty))
        args = (P(mc.wholerot, C.c_float), P(mc.wholetrans, C.c_float), P(mc.nwhole), P(mc.nmovem),
                _ptr(mc.nstepreject), P(ran1_state, C.c_int32))
        cached = (objs, args)   # holds the arrays, so their ids and buffers stay valid
        mc._args = cached
    return cached[1]


def molfit(ctx: Context, box: DeviceBox, tables: DeviceTables, ledger: Ledger, mc: MCState, ran1_state, ncycles=1):
    """ncycles MC cycles on the GPU (one call of epsr_molfit, or epsr_molfit_cycles for
    several back to back); updates coordinates, ledger, mc and ran1_state in place."""
    c = box.coords()
    t = tables.struct()
    fad = C.c_float(float(mc.fadds))
    if ncycles == 1:
        ctx.check(ctx.lib.epsr_molfit(ctx.h, box.h, C.byref(c), C.byref(t), ledger.h, C.byref(mc.params),
                                      *_mc_args(mc, ran1_state), C.byref(fad), C.byref(mc.stats)))
    else:
        ctx.check(ctx.lib.epsr_molfit_cycles(ctx.h, box.h, C.byref(c), C.byref(t), ledger.h, C.byref(mc.params),
                                             int(ncycles), *_mc_args(mc, ran1_state), C.byref(fad),
                                             C.byref(mc.stats)))
    mc.fadds = F32(fad.value)
    return mc


def en_mole_intra(ctx: Context, box: DeviceBox, tables: DeviceTables, simt, tempang, tempdih, xav=None, xav1=None,
                  out=None):
    """en_mole_intra of every molecule at the current coordinates: float32 tensor [nmol]."""
    out = ctx.zeros(box.nmol) if out is None else out
    c = box.coords()
    t = tables.struct()
    ctx.check(ctx.lib.epsr_en_mole_intra(ctx.h, box.h, C.byref(c), C.byref(t), _ptr(xav) if xav is not None else None,
                                         _ptr(xav1) if xav1 is not None else None, F32(simt), F32(tempang),
                                         F32(tempdih), _ptr(out)))
    return out


class MCFullState:
    """Everything molfit keeps between cycles (clustercom step sizes and counters,
    nstepreject, fadds/faddbs, nisum) plus the run parameters (molecom.stepr/stepm/stepri/
    stepmi, the intramolecular weights, rejrate, ireset, intra, rotfreq)."""

    def __init__(self, ctx: Context, nmol, nmoltype, simt, *, stepr=1.0, stepm=1.0, stepri=0.0, stepmi=0.0,
                 rejrate=0.75, tempangwt=1.0, tempdihwt=0.0, ireset=0, intra=100, rotfreq=5, nisum=None,
                 wholerot=None, wholetrans=None, internrot=None, interntrans=None, nstepreject=None):
        self.params = _native.MCFullParams(F32(simt), F32(stepr), F32(stepm), F32(stepri), F32(stepmi), F32(rejrate),
                                           F32(tempangwt), F32(tempdihwt), int(ireset), int(intra), int(rotfreq))
        z = lambda a: np.ascontiguousarray(np.zeros(nmoltype, F32) if a is None else a, dtype=F32).copy()
        self.wholerot, self.wholetrans = z(wholerot), z(wholetrans)
        self.internrot, self.interntrans = z(internrot), z(interntrans)
        self.nwhole, self.nmovem = np.zeros(2 * nmoltype, np.int32), np.zeros(2 * nmoltype, np.int32)
        self.ninter, self.nmovei = np.zeros(2 * nmoltype, np.int32), np.zeros(2 * nmoltype, np.int32)
        self.nstepreject = ctx.zeros(nmol, torch.int32)
        if nstepreject is not None:
            self.nstepreject.copy_(torch.from_numpy(np.ascontiguousarray(nstepreject, dtype=np.int32)))
        self.fadds = F32(0.0)
        self.faddbs = F32(0.0)
        self.nisum = int(intra if nisum is None else nisum)
        self.stats = _native.MCStats()

    def steps(self):
        P = lambda a, ty=C.c_int: a.ctypes.data_as(C.POINTER(ty))
        return _native.MCSteps(P(self.wholerot, C.c_float), P(self.wholetrans, C.c_float),
                               P(self.internrot, C.c_float), P(self.interntrans, C.c_float), P(self.nwhole),
                               P(self.nmovem), P(self.ninter), P(self.nmovei))


def molfit_full(ctx: Context, box: DeviceBox, tables: DeviceTables, ledger: Ledger, mc: MCFullState, ran1_state,
                xav=None, xav1=None):
    """One MC cycle of any kind (molfit.molfit): updates coordinates, ledger, mc and
    ran1_state in place."""
    c = box.coords()
    t = tables.struct()
    st = mc.steps()
    fad, fadb, nis = C.c_float(float(mc.fadds)), C.c_float(float(mc.faddbs)), C.c_int(int(mc.nisum))
    assert ran1_state.dtype == np.int32 and ran1_state.size == 35
    ctx.check(ctx.lib.epsr_molfit_full(ctx.h, box.h, C.byref(c), C.byref(t), _ptr(xav) if xav is not None else None,
                                       _ptr(xav1) if xav1 is not None else None, ledger.h, C.byref(mc.params),
                                       C.byref(st), _ptr(mc.nstepreject), ran1_state.ctypes.data_as(C.POINTER(C.c_int32)),
                                       C.byref(fad), C.byref(fadb), C.byref(nis), C.byref(mc.stats)))
    mc.fadds, mc.faddbs, mc.nisum = F32(fad.value), F32(fadb.value), int(nis.value)
    return mc


def molfit_da(ctx: Context, box: DeviceBox, tables: DeviceTables, ledger: Ledger, mc: MCState, ran1_state,
              da=None, record=None):
    """One MC cycle with delayed acceptance (da: a _native.DA, or None) and/or recording
    (record: a _native.MCRecord, or None); with both None it is molfit."""
    c = box.coords()
    t = tables.struct()
    fad = C.c_float(float(mc.fadds))
    ctx.check(ctx.lib.epsr_molfit_da(ctx.h, box.h, C.byref(c), C.byref(t), ledger.h, C.byref(mc.params),
                                     *_mc_args(mc, ran1_state), C.byref(fad),
                                     C.byref(da) if da is not None else None,
                                     C.byref(record) if record is not None else None, C.byref(mc.stats)))
    mc.fadds = F32(fad.value)
    return mc


def descriptors(ctx: Context, box: DeviceBox, ledger: Ledger, trial: dict, dp, ldx, with_f32=False):
    """Descriptor rows 2k (current) and 2k+1 (trial) of every chain position; trial maps
    xa..zadd to device tensors of the trial coordinates.  Returns (x bf16 bits as int16
    [2 nmol, ldx], x fp32 or None)."""
    n = 2 * box.nmol
    x = ctx.zeros(n * ldx, torch.int16)
    xf = ctx.zeros(n * ldx) if with_f32 else None
    c = box.coords()
    tc = _native.Coords(*(_ptr(trial[k]) for k in ("xa", "ya", "za", "xadd", "yadd", "zadd")))
    ctx.check(ctx.lib.epsr_descriptors(ctx.h, box.h, C.byref(c), C.byref(tc), ledger.h, C.byref(dp), int(ldx),
                                       _ptr(x), _ptr(xf)))
    return x.view(n, ldx), (xf.view(n, ldx) if xf is not None else None)


# ---------------------------------------------------------------- EP fit

SUMMAX = F32(np.exp(25.0))   # expf(25) (epsr_standalone_rev3.py:4722)


def expand_all(ctx: Context, nterms, npt, expr, coeff, summax=SUMMAX, out=None):
    """expand(nterms, j, expr, coeff, summax) for j = 1..npt (dev tensors, 1-based)."""
    if out is None:
        out = ctx.zeros(npt + 1)
    ctx.check(ctx.lib.epsr_expand_all(ctx.h, int(nterms), int(npt), _ptr(expr), _ptr(coeff), F32(summax),
                                      _ptr(out)))
    return out


def smoop(ctx: Context, nqmin, nqmax, nsmoo, diff):
    """auxiliary_routines.smoop on a device tensor, in place."""
    ctx.check(ctx.lib.epsr_smoop(ctx.h, int(nqmin), int(nqmax), int(nsmoo), _ptr(diff)))
    return diff


class FitResult:
    def __init__(self, stepr, acceptr, rejrat, chisqnew, ntries):
        self.stepr, self.acceptr, self.rejrat, self.chisqnew, self.ntries = stepr, acceptr, rejrat, chisqnew, ntries


def fit_expansion(ctx: Context, yin, expr, yout, wt, wtold, stepr, acceptr, rejrat, chisqnew, nmin, npt, ncfirst,
                  nterms, nits, nfreq, ran1_state, mode=_native.FIT_MC, nchain=1, ldy=None, lde=0, ldo=None,
                  ldw=None):
    """fit_expansion for nchain chains (device tensors yin/expr/yout/wt/wtold, chain c at
    c*ld).  stepr/acceptr/rejrat/chisqnew: scalars or arrays of nchain; ran1_state int32
    [nchain*35], updated in place.  Returns FitResult with float32 arrays."""
    ldy = ldy or npt + 1
    ldo = ldo or npt + 1
    ldw = ldw or nterms + 1
    st = np.ascontiguousarray(np.broadcast_to(np.asarray(stepr, F32), (nchain,)), dtype=F32).copy()
    ac = np.ascontiguousarray(np.broadcast_to(np.asarray(acceptr, F32), (nchain,)), dtype=F32).copy()
    rj = np.ascontiguousarray(np.broadcast_to(np.asarray(rejrat, F32), (nchain,)), dtype=F32).copy()
    cs = np.ascontiguousarray(np.broadcast_to(np.asarray(chisqnew, F32), (nchain,)), dtype=F32).copy()
    nt = np.zeros(nchain, np.int32)
    if ran1_state is None:
        r1p = None
    else:
        assert ran1_state.dtype == np.int32 and ran1_state.size == 35 * nchain and ran1_state.flags["C_CONTIGUOUS"]
        r1p = ran1_state.ctypes.data_as(C.POINTER(C.c_int32))
    fp = C.POINTER(C.c_float)
    ctx.check(ctx.lib.epsr_fit_expansion(
        ctx.h, int(nchain), _ptr(yin), int(ldy), _ptr(expr), int(lde), _ptr(yout), int(ldo), _ptr(wt), _ptr(wtold),
        int(ldw), st.ctypes.data_as(fp), ac.ctypes.data_as(fp), rj.ctypes.data_as(fp), cs.ctypes.data_as(fp),
        int(nmin), int(npt), int(ncfirst), int(nterms), int(nits), int(nfreq), r1p, int(mode),
        nt.ctypes.data_as(C.POINTER(C.c_int))))
    return FitResult(st, ac, rj, cs, nt)


def calc_pwrexpq(ctx: Context, lpt, nterms, dmin, dstep, dmax, sigmab, rho, q, expr=None, rmin=0.0):
    """The Q-space Poisson basis; returns (expr tensor, ncfirst)."""
    if expr is None:
        expr = ctx.zeros(lpt * nterms + 1)
    nc = C.c_int()
    ctx.check(ctx.lib.epsr_calc_pwrexpq(ctx.h, int(lpt), int(nterms), F32(dmin), F32(dstep), F32(dmax), F32(sigmab),
                                        F32(rho), _ptr(q), _ptr(expr), F32(rmin), C.byref(nc)))
    return expr, nc.value


def ptgnsq_diff(ctx: Context, nfile, nq, npairs, q, data, ldq, qwind, cwt, szeros, nqmin, nqmax, qminptemp,
                qtemp=None, weightx=None, xfile=None):
    """ptgnsq's differences; weightx (device float64 [nxfile*ldq*npairs]) and xfile (host
    int [nfile], 0 for neutron data) for X-ray datasets."""
    if qtemp is None:
        qtemp = ctx.zeros(nfile * ldq)
    cwt_, cwtp = _fp(cwt)
    sz_, szp = _fp(szeros)
    mn_, mnp = _ip(nqmin)
    mx_, mxp = _ip(nqmax)
    if xfile is not None:
        xf_, xfp = _ip(xfile)
        nx = int(np.max(xf_)) if len(xf_) else 0
    else:
        xfp, nx = None, 0
    ctx.check(ctx.lib.epsr_ptgnsq_diff(ctx.h, int(nfile), int(nq), int(npairs), _ptr(q), _ptr(data), int(ldq),
                                       _ptr(qwind), cwtp, szp, mnp, mxp, F32(qminptemp), _ptr(qtemp), _ptr(weightx),
                                       xfp, nx))
    return qtemp


def data_minus_sim(ctx: Context, q, datai, ldq, nq, soqintra, soqinter, ciwt, cwt, nqmind, nqmaxd, nqmin, nqmax,
                   szeros, porodamp, rshmin, porodq, porodpwr, nsmoop, x, npt1, rho, grt, weightx=None, xfile=None,
                   datat=None, szerostemp=None):
    """The data-minus-simulation block of the epsr program (epsr_data_minus_sim).  Device:
    q [ldq], datai [nfile*ldq], soqintra/soqinter [npairs*ldq], x [npt1+1], grt [npt1+1]
    (kept between calls), weightx float64 [nxfile*ldq*npairs]; host: ciwt/cwt [nfile,
    npairs], the per-dataset ranges and scalars.  Returns (datat, szerostemp) on the device."""
    ciwt = np.ascontiguousarray(ciwt, F32)
    cwt = np.ascontiguousarray(cwt, F32)
    nfile, npairs = ciwt.shape
    xf = np.zeros(nfile, np.int32) if xfile is None else np.ascontiguousarray(xfile, np.int32)
    keep = dict(ciwt=ciwt, cwt=cwt, xfile=xf, nqmind=np.ascontiguousarray(nqmind, np.int32),
                nqmaxd=np.ascontiguousarray(nqmaxd, np.int32), nqmin=np.ascontiguousarray(nqmin, np.int32),
                nqmax=np.ascontiguousarray(nqmax, np.int32), szeros=np.ascontiguousarray(szeros, F32),
                porodamp=np.ascontiguousarray(porodamp, F32), rshmin=np.ascontiguousarray(rshmin, F32))
    if datat is None:
        datat = ctx.zeros(nfile * ldq)
    if szerostemp is None:
        szerostemp = ctx.zeros(nfile)
    dts = ctx.zeros(ldq)
    h = lambda k: keep[k].ctypes.data
    dv = lambda t: t.data_ptr() if t is not None else None
    d = _native.DMSDesc(nfile, int(nq), int(ldq), npairs, int(xf.max()) if nfile else 0, dv(q), dv(datai),
                        dv(soqintra), dv(soqinter), dv(weightx), dv(x), h("ciwt"), h("cwt"), h("xfile"),
                        h("nqmind"), h("nqmaxd"), h("nqmin"), h("nqmax"), h("szeros"), h("porodamp"), h("rshmin"),
                        F32(porodq), F32(porodpwr), int(nsmoop), int(npt1), F32(rho))
    ctx.check(ctx.lib.epsr_data_minus_sim(ctx.h, C.byref(d), _ptr(datat), _ptr(szerostemp), _ptr(dts), _ptr(grt)))
    return datat, szerostemp


def ptgnsq_rfactor(ctx: Context, i0, i1, fit, qtemp):
    r = C.c_float()
    ctx.check(ctx.lib.epsr_ptgnsq_rfactor(ctx.h, int(i0), int(i1), _ptr(fit), _ptr(qtemp), C.byref(r)))
    return F32(r.value)


def ptgnsq_update(ctx: Context, id_, nterms, npairs, pwtold, irefpcof, efacp, cwtpot, ldc, pfact, pcoefd, pcoefp,
                  ldp):
    """pcoefd/pcoefp: host float32 arrays updated in place."""
    pw_, pwp = _fp(pwtold)
    ir_, irp = _ip(irefpcof)
    ef_, efp = _fp(efacp)
    cw_, cwp = _fp(cwtpot)
    assert pcoefd.dtype == F32 and pcoefp.dtype == F32 and pcoefd.flags["C_CONTIGUOUS"] and pcoefp.flags["C_CONTIGUOUS"]
    fp = C.POINTER(C.c_float)
    ctx.check(ctx.lib.epsr_ptgnsq_update(ctx.h, int(id_), int(nterms), int(npairs), pwp, irp, efp, cwp, int(ldc),
                                         F32(pfact), pcoefd.ctypes.data_as(fp), pcoefp.ctypes.data_as(fp), int(ldp)))
