"""One box's refinement-loop state resident on the GPU, and the steps of the loop.

A :class:`Refinement` holds everything the hot path reads and writes, in HBM:
coordinates, radial/Q grids, transform matrix, truncation functions, reference and
empirical potential tables, cell lists + pair-energy ledger, MC step sizes and counters,
histograms and partial structure factors.  Its methods are the steps of EPSR's
iteration (src/epsr26/programs/epsr_standalone_rev3.py:1235-1400): fmolec1, molfit
sweeps, atogen, gensq, stog, generate_rp/generate_ep.  Only the ran1 state and a few
scalars live on the host.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _native
from . import device as D
from . import tables as T

F32 = np.float32


def sum_factors(nsum: int, nsums: int = 0, nesum: int = 0):
    """The accumulation factors of common /sumvalues/ that runepsr17 sets before sumsq and
    sumgr (epsr_standalone_rev3.py:1613-1618): (fc1, fc2) as doubles."""
    if nsums >= 0:
        fc1 = float(F32(1.0) / F32(nsum))
        fc2 = float(F32(nsum - 1)) * fc1
    else:
        fc1 = float(F32(1.0) / F32(nesum + 1))
        fc2 = float(F32(nesum)) * fc1
    return fc1, fc2


def invert_cwt(ctx, cwt, cwtt, feedback, efacm, ran1_state):
    """invert_cwt (epsr_standalone_rev3.py:5423-5672) through the library: cwt [nfile,
    npairs] and cwtt [npairs, npairs + nfile] are float64 arrays updated in place (cwt
    scaled by feedback * efacm, cwtt the transpose); ran1_state int32[35] in place.
    Returns (carray, carrayold, itimes)."""
    cwt = np.asarray(cwt)
    cwtt = np.asarray(cwtt)
    assert cwt.dtype == np.float64 and cwtt.dtype == np.float64 and cwt.flags["C_CONTIGUOUS"] and \
        cwtt.flags["C_CONTIGUOUS"]
    assert ran1_state.dtype == np.int32 and ran1_state.size == 35
    nfile, npairs = cwt.shape
    nnew = npairs + nfile
    assert cwtt.shape == (npairs, nnew)
    ef = np.ascontiguousarray(efacm, dtype=F32)
    carray = np.zeros(nnew)
    carrayold = np.zeros((nnew, nnew))
    it = C.c_int()
    ctx.check(ctx.lib.epsr_invert_cwt(ctx.h, cwt.ctypes.data, npairs, cwtt.ctypes.data, nnew, carray.ctypes.data,
                                      carrayold.ctypes.data, nnew, F32(feedback), ef.ctypes.data, nfile, npairs,
                                      ran1_state.ctypes.data, C.byref(it)))
    return carray, carrayold, it.value


_LIBM = None


def _expf(x):
    """glibc expf, the function the reference's fmath.expf calls (fortran/fmath.py:46-145)"""
    global _LIBM
    if _LIBM is None:
        import ctypes.util
        _LIBM = C.CDLL(ctypes.util.find_library("m") or "libm.so.6")
        _LIBM.expf.restype = C.c_float
        _LIBM.expf.argtypes = [C.c_float]
    return F32(_LIBM.expf(C.c_float(x)))


def _ifix(x):
    return int(np.trunc(F32(x)))


def _qtypea(s):
    """molecom.qtypea: per atom type, the Q-type flag of its first atom (1 for labels
    starting with Q, ato_manipulation_rev4.py:269)"""
    out = np.zeros(s.ntype, np.int32)
    if getattr(s, "intra", None) is not None:
        seen = set()
        for ia, it in enumerate(np.asarray(s.itype)):
            if int(it) not in seen:
                seen.add(int(it))
                out[int(it)] = int(s.intra.qtype[ia])
    return out


def xray_form_factors(q, nq, w, labels, ntype):
    """wxtype (epsr_standalone_rev3.py:4571-4610) for every record of X-ray weights file w:
    f0x, fm0x float64 [nq + 1, ntype] (Q index 1..nq; types without a record stay 0)."""
    f0x = np.zeros((nq + 1, ntype))
    fm0x = np.zeros((nq + 1, ntype))
    pi4 = 16.0 * float(np.arctan(1.0))
    import math
    for r in w.records:
        n1 = max([k for k, lab in enumerate(labels) if lab == r.a1], default=-1)
        if n1 < 0:
            continue
        for iq in range(1, nq + 1):
            qx = float(q[iq]) / pi4
            qx2 = qx * qx
            sm = r.cxt
            for a, b in zip(r.axt, r.bxt):
                term = -b * qx2
                if -40.0 < term < 40.0:
                    sm = sm + a * math.exp(term)
            f0x[iq, n1] = sm
            fac = F32(1.0)
            if r.alphaxt != 0.0 and r.deltaxt > 0.0:
                expon = F32(F32(q[iq] * q[iq]) / F32(F32(F32(2.0) * r.deltaxt) * r.deltaxt))
                if expon < 20.0:
                    fac = F32(1.0 - (float(r.alphaxt) / f0x[1, n1]) * float(_expf(-expon)))
            fm0x[iq, n1] = float(fac) * sm
    return f0x, fm0x


def xray_weights(nq, w, f0x, fm0x, s, qtypea, cfraccg):
    """wxcreate (epsr_standalone_rev3.py:4613-4705) for one X-ray file: the Q-dependent pair
    weights weightx float64 [nq + 1, npairs] and the self-scattering shift backx [nq + 1]
    (before get_wts scales weightx by the concentrations)."""
    nrho, normx, iafformaff = _ifix(w.rhofac), int(w.normx), int(w.iafformaff)
    amolrat = F32(F32(s.natomr) / F32(s.nmol))
    weightx = np.zeros((nq + 1, s.npairs))
    backx = np.zeros(nq + 1)
    for iq in range(1, nq + 1):
        singlem = F32(0.0)
        singlei = F32(0.0)
        bbar = 0.0
        for it in range(s.ntype):
            if qtypea[it] != 0:
                continue
            cf = float(cfraccg[it])
            fm, f0 = fm0x[iq, it], f0x[iq, it]
            singlem = F32(float(singlem) + cf * (fm * fm))
            singlei = F32(float(singlei) + cf * (f0 * f0))
            f = fm if iafformaff == 2 else f0
            if normx == 0:
                bbar = bbar + cf
            elif normx == 1:
                bbar = bbar + cf * f
            elif normx == 2:
                bbar = bbar + cf * (f * f)
        if normx == 1:
            bbar = bbar * bbar
        for ip in range(s.npairs):
            n1, n2 = int(s.n1extr[ip]), int(s.n2extr[ip])
            weightx[iq, ip] = fm0x[iq, n1] * fm0x[iq, n2] / bbar
            if nrho == 2 and normx == 1:
                weightx[iq, ip] = weightx[iq, ip] / float(amolrat)
        backx[iq] = float(F32(singlem - singlei)) / bbar if iafformaff == 1 else 0.0
        if nrho == 2 and normx == 1:
            backx[iq] = backx[iq] / float(amolrat)
    return weightx, backx


def get_wts(ctx, s, wts, feedback, efacm, ran1_state, cwtt=None, recompute=True, q=None, nq=None):
    """get_wts (epsr_standalone_rev3.py:4180-4479) for box `s` and the parsed weights files
    `wts` (formats.read_wts, one per dataset, neutron or X-ray): per dataset and type pair
    the weights ciwt (intra) and cwt (inter) -- for X-ray data the Q = 0 values of the
    Q-dependent weights -- wsum = sum |cwt| and xwt = sum cwt, and, when `recompute`
    (ireset = 1 or feedback/efacm changed), the inverted matrix through invert_cwt from
    cwtt = [0 | I] (else from the given cwtt; with neither, cwtpot is None): at one Q, or
    with X-ray data at every 2/qstep-th Q point (at most 15), continuing from the previous
    point's matrix.  X-ray data need the Q grid q [nq + 1] (float32).
    Returns a dict with ciwt, cwt [nfile, npairs] (float32), wsum, xwt [nfile], cwtt
    [npairs, npairs + nfile] and cwtpot [npairs, nfile] (float64), and for X-ray data
    xfile [nfile] (0 neutron, else the 1-based X-ray file), weightx float64 [nxfile, nq + 1,
    npairs] (scaled by the concentrations), backx [nxfile, nq + 1], f0x/fm0x [nxfile,
    nq + 1, ntype], qinv (1-based Q points of the inversion) and cwttq [nqinv, npairs,
    npairs + nfile]."""
    nfile, npairs = len(wts), s.npairs
    labels = [fa3(t) for t in s.type_names]
    weight = np.zeros((2, nfile, npairs), F32)
    cfraccg = s.cfraccg if s.cfraccg is not None else s.cfrac
    xfile = np.zeros(nfile, np.int32)
    xr = []                                             # per X-ray file: (weightx, backx, f0x, fm0x)
    for i, w in enumerate(wts):
        if getattr(w, "xray", False):
            if q is None or nq is None:
                raise ValueError("get_wts: X-ray weights need the Q grid (q, nq)")
            if len(xr) >= 2:                            # x_rayf0.mxfile: further X-ray files get no weights
                continue
            f0x, fm0x = xray_form_factors(q, nq, w, labels, s.ntype)
            wx, bx = xray_weights(nq, w, f0x, fm0x, s, _qtypea(s), cfraccg)
            xr.append((wx, bx, f0x, fm0x))
            xfile[i] = len(xr)
            weight[0, i, :] = wx[1, :].astype(F32)     # the Q = 0 weights (wxcreate :4690-4693)
            weight[1, i, :] = weight[0, i, :]
            continue
        for a1, a2, w1, w2 in w.pairs:                 # wtype (:4540-4558)
            n1 = max([k + 1 for k, lab in enumerate(labels) if lab == a1], default=0)
            n2 = max([k + 1 for k, lab in enumerate(labels) if lab == a2], default=0)
            if n1 > 0 and n2 > 0:
                ic = int(s.npairref[n1 - 1, n2 - 1])
                weight[0, i, ic], weight[1, i, ic] = w1, w2
    ciwt = np.zeros((nfile, npairs), F32)
    cwt = np.zeros((nfile, npairs), F32)
    wsum = np.zeros(nfile, F32)
    xwt = np.zeros(nfile, F32)
    cprods = np.zeros(npairs, F32)
    for ic in range(npairs):
        it, jt = int(s.n1extr[ic]), int(s.n2extr[ic])
        cprod = F32(cfraccg[it]) * F32(cfraccg[jt])
        cprods[ic] = F32(2.0) * cprod if it != jt else cprod
    for i in range(nfile):
        for ic in range(npairs):
            ciwt[i, ic] = cprods[ic] * weight[0, i, ic]
            cwt[i, ic] = cprods[ic] * weight[1, i, ic]
            wsum[i] = F32(float(wsum[i]) + float(abs(cwt[i, ic])))
            xwt[i] = F32(float(xwt[i]) + float(cwt[i, ic]))
    for wx, _, _, _ in xr:
        wx *= cprods.astype(np.float64)[None, :]
    out = dict(ciwt=ciwt, cwt=cwt, wsum=wsum, xwt=xwt, xfile=xfile)
    if xr:
        out["weightx"] = np.stack([x[0] for x in xr])
        out["backx"] = np.stack([x[1] for x in xr])
        out["f0x"] = np.stack([x[2] for x in xr])
        out["fm0x"] = np.stack([x[3] for x in xr])
        nqinvstep = _ifix(F32(2.0) / F32(q[2] - q[1]))
    else:
        nqinvstep = nq if nq is not None else 1
    nnew = npairs + nfile
    if recompute:
        cwtt = np.zeros((npairs, nnew))
        for ip in range(npairs):
            cwtt[ip, nfile + ip] = 1.0
        rat = np.zeros(npairs, F32)
        for ic in range(npairs):
            it, jt = int(s.n1extr[ic]), int(s.n2extr[ic])
            if cfraccg[it] > 0.0 and cfraccg[jt] > 0.0:
                rat[ic] = F32(s.cfrac[it]) * F32(s.cfrac[jt]) / (F32(cfraccg[it]) * F32(cfraccg[jt]))
        qinv, cwttq = [], []
        for iqinv in range(1, (nq if nq is not None else 1) + 1, max(1, nqinvstep)):
            if len(qinv) >= 15:                         # weights.mqinv
                break
            qinv.append(iqinv)
            cwtpot = np.zeros((nfile, npairs))
            for i in range(nfile):
                for ic in range(npairs):
                    wv = out["weightx"][xfile[i] - 1, iqinv, ic] if xfile[i] > 0 else float(cwt[i, ic])
                    cwtpot[i, ic] = float(rat[ic]) * wv
            invert_cwt(ctx, cwtpot, cwtt, feedback, efacm, ran1_state)
            cwttq.append(cwtt.copy())
        out["qinv"] = np.array(qinv, np.int32)
        out["cwttq"] = np.stack(cwttq)
        cwtt = cwttq[0]
    out["cwtt"] = cwtt
    if cwtt is None:           # no inversion asked for and none given
        out["cwtpot"] = None
        return out
    pot = cwtt[:, :nfile].copy()
    for ic in range(npairs):
        if s.n1extr[ic] != s.n2extr[ic]:
            pot[ic] = 0.5 * pot[ic]
    out["cwtpot"] = pot
    return out


def fa3(label):
    return (label + "   ")[:3]


class Refinement:
    def __init__(self, ctx: D.Context, system, idum: int = -1, timing: bool = False):
        self.ctx = ctx
        self.s = s = system
        self.g = g = T.make_grid(s)
        self.box = D.DeviceBox.from_system(ctx, s)
        self.ran1 = D.ran1_state_array(idum)
        # grids on the device
        self.x = ctx.tensor(g.x)
        self.xstar = ctx.tensor(g.xstar)
        self.ycoh = ctx.tensor(g.ycoh)
        self.q = ctx.tensor(g.q)
        self.ld = g.ld
        self.ldw = g.npt1 + 1
        self.ldq = g.nq + 1
        npairs = s.npairs
        self.irefpcof_h = np.arange(npairs, dtype=np.int32)
        self.irefpcof = ctx.tensor(self.irefpcof_h)
        self.rminex = ctx.tensor(np.asarray(s.rminex, F32)) if s.rminex is not None else ctx.zeros(npairs)
        self.pcoefp = np.ascontiguousarray(s.pcoefp.astype(F32).copy())
        # tables
        z = lambda: ctx.zeros(npairs * self.ld)
        self.wfac = ctx.zeros(self.ld)
        self.wfacc = ctx.zeros(self.ld)
        self.epot, self.ecoul, self.epot1, self.ecoul1 = z(), z(), z(), z()
        self.xavs, self.xavs1, self.xavtot, self.xavtot1 = z(), z(), z(), z()
        self.foqols = None
        self.ldf = g.npt1 + 1
        self.wav = ctx.zeros(2 * npairs * self.ldw)
        self.potex = ctx.zeros(npairs)
        self.potexmax = ctx.zeros(npairs)
        self.sofq = ctx.zeros(2 * npairs * self.ldq)
        self.ledger = D.Ledger(ctx, self.box)
        r = s.mc_run
        if s.intra is not None:
            # a box from files: intramolecular energies from its topology, every branch of
            # molfit, the file's step sizes; fmolec1 sets the angle and dihedral weights to
            # the requested ones (ato_manipulation_rev4.py:2416-2419), which molfit then uses
            self.box.set_intra(s.intra)
            self.eintra = None
            self.mc = D.MCFullState(ctx, s.nmol, s.nmoltype, s.simt, stepr=r["stepr"], stepm=r["stepm"],
                                    stepri=r["stepri"], stepmi=r["stepmi"], rejrate=s.rejrate,
                                    tempangwt=r["tempangreq"], tempdihwt=r["tempdihreq"], intra=r["intra"],
                                    rotfreq=r["rotfreq"], nisum=r["intra"] if r["iinit"] != 0 else 1,
                                    wholerot=r["wholerot"], wholetrans=r["wholetrans"], internrot=r["internrot"],
                                    interntrans=r["interntrans"], nstepreject=r["nstepreject"])
        else:
            self.eintra = ctx.tensor(s.eintra.astype(F32))
            wr = np.full(s.nmoltype, 0.05, F32)
            wt = np.full(s.nmoltype, F32(0.1) / F32(s.dmax), F32)
            self.mc = D.MCState(ctx, s.nmol, s.nmoltype, s.simt, rejrate=s.rejrate, wholerot=wr, wholetrans=wt)
        self.pressfac = F32(1.0)

    # ---------------------------------------------------------------- setup (inita/initpt)
    def build_tables(self):
        """truncation functions, foqols, generate_rp and generate_ep (initpt + inita)."""
        ctx, s, g = self.ctx, self.s, self.g
        ctx.check(ctx.lib.epsr_truncation(ctx.h, g.npt, D._ptr(self.x), F32(s.rminpt), F32(s.rmaxpt),
                                          D._ptr(self.wfac), D._ptr(self.wfacc)))
        self.foqols, self.ldf = D.foqols(ctx, g, s.rho, s.rmaxgr)
        self.generate_rp()
        self.generate_ep()
        return self

    def generate_rp(self):
        ctx, s, g = self.ctx, self.s, self.g
        qcor = s.qcor.astype(F32)
        if s.coulfac is not None:   # qct = qcor * coulfac (epsr_standalone_rev3.py:5739)
            qcor = (qcor * np.asarray(s.coulfac, F32)).astype(F32)
        self._rp_in = [ctx.tensor(a.astype(F32)) for a in (s.ecor, s.dcor, qcor)]
        extra = {}
        if s.pcoefex is not None and np.any(np.asarray(s.pcoefex) != 0.0):
            extra["pcoefex"] = ctx.tensor(np.asarray(s.pcoefex, F32))
            extra["rminex"] = self.rminex
        if s.qradius is not None and np.any(np.asarray(s.qradius) != 0.0):
            qr = np.asarray(s.qradius, F32)
            extra["qradius_i"] = ctx.tensor(qr[s.n1extr])
            extra["qradius_j"] = ctx.tensor(qr[s.n2extr])
        self._rp_extra = extra
        p = _native_rp_params(s, self._rp_in, **extra)
        ctx.check(ctx.lib.epsr_generate_rp(ctx.h, g.npt, g.npt2, D._ptr(self.x), D._ptr(self.xstar),
                                           D._ptr(self.wfac), D._ptr(self.wfacc), C.byref(p), D._ptr(self.epot),
                                           D._ptr(self.ecoul), D._ptr(self.epot1), D._ptr(self.ecoul1), self.ld))

    def generate_ep(self, absolute_energy=10.0, erequnit=0.0, ereqstep=0.0):
        ctx, s, g = self.ctx, self.s, self.g
        p = EPParams(F32(s.pdmin), F32(s.pdstep), F32(s.rmaxpt), F32(s.psigma2), F32(s.rho), F32(0.0),
                     F32(s.dmax), F32(absolute_energy), F32(erequnit), F32(ereqstep))
        pf, ea = C.c_float(), C.c_float()
        ir = self.irefpcof_h
        ctx.check(ctx.lib.epsr_generate_ep(ctx.h, g.npt, g.npt2, D._ptr(self.x), D._ptr(self.xstar),
                                           D._ptr(self.wfac), s.npairs, ir.ctypes.data_as(C.POINTER(C.c_int)),
                                           s.nterms, self.pcoefp.ctypes.data_as(C.POINTER(C.c_float)),
                                           s.nterms + 1, C.byref(p), D._ptr(self.ecoul), D._ptr(self.xavs),
                                           D._ptr(self.xavs1), D._ptr(self.xavtot), D._ptr(self.xavtot1), self.ld,
                                           C.byref(pf), C.byref(ea)))
        self.pressfac = F32(pf.value)
        self.energabs = F32(ea.value)

    @property
    def tables(self):
        return D.DeviceTables(self.ctx, self.ld, self.g.npt2, self.xstar, self.xavs, self.xavs1, self.ecoul,
                              self.ecoul1)

    # ---------------------------------------------------------------- the loop
    def cells_and_ledger(self):
        """initialise_cellneighbourlist + fmolec1 (:1259-1281).  Returns (ftest, ftestb)."""
        self.ledger.cells_init(self.s.nmolcell, T.sepmax(self.s, self.g))
        if self.s.intra is not None:
            r = self.s.mc_run
            tempdih = F32(F32(r["tempdihreq"]) * F32(self.s.intra.tempb))
            self.eintra = D.en_mole_intra(self.ctx, self.box, self.tables, self.s.simt, r["tempangreq"], tempdih,
                                          out=self.eintra)
        return self.ledger.fmolec1(self.tables, self.eintra)

    def accumulate(self, nsum: int, nsums: int = 0, nesum: int = 0):
        """sumsq + sumgr (epsr_standalone_rev3.py:1632-1642) on the current sofq and wav: the
        running averages soq/sq2 and gor/gr2 and the zero limits entsig."""
        s, g, ctx = self.s, self.g, self.ctx
        if getattr(self, "soq", None) is None:
            self.soq = ctx.zeros(2 * s.npairs * self.ldq)
            self.sq2 = ctx.zeros(2 * s.npairs * self.ldq)
            self.gor = ctx.zeros(2 * s.npairs * self.ldw)
            self.gr2 = ctx.zeros(2 * s.npairs * self.ldw)
            self.entsig = ctx.zeros(s.npairs * self.ldw)
        fc1, fc2 = sum_factors(nsum, nsums, nesum)
        D.sumsq(ctx, s.npairs, g.nq, self.ldq, self.sofq, fc1, fc2, self.soq, self.sq2)
        D.sumgr(ctx, s.npairs, g.npt1, self.ldw, self.wav, self.ycoh, fc1, fc2, self.gor, self.gr2, self.entsig)

    def data_minus_sim(self, datai, wts, nqmind, nqmaxd, nqmin, nqmax, szeros, porodamp, rshmin, porodq=0.0,
                       porodpwr=0.0, nsmoop=1):
        """The data-minus-simulation block of the epsr program (epsr_standalone_rev3.py:
        1703-1830) on the running averages (accumulate first): datai float32 [nfile, nq + 1]
        (the data on this box's Q grid), wts = get_wts(...) (ciwt, cwt, xfile, weightx),
        the per-dataset Q ranges (1-based), zero limits, Porod amplitudes and ftback radii.
        Returns (datat, szerostemp) on the device: the data and zero limits ptgnsq takes."""
        ctx, s, g = self.ctx, self.s, self.g
        nfile, npairs, ldq = datai.shape[0], s.npairs, self.ldq
        dd = np.zeros((nfile, ldq), F32)
        dd[:, :min(ldq, datai.shape[1])] = np.asarray(datai, F32)[:, :ldq]
        if getattr(self, "grt", None) is None:
            self.grt = ctx.zeros(g.npt1 + 1)     # the program's r-indexed work array of ftback
        wx = wts.get("weightx")
        dwx = None
        if wx is not None:
            w = np.zeros((wx.shape[0], ldq, npairs))
            w[:, :wx.shape[1], :] = wx
            dwx = ctx.tensor(w.reshape(-1))
        return D.data_minus_sim(ctx, self.q, ctx.tensor(dd.reshape(-1)), ldq, g.nq, self.soq[:npairs * ldq],
                                self.soq[npairs * ldq:2 * npairs * ldq], wts["ciwt"], wts["cwt"], nqmind, nqmaxd,
                                nqmin, nqmax, szeros, porodamp, rshmin, porodq, porodpwr, nsmoop, self.x, g.npt1,
                                s.rho, self.grt, weightx=dwx, xfile=wts.get("xfile"))

    def rfactor_data(self, datai, wts, nqmind=None, nqmaxd=None, nqmin=None, nqmax=None, porodamp=None,
                     porodq=0.0, porodpwr=0.0):
        """The fit to the data that the epsr program reports after each iteration, its
        "R-factor for data file" (epsr_standalone_rev3.py:2070-2131), on the running averages
        (accumulate first): per dataset the mean over Q points nqmin..nqmax of
        (data - sim)^2, where sim = sum over pairs of ciwt S_intra + cwt S_inter (X-ray datasets:
        weightx (S_intra + S_inter) + backx) plus the Porod term, and data - sim is taken only
        inside the data's range nqmind..nqmaxd; and the mean over datasets (rfactot).  A
        report, formed on the host in FP64.  Returns (rfacd [nfile], rfactot)."""
        s, g = self.s, self.g
        nq, npairs, ldq = g.nq, s.npairs, self.ldq
        datai = np.asarray(datai, np.float64)
        nfile = datai.shape[0]
        soq = self.soq.cpu().numpy().astype(np.float64).reshape(2, npairs, ldq)
        intra, inter = soq[0][:, 1:nq + 1], soq[1][:, 1:nq + 1]
        one = lambda v, d: np.full(nfile, d, np.int32) if v is None else np.asarray(v, np.int32)
        nqmind, nqmaxd = one(nqmind, 1), one(nqmaxd, nq)
        nqmin, nqmax = one(nqmin, 1), one(nqmax, nq)
        porodamp = np.zeros(nfile) if porodamp is None else np.asarray(porodamp, np.float64)
        q = np.asarray(g.q, np.float64)[1:nq + 1]
        ciwt = np.asarray(wts.get("ciwt", np.zeros((nfile, npairs))), np.float64)
        cwt = np.asarray(wts["cwt"], np.float64)
        xfile = wts.get("xfile")
        idx = np.arange(1, nq + 1)
        rfacd = np.zeros(nfile)
        for f in range(nfile):
            ix = int(xfile[f]) if xfile is not None else 0
            if ix > 0:
                wx = np.asarray(wts["weightx"], np.float64)[ix - 1][1:nq + 1]          # [nq, npairs]
                sim = (wx * (intra + inter).T).sum(1) + np.asarray(wts["backx"], np.float64)[ix - 1][1:nq + 1]
            else:
                sim = ciwt[f] @ intra + cwt[f] @ inter
            if porodamp[f] > 0.0:
                sim = sim + np.where(q > 0.0, porodamp[f] * np.power(porodq / np.where(q > 0, q, 1.0), porodpwr), 0.0)
            inside = (idx >= nqmind[f]) & (idx <= nqmaxd[f])
            diff = np.where(inside, datai[f, 1:nq + 1] - sim, 0.0)
            fit = (idx >= nqmin[f]) & (idx <= nqmax[f])
            npts = int(nqmax[f] - nqmin[f] + 1)
            rfacd[f] = float((diff[fit] ** 2).sum() / npts) if npts > 0 else 0.0
        return rfacd, float(rfacd.mean())

    def ptgnsq(self, data, cwt, cwtpot, nqmin=None, szeros=None, potfac=1.0, fit_mode=None, npits=1000,
               paccept=5.0e-4, pstepd=0.002, weightx=None, xfile=None, nqmax=None):
        """One refinement of the empirical potential from the running intermolecular S(Q)
        (accumulate first), as ptgnsq does it (epsr_standalone_rev3.py:3789-4021, neutron
        weights): per dataset the difference data - sum_ic cwt soqinter, a Poisson-basis fit
        of it (fit_expansion), the update of pcoefd/pcoefp, then generate_ep.
        data: float32 [nfile, nq + 1] (Q points 1..nq); cwt float32 [nfile, npairs];
        cwtpot float64/32 [npairs, nfile] (get_wts).  Returns (r-factors, chi^2) per dataset."""
        ctx, s, g = self.ctx, self.s, self.g
        nfile, npairs, nq, ldq = np.asarray(cwt).shape[0], s.npairs, g.nq, self.ldq
        fit_mode = _native.FIT_MC if fit_mode is None else fit_mode
        nterms = s.nterms
        if getattr(self, "pcoefd", None) is None:
            self.pcoefd = np.zeros((nfile + 1, nterms + 1), F32)
            self.pstepd = np.full(nfile + 1, F32(pstepd), F32)
        nqmin = np.full(nfile, 2, np.int32) if nqmin is None else np.asarray(nqmin, np.int32)
        nqmax = np.full(nfile, nq, np.int32) if nqmax is None else np.asarray(nqmax, np.int32)
        if isinstance(szeros, torch.Tensor):
            szeros = szeros.cpu().numpy()
        szeros = np.zeros(nfile, F32) if szeros is None else np.asarray(szeros, F32)
        if isinstance(data, torch.Tensor):       # data_minus_sim's datat, [nfile * ldq] on the device
            dd = data.reshape(-1)
        else:
            dd = ctx.tensor(np.ascontiguousarray(np.asarray(data, F32)[:, :ldq]).reshape(-1))
        qwind = self.soq[npairs * ldq:2 * npairs * ldq]
        cw = np.ascontiguousarray(np.asarray(cwt, F32)).reshape(-1)
        dwx = None
        if weightx is not None:
            w = np.zeros((weightx.shape[0], ldq, npairs))
            w[:, :weightx.shape[1], :] = weightx
            dwx = ctx.tensor(w.reshape(-1))
        qt = D.ptgnsq_diff(ctx, nfile, nq, npairs, self.q, dd, ldq, qwind, cw, szeros, nqmin, nqmax, F32(0.0),
                           weightx=dwx, xfile=xfile)
        pcp = np.zeros((npairs + 1, nterms + 1), F32)
        pcp[1:] = self.pcoefp
        irefp = np.arange(1, npairs + 1, dtype=np.int32)
        efacp = np.ones(npairs + 1, F32)
        cwp = np.zeros((npairs, nfile + 1), F32)
        cwp[:, 1:] = np.asarray(cwtpot, F32)
        rfac = np.zeros(nfile, F32)
        chisq = np.zeros(nfile, F32)
        for i in range(nfile):
            idn = i + 1
            nqm = int(nqmax[i])
            expsq, _ = D.calc_pwrexpq(ctx, nqm, nterms, s.pdmin, s.pdstep, s.rmaxpt, s.psigma2, s.rho, self.q)
            dt = -qt[i * ldq:i * ldq + nqm + 1]
            pw = np.zeros(nterms + 1, F32)
            pw[1:] = self.pcoefd[idn, 1:]
            dts, dw, dwo = ctx.zeros(nqm + 1), ctx.tensor(pw), ctx.tensor(pw)
            r = D.fit_expansion(ctx, dt.contiguous(), expsq, dts, dw, dwo, self.pstepd[idn], F32(paccept), 0.0, 0.0,
                                1, nqm, 1, nterms, npits, 10, self.ran1 if fit_mode == _native.FIT_MC else None,
                                mode=fit_mode)
            self.pstepd[idn] = r.stepr[0]
            chisq[i] = r.chisqnew[0]
            D.ptgnsq_update(ctx, idn, nterms, npairs, dwo.cpu().numpy(), irefp, efacp, cwp.reshape(-1), nfile + 1,
                            F32(potfac), self.pcoefd, pcp, nterms + 1)
            i0 = int(nqmin[i]) if szeros[i] == 0.0 else 1
            fit_full = ctx.zeros(ldq)
            fit_full[:nqm + 1] = dts
            rfac[i] = D.ptgnsq_rfactor(ctx, i0, nq, fit_full, qt[i * ldq:(i + 1) * ldq].contiguous())
        self.pcoefp = np.ascontiguousarray(pcp[1:])
        self.generate_ep()
        return rfac, chisq

    def press(self, energf=None, pressf=None, wsum=None):
        """press (ato_manipulation_rev4.py:3901-4063) over the current ledger: (energy per
        molecule, pressure in kbar, energabs, energthreeb, prthreebout)."""
        from . import press as P
        s = self.s
        vol = F32(s.dmax) * F32(s.dmax) * F32(s.dmax)
        return P.press(self.ledger, self.tables, self.g.npt1, s.simt, vol, pcoefp=self.pcoefp, energf=energf,
                       pressf=pressf, wsum=wsum, nmol=s.nmol)

    def molfit(self, tables=None, ncycles=1):
        """ncycles molfit cycles (every branch when the box has an intramolecular topology,
        one call per cycle; rigid boxes run them back to back on the GPU)"""
        if isinstance(self.mc, D.MCFullState):
            for _ in range(ncycles):
                D.molfit_full(self.ctx, self.box, tables or self.tables, self.ledger, self.mc, self.ran1)
        else:
            D.molfit(self.ctx, self.box, tables or self.tables, self.ledger, self.mc, self.ran1, ncycles=ncycles)
        return self.mc

    def atogen(self, ngrsamples=0):
        _, _, _, ngr = D.atogen(self.ctx, self.box, self.g.npt1, self.ldw, self.ycoh, self.irefpcof, self.rminex,
                                ngrsamples, self.ran1, wav=self.wav, potex=self.potex, potexmax=self.potexmax)
        return ngr

    def gensq(self, mode=1, ctx=None):
        D.gensq(ctx or self.ctx, self.g.nq, self.s.npairs, self.g.npt1, self.ldw, self.wav, self.ycoh, self.foqols,
                self.ldf, self.sofq, self.ldq, mode=mode)
        return self.sofq

    def stog(self, mode=1, ctx=None):
        """g(r) of the intermolecular partials on the r grid 1..npt1 (stog, Q points 2..nq, no
        weights), into self.gofr [npairs][npt1 + 1].  ctx: another context on the same
        device (a graph's), else the box's own."""
        s = self.s
        ctx = ctx or self.ctx
        self.stog_setup(ctx)
        D.stog(ctx, s.rho, self.q, self.sofq[s.npairs * self.ldq:], self.ldq, self.r, self.gofr, self.ldg,
               self.nqmin, self.nqmax, self.nr, s.npairs, self.xwt, mode=mode)
        return self.gofr

    def stog_setup(self, ctx=None):
        """the r grid, g(r) array and Q limits of stog (once)"""
        s, g = self.s, self.g
        ctx = ctx or self.ctx
        if getattr(self, "gofr", None) is None:
            self.nr = g.npt1
            self.r = ctx.tensor(g.x[:self.nr + 1])
            self.ldg = self.nr + 1
            self.gofr = ctx.zeros(s.npairs * self.ldg)
            self.nqmin = ctx.tensor(np.full(s.npairs, 2, np.int32))
            self.nqmax = ctx.tensor(np.full(s.npairs, g.nq, np.int32))
            self.xwt = ctx.tensor(np.zeros(s.npairs, F32))


def transforms_batch(ctx, boxes):
    """gensq and stog (mode 1) of refinements on one grid (replicas of one system) as one
    GEMM each (epsr_gensq_batch, epsr_stog_batch): every box's S(Q) and g(r) equal its own
    gensq(mode=1) and stog(mode=1) bit for bit.  The boxes' inputs must be complete (their
    streams synchronised) before the call; the outputs are complete when it returns."""
    R0 = boxes[0]
    g = R0.g
    for R in boxes:
        assert (R.g.nq, R.g.npt1, R.ldf, R.ldq) == (g.nq, g.npt1, R0.ldf, R0.ldq), "boxes on different grids"
        R.stog_setup(R.ctx)
    D.gensq_batch(ctx, g.nq, g.npt1, R0.ycoh, R0.foqols, R0.ldf,
                  [(R.s.npairs, R.wav, R.ldw, R.sofq, R.ldq) for R in boxes])
    D.stog_batch(ctx, R0.s.rho, R0.q, R0.ldq, R0.r, R0.nr,
                 [(R.sofq[R.s.npairs * R.ldq:], R.gofr, R.ldg, R.nqmin, R.nqmax, R.s.npairs, R.xwt) for R in boxes])


class EPParams(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("pdmin", "pdstep", "pdmax", "psigma2", "rho", "rbroad", "dmax",
                                         "absolute_energy", "erequnit", "ereqstep")]


class RPParams(C.Structure):
    _fields_ = [("npairs", C.c_int), ("ecor", C.c_void_p), ("dcor", C.c_void_p), ("qcor", C.c_void_p),
                ("qradius_i", C.c_void_p), ("qradius_j", C.c_void_p), ("pcoefex", C.c_void_p),
                ("rminex", C.c_void_p), ("power", C.c_float), ("refpotfac", C.c_float), ("expecf", C.c_float),
                ("expecfpwr", C.c_float)]


def _native_rp_params(s, tensors, pcoefex=None, rminex=None, qradius_i=None, qradius_j=None):
    ecor, dcor, qcor = tensors
    return RPParams(s.npairs, D._ptr(ecor), D._ptr(dcor), D._ptr(qcor), D._ptr(qradius_i), D._ptr(qradius_j),
                    D._ptr(pcoefex), D._ptr(rminex), F32(s.power), F32(s.refpotfac), F32(s.expecf),
                    F32(s.expecfpwr))
