"""epsr_standalone_rev3's hot-path routines over the COMMON mirror
(src/epsr26/programs/epsr_standalone_rev3.py): inita's transform matrix (:3025-3071),
initpt's truncation functions (:3755-3767), gensq (:4730-4802), expand (:4924-4939),
fit_expansion (:4806-4920), calc_pwrexpq (:5100-5153), generate_rp (:5675-5872),
generate_ep (:5875-6023) and ptgnsq (:3789-4021, neutron and X-ray weights)."""

from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _native
from .. import device as D
from .. import pipeline as PL
from .. import tables as T
from ..common import clustercom, epsr16, gpu, molecom, potential_par, ran1com, weights, x_rayf0

F32 = np.float32
SUMMAX = D.SUMMAX

# gensq: False = the reference's sequential sums (bit-identical); True = FP32 MFMA
FAST_TRANSFORMS = False
# fit_expansion: the exact Metropolis replay (reference semantics) or the FP64
# normal-equation solve of the same least-squares problem
FIT_MODE = _native.FIT_MC


def _grid():
    return T.Grid(x=epsr16.x, xstar=epsr16.xstar, ycoh=epsr16.ycoh, r2=T.r2_chain(epsr16.x, epsr16.npt1),
                  q=epsr16.q, npt=epsr16.npt, npt1=epsr16.npt1, npt2=epsr16.npt2, nptmin=epsr16.nptmin,
                  nptmax=epsr16.nptmax, nq=epsr16.nq)


def _cols(a, ncol):
    return np.ascontiguousarray(np.asfortranarray(a)[:, 1:ncol + 1].T).reshape(-1)


def inita_foqols(fwhm=0.0, fwhmq=0.0, lorchqwidth=0.0, lorchqwidthq=0.0):
    """epsr16.foqols(j, i), the histogram -> S(Q) transform of inita, built on the GPU."""
    c = gpu.ctx()
    fo, ldf = D.foqols(c, _grid(), epsr16.rho, epsr16.rmaxgr, fwhm, fwhmq, lorchqwidth, lorchqwidthq)
    h = fo.cpu().numpy().reshape(epsr16.nq + 1, ldf)
    epsr16.foqols[:ldf, :] = h.T
    return epsr16.foqols


def initpt_truncation():
    """potential_par.wfac = truncate(x, rminpt, rmaxpt), wfacc = truncatec(x, rmaxpt)."""
    c = gpu.ctx()
    npt = epsr16.npt
    wf, wfc = c.zeros(npt + 1), c.zeros(npt + 1)
    x = c.tensor(epsr16.x[:npt + 1])   # held for the call: a temporary would be recycled
    c.check(c.lib.epsr_truncation(c.h, npt, D._ptr(x), F32(epsr16.rminpt), F32(epsr16.rmaxpt), D._ptr(wf),
                                  D._ptr(wfc)))
    potential_par.wfac[:npt + 1] = wf.cpu().numpy()
    potential_par.wfacc[:npt + 1] = wfc.cpu().numpy()


def gensq(nq, q, sofq, wav):
    """Partial S(Q): sofq(1,i,ic) intramolecular, sofq(2,i,ic) intermolecular, in place."""
    c = gpu.ctx()
    npt1, npairs = epsr16.npt1, molecom.npairs
    ldw, ldq = npt1 + 1, int(nq) + 1
    dw = np.zeros((2, npairs, ldw), F32)
    dw[0, :, 1:] = wav[1, 1:npt1 + 1, 1:npairs + 1].T
    dw[1, :, 1:] = wav[2, 1:npt1 + 1, 1:npairs + 1].T
    ldf = epsr16.foqols.shape[0]
    fo = np.ascontiguousarray(np.asfortranarray(epsr16.foqols).T).reshape(-1)
    out = c.zeros(2 * npairs * ldq)
    D.gensq(c, int(nq), npairs, npt1, ldw, c.tensor(dw.reshape(-1)), c.tensor(epsr16.ycoh[:ldw]), c.tensor(fo), ldf,
            out, ldq, mode=1 if FAST_TRANSFORMS else 0)
    h = out.cpu().numpy().reshape(2, npairs, ldq)
    sofq[1, 1:nq + 1, 1:npairs + 1] = h[0, :, 1:nq + 1].T
    sofq[2, 1:nq + 1, 1:npairs + 1] = h[1, :, 1:nq + 1].T


def expand(nterms, j, expr, coeff, summax):
    """sum_i expr((j-1)*nterms+i) * coeff(i), clamped at +-summax after every term."""
    c = gpu.ctx()
    out = D.expand_all(c, int(nterms), int(j), c.tensor(np.asarray(expr, F32)), c.tensor(np.asarray(coeff, F32)),
                       F32(summax))
    return F32(out[int(j)].item())


def fit_expansion(yin, expr, yout, wt, wtold, stepr, acceptr, rejrat, chisqnew, nmin, npt, ncfirst, nterms, nits,
                  nfreq):
    """Fit wt so that expand(wt) matches yin(nmin..npt); yout, wt, wtold filled in place,
    ran1com advanced as the reference's fit advances it.  Returns
    (stepr, acceptr, rejrat, chisqnew)."""
    c = gpu.ctx()
    dy, de = c.tensor(np.asarray(yin[:npt + 1], F32)), c.tensor(np.asarray(expr, F32))
    dw, dwo, dyo = c.tensor(np.asarray(wt[:nterms + 1], F32)), c.tensor(np.asarray(wtold[:nterms + 1], F32)), \
        c.zeros(npt + 1)
    st = ran1com.state()
    r = D.fit_expansion(c, dy, de, dyo, dw, dwo, stepr, acceptr, rejrat, chisqnew, nmin, npt, ncfirst, nterms, nits,
                        nfreq, st if FIT_MODE == _native.FIT_MC else None, mode=FIT_MODE)
    if FIT_MODE == _native.FIT_MC:
        ran1com.set_state(st)
    yout[1:npt + 1] = dyo.cpu().numpy()[1:]
    wt[1:nterms + 1] = dw.cpu().numpy()[1:]
    wtold[1:nterms + 1] = dwo.cpu().numpy()[1:]
    return F32(r.stepr[0]), F32(r.acceptr[0]), F32(r.rejrat[0]), F32(r.chisqnew[0])


def calc_pwrexpq(lpt, nterms, dmin, dstep, dmax, sigmab, rho, xin, expr, rmin, ncfirst):
    """Q-space Poisson basis into expr((i-1)*nterms+m).  Returns ncfirst."""
    c = gpu.ctx()
    de, nc = D.calc_pwrexpq(c, lpt, nterms, dmin, dstep, dmax, sigmab, rho, c.tensor(np.asarray(xin, F32)),
                            rmin=rmin)
    expr[:lpt * nterms + 1] = de.cpu().numpy()
    return nc


def generate_rp():
    """Reference-potential tables potential_par.epot/ecoul (+ gradients) from the n-6
    LJ + Coulomb parameters in molecom, on the truncation functions of initpt."""
    c = gpu.ctx()
    npt, npairs = epsr16.npt, molecom.npairs
    ld = npt + 1
    t_in = [c.tensor(np.asarray(a[1:npairs + 1], F32)) for a in (molecom.ecor, molecom.dcor, molecom.qcor)]
    p = PL.RPParams(npairs, D._ptr(t_in[0]), D._ptr(t_in[1]), D._ptr(t_in[2]), None, None, None, None,
                    F32(molecom.power), F32(1.0), F32(0.3), F32(0.0))
    outs = [c.zeros(npairs * ld) for _ in range(4)]
    # inputs held in locals for the duration of the call
    x, xs, wf, wfc = (c.tensor(a[:ld]) for a in (epsr16.x, epsr16.xstar, potential_par.wfac, potential_par.wfacc))
    c.check(c.lib.epsr_generate_rp(c.h, npt, epsr16.npt2, D._ptr(x), D._ptr(xs), D._ptr(wf), D._ptr(wfc), C.byref(p),
                                   *[D._ptr(o) for o in outs], ld))
    for name, o in zip(("epot", "ecoul", "epot1", "ecoul1"), outs):
        getattr(potential_par, name)[:ld, 1:npairs + 1] = o.cpu().numpy().reshape(npairs, ld).T


def generate_ep(changepotfac=False):
    """Empirical-potential tables potential_par.xavs/xavs1/xavtot/xavtot1 from the Poisson
    coefficients pcoefp, scaled by pressfac (multiply_pot_coeffs)."""
    c = gpu.ctx()
    npt, npairs = epsr16.npt, molecom.npairs
    ld = npt + 1
    nterms = potential_par.pcoefp.shape[0] - 1
    nref = potential_par.pcoefp.shape[1] - 1
    pc = np.zeros((nref, nterms + 1), F32)
    pc[:, :] = potential_par.pcoefp[:, 1:].T
    pc = np.ascontiguousarray(pc)
    ir = (potential_par.irefpcof[1:npairs + 1] - 1).astype(np.int32)
    p = PL.EPParams(F32(potential_par.pdmin), F32(potential_par.pdstep), F32(potential_par.pdmax),
                    F32(potential_par.psigma2), F32(epsr16.rho), F32(potential_par.rbroad), F32(clustercom.dmax),
                    F32(potential_par.absolute_energy), F32(potential_par.erequnit), F32(potential_par.ereqstep))
    ec = c.tensor(_cols(potential_par.ecoul, npairs))
    outs = [c.tensor(_cols(getattr(potential_par, n), npairs)) for n in ("xavs", "xavs1", "xavtot", "xavtot1")]
    pf, ea = C.c_float(), C.c_float()
    x, xs, wf = (c.tensor(a[:ld]) for a in (epsr16.x, epsr16.xstar, potential_par.wfac))
    c.check(c.lib.epsr_generate_ep(c.h, npt, epsr16.npt2, D._ptr(x), D._ptr(xs), D._ptr(wf), npairs,
                                   ir.ctypes.data_as(C.POINTER(C.c_int)), nterms,
                                   pc.ctypes.data_as(C.POINTER(C.c_float)), nterms + 1, C.byref(p), D._ptr(ec),
                                   *[D._ptr(o) for o in outs], ld, C.byref(pf), C.byref(ea)))
    potential_par.pcoefp[:, 1:] = pc.T
    potential_par.pressfac, potential_par.energabs = F32(pf.value), F32(ea.value)
    for name, o in zip(("xavs", "xavs1", "xavtot", "xavtot1"), outs):
        getattr(potential_par, name)[:ld, 1:npairs + 1] = o.cpu().numpy().reshape(npairs, ld).T


def ptgnsq(nsum, nq, q, wav, data, qwind, rminpt, rmaxpt, szeros, overallfac, gfac, rminp, potnoi, wtentp, rbroad,
           sigma0, nsmoop):
    """Refine the empirical potential from data minus the weighted intermolecular
    partials qwind(i, ic): per dataset the difference, a Poisson-basis fit of it
    (fit_expansion), the update of pcoefd/pcoefp, then generate_ep.  Neutron weights
    for X-ray datasets (weights.xfile(id) = ix > 0) the Q-dependent weights
    x_rayf0.weightx(i, ic, ix).  potnoi(id) receives chi^2.  Returns the r-factors."""
    c = gpu.ctx()
    nfile, npairs = epsr16.nfile, molecom.npairs
    xfile = np.ascontiguousarray(weights.xfile[1:nfile + 1], np.int32)
    ldq = int(nq) + 1
    qminptemp = epsr16.qminm if epsr16.qminp <= 0.0 else epsr16.qminp
    dq = c.tensor(np.asarray(q[:ldq], F32))
    dd = c.tensor(_cols(data, nfile)) if data.ndim == 2 else c.tensor(np.asarray(data, F32))
    dqw = c.tensor(_cols(qwind, npairs))
    cwt = np.ascontiguousarray(weights.cwt[1:nfile + 1, 1:npairs + 1], F32).reshape(-1)
    szr = np.asarray(szeros[1:nfile + 1], F32)
    wx = None
    if np.any(xfile > 0):
        nx = int(xfile.max())
        w = np.zeros((nx, ldq, npairs))
        w[:, :, :] = np.transpose(x_rayf0.weightx[:ldq, 1:npairs + 1, 1:nx + 1], (2, 0, 1))
        wx = c.tensor(np.ascontiguousarray(w).reshape(-1))
    qt = D.ptgnsq_diff(c, nfile, nq, npairs, dq, dd, ldq, dqw, cwt, szr, epsr16.nqmin[1:nfile + 1],
                       epsr16.nqmax[1:nfile + 1], qminptemp, weightx=wx, xfile=xfile)
    rfac = np.zeros(nfile + 1, F32)
    if not (potential_par.feedback > 0.0 and potential_par.potfac > 0.0):
        potential_par.npits = potential_par.npitss
        return rfac
    nterms_all = potential_par.pcoefp.shape[0] - 1
    for id_ in range(1, nfile + 1):
        nterms = int(potential_par.npcoeff[id_])
        nqmax = int(epsr16.nqmax[id_])
        pwtold = np.zeros(nterms + 1, F32)
        pwtold[1:] = potential_par.pcoefd[1:nterms + 1, id_]
        expsq, _ = D.calc_pwrexpq(c, nqmax, nterms, potential_par.pdmin, potential_par.pdstep, potential_par.pdmax,
                                  epsr16.sigma2[id_], epsr16.rho, dq)
        dt = -qt[(id_ - 1) * ldq:(id_ - 1) * ldq + nqmax + 1]
        dts, dw, dwo = c.zeros(nqmax + 1), c.tensor(pwtold), c.tensor(pwtold)
        st = ran1com.state()
        r = D.fit_expansion(c, dt.contiguous(), expsq, dts, dw, dwo, potential_par.pstepd[id_], potential_par.paccept,
                            0.0, 0.0, 1, nqmax, 1, nterms, potential_par.npits, 10,
                            st if FIT_MODE == _native.FIT_MC else None, mode=FIT_MODE)
        if FIT_MODE == _native.FIT_MC:
            ran1com.set_state(st)
        potential_par.pstepd[id_] = r.stepr[0]
        potnoi[id_] = r.chisqnew[0]
        pw = dwo.cpu().numpy()
        pcd = np.ascontiguousarray(potential_par.pcoefd.T, F32)       # [nfile+1][nterms_all+1]
        pcp = np.ascontiguousarray(potential_par.pcoefp.T, F32)       # [nref+1][nterms_all+1]
        irefp = np.asarray(potential_par.irefpcof[1:npairs + 1], np.int32)   # 1-based rows of pcp
        cwtpot = np.ascontiguousarray(weights.cwtpot[1:npairs + 1, :], F32).reshape(-1)   # [ip][id]
        D.ptgnsq_update(c, id_, nterms, npairs, pw, irefp, np.asarray(potential_par.efacp, F32), cwtpot, nfile + 1,
                        F32(potential_par.potfac), pcd, pcp, nterms_all + 1)
        potential_par.pcoefd[:, :] = pcd.T
        potential_par.pcoefp[:, :] = pcp.T
        i0 = int(epsr16.nqmin[id_]) if szeros[id_] == 0.0 else 1
        fit_full = c.zeros(ldq)
        fit_full[:nqmax + 1] = dts
        rfac[id_] = D.ptgnsq_rfactor(c, i0, int(nq), fit_full, qt[(id_ - 1) * ldq:id_ * ldq].contiguous())
    generate_ep(False)
    potential_par.npits = potential_par.npitss
    return rfac
