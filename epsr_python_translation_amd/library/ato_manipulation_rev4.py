"""ato_manipulation_rev4.atogen, stog and fmolec1 over the COMMON mirror
(src/epsr26/library/ato_manipulation_rev4.py:3640-3899, 4066-4124, 2374-2528)."""

from __future__ import annotations

import numpy as np

from .. import device as D
from .. import tables as T
from ..common import clustercom, epsr16, gpu, molecom, potential_par, ran1com

F32 = np.float32

# stog: False = the reference's per-output sequential sums (mode 0, correctly rounded
# sinf); True = the FP32 MFMA contraction (mode 1, within 1e-5 of the sum of |terms|)
FAST_TRANSFORMS = False


def atogen(wav, ngrsamples):
    """Histogram the configuration into wav(1,iz,ic) (intramolecular) and wav(2,iz,ic)
    (intermolecular), normalised as the reference does; close contacts into
    potential_par.potex/potexmax.  Consumes ran1 when sampling origins.  Returns
    ngrsamples."""
    c = gpu.ctx()
    b = gpu.upload_coords()
    npt1, npairs = epsr16.npt1, molecom.npairs
    ldw = npt1 + 1
    st = ran1com.state()
    irefp = (potential_par.irefpcof[1:npairs + 1] - 1).astype(np.int32)
    rminex = np.ascontiguousarray(potential_par.rminex[1:], dtype=F32)
    dwav, potex, potexmax, ngr = D.atogen(c, b, npt1, ldw, c.tensor(epsr16.ycoh[:ldw]), c.tensor(irefp),
                                          c.tensor(rminex), int(ngrsamples), st)
    ran1com.set_state(st)
    h = dwav.cpu().numpy().reshape(2, npairs, ldw)
    wav[1, 1:npt1 + 1, 1:npairs + 1] = h[0, :, 1:npt1 + 1].T
    wav[2, 1:npt1 + 1, 1:npairs + 1] = h[1, :, 1:npt1 + 1].T
    potential_par.potex[1:npairs + 1] = potex.cpu().numpy()
    potential_par.potexmax[1:npairs + 1] = potexmax.cpu().numpy()
    return ngr


def stog(rho, q, sofq, r, gofr, nqmin, nqmax, nr, npairs, nqdim, nrdim, npdim, xwt):
    """Lorch-windowed sine transform of sofq(nqdim,npdim) to gofr(nrdim,npdim), in place."""
    c = gpu.ctx()
    ldq, ldg = sofq.shape[0], gofr.shape[0]
    s_cols = np.ascontiguousarray(np.asfortranarray(sofq)[:, 1:npairs + 1].T).reshape(-1)
    dg = c.zeros(npairs * ldg)
    D.stog(c, F32(rho), c.tensor(np.asarray(q, F32)), c.tensor(s_cols), ldq, c.tensor(np.asarray(r, F32)), dg, ldg,
           c.tensor(np.asarray(nqmin[1:npairs + 1], np.int32)), c.tensor(np.asarray(nqmax[1:npairs + 1], np.int32)),
           int(nr), int(npairs), c.tensor(np.asarray(xwt[1:npairs + 1], F32)), mode=1 if FAST_TRANSFORMS else 0)
    h = dg.cpu().numpy().reshape(npairs, ldg)
    gofr[1:nr + 1, 1:npairs + 1] = h[:, 1:nr + 1].T


def fmolec1(ierr, num_threads_input, ftest, ftestb, ftethered, fcoordconstraints):
    """Cell lists and the pair-energy ledger on the GPU (kept there for molfit and
    en_change).  Energies per molecule come back summed in FP64 and rounded to float32.
    Returns (ierr, num_threads_input, ftest, ftestb, ftethered, fcoordconstraints)."""
    c = gpu.ctx()
    gpu.upload_coords()
    led = gpu.ledger()
    grid = T.Grid(x=epsr16.x, xstar=epsr16.xstar, ycoh=epsr16.ycoh, r2=None, q=epsr16.q, npt=epsr16.npt,
                  npt1=epsr16.npt1, npt2=epsr16.npt2, nptmin=epsr16.nptmin, nptmax=epsr16.nptmax, nq=epsr16.nq)
    led.cells_init(epsr16.nmolcell, _sepmax(grid))
    if gpu.intra is not None:
        # en_mole_intra of every molecule, with the weights fmolec1 sets (:2416-2419)
        molecom.tempangwt, molecom.tempdihwt = molecom.tempangreq, molecom.tempdihreq
        ei = D.en_mole_intra(c, gpu.box(), gpu.tables(), molecom.simt, molecom.tempangreq,
                             F32(molecom.tempdihreq * molecom.tempb))
    else:
        ei = c.tensor(gpu.eintra if gpu.eintra is not None else np.zeros(clustercom.nmol, F32))
    f64, fb64 = led.fmolec1(gpu.tables(), ei)
    return 0, max(1, int(num_threads_input)), F32(f64), F32(fb64), F32(0.0), F32(fcoordconstraints)


def _sepmax(grid):
    """xstar(nptmax) + maxmolsize + 3/dmax (epsr_standalone_rev3.py:1259)."""
    n = molecom.natom
    r = np.sqrt(molecom.xa[1:n + 1].astype(np.float64) ** 2 + molecom.ya[1:n + 1].astype(np.float64) ** 2 +
                molecom.za[1:n + 1].astype(np.float64) ** 2)
    nat = clustercom.nato[1:clustercom.nmol + 1]
    maxmolsize = F32(2.0 * r.max()) if nat.max() > 1 else F32(0.0)
    return F32(F32(grid.xstar[grid.nptmax] + maxmolsize) + F32(F32(3.0) / F32(clustercom.dmax)))
