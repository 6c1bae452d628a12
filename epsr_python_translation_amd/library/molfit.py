"""molfit.molfit, en_change and en_mole_pair over the COMMON mirror
(src/epsr26/library/molfit.py:51-541, 805-919, 1527-1578).  With an intramolecular topology
loaded (common.setup from a box read from files) molfit runs every branch of the
reference's cycle on the GPU: the shake every `intra` cycles, internal rotations every
`rotfreq` cycles and the whole-molecule moves.  Without one it runs the whole-molecule
cycles and refuses the cycles that would move atoms inside molecules."""

from __future__ import annotations

import numpy as np
import torch

from .. import _native
from .. import device as D
from ..common import clustercom, epsr16, gpu, molecom, ran1com

F32 = np.float32
last_stats = None


def _moves_inside_molecules(nisum, intra, rotfreq):
    """True when the reference's cycle would move atoms inside molecules (molfit.py:124,295)."""
    if nisum >= intra:
        return True
    if molecom.stepri > 0.0 and np.any(clustercom.ngrot[1:clustercom.nmol + 1] > 0):
        if rotfreq == 0:
            raise ZeroDivisionError("molfit: rotfreq = 0 with internal rotations (nisum / rotfreq)")
        return nisum == rotfreq * int(nisum / rotfreq)
    return False


def molfit(fadds, faddbs, nisum, intra, rotfreq, ireset):
    """One MC cycle on the ledger of the last fmolec1.  Returns (fadds, faddbs, nisum)."""
    global last_stats
    if np.any(clustercom.tethered[1:clustercom.nmol + 1]):
        # the reference's moves keep a tethered molecule's anchor atom in place
        # (check_tethered, molfit.py:219, :317, :397); the GPU cycle has no such moves
        raise NotImplementedError("molfit: tethered molecules are not handled by the GPU path")
    if gpu.intra is None and _moves_inside_molecules(nisum, intra, rotfreq):
        raise NotImplementedError("molfit: this cycle moves atoms inside molecules (shake or internal rotations), "
                                  "which needs the box's intramolecular topology (load it from .ato files)")
    if not gpu.has_ledger():
        raise RuntimeError("molfit: no pair-energy ledger (call fmolec1 first)")
    c = gpu.ctx()
    b = gpu.upload_coords()
    led = gpu.ledger()
    nmol, nmt = clustercom.nmol, clustercom.nmoltype
    st = ran1com.state()
    if gpu.intra is not None:
        mc = D.MCFullState(c, nmol, nmt, molecom.simt, stepr=molecom.stepr, stepm=molecom.stepm,
                           stepri=molecom.stepri, stepmi=molecom.stepmi, rejrate=epsr16.rejrate,
                           tempangwt=molecom.tempangwt, tempdihwt=molecom.tempdihwt, ireset=int(ireset),
                           intra=int(intra), rotfreq=int(rotfreq), nisum=int(nisum),
                           wholerot=clustercom.wholerot[1:nmt + 1], wholetrans=clustercom.wholetrans[1:nmt + 1],
                           internrot=clustercom.internrot[1:nmt + 1], interntrans=clustercom.interntrans[1:nmt + 1],
                           nstepreject=clustercom.nstepreject[1:nmol + 1])
        mc.fadds, mc.faddbs = F32(fadds), F32(faddbs)
        D.molfit_full(c, b, gpu.tables(), led, mc, st)
        clustercom.internrot[1:nmt + 1] = mc.internrot
        clustercom.interntrans[1:nmt + 1] = mc.interntrans
        clustercom.ninter[1, 1:nmt + 1], clustercom.ninter[2, 1:nmt + 1] = mc.ninter[0::2], mc.ninter[1::2]
        clustercom.nmovei[1, 1:nmt + 1], clustercom.nmovei[2, 1:nmt + 1] = mc.nmovei[0::2], mc.nmovei[1::2]
        nisum, faddbs = mc.nisum, mc.faddbs
    else:
        mc = D.MCState(c, nmol, nmt, molecom.simt, stepr=molecom.stepr, stepm=molecom.stepm, rejrate=epsr16.rejrate,
                       ireset=int(ireset), wholerot=clustercom.wholerot[1:nmt + 1].copy(),
                       wholetrans=clustercom.wholetrans[1:nmt + 1].copy())
        mc.nstepreject.copy_(torch.from_numpy(np.ascontiguousarray(clustercom.nstepreject[1:nmol + 1])))
        mc.fadds = F32(fadds)
        D.molfit(c, b, gpu.tables(), led, mc, st)
        # molecules without rotational groups zero their type's internal rotation step
        for im in range(1, nmol + 1):
            if clustercom.ngrot[im] == 0:
                clustercom.internrot[clustercom.mtype[im]] = 0.0
        nisum = nisum + 1
        faddbs = F32(F32(faddbs) + F32(F32(0.0) / F32(nmol)))
    ran1com.set_state(st)
    gpu.download_coords()
    clustercom.wholerot[1:nmt + 1] = mc.wholerot
    clustercom.wholetrans[1:nmt + 1] = mc.wholetrans
    clustercom.nwhole[1, 1:nmt + 1] = mc.nwhole[0::2]
    clustercom.nwhole[2, 1:nmt + 1] = mc.nwhole[1::2]
    clustercom.nmovem[1, 1:nmt + 1] = mc.nmovem[0::2]
    clustercom.nmovem[2, 1:nmt + 1] = mc.nmovem[1::2]
    clustercom.nstepreject[1:nmol + 1] = mc.nstepreject.cpu().numpy()
    last_stats = mc.stats
    return mc.fadds, F32(faddbs), nisum


def en_change(im):
    """Energy change of molecule im (1-based) at its current coordinates against the
    ledger of the last fmolec1."""
    if not gpu.has_ledger():
        raise RuntimeError("en_change: no pair-energy ledger (call fmolec1 first)")
    gpu.upload_coords()
    return gpu.ledger().en_change(gpu.tables(), int(im) - 1)


def en_mole_pair(im, jm):
    """Pair energy of molecules im and jm (1-based)."""
    c = gpu.ctx()
    b = gpu.upload_coords()
    return F32(D.en_mole_pair(c, b, gpu.tables(), [int(im) - 1], [int(jm) - 1]).cpu().numpy()[0])
