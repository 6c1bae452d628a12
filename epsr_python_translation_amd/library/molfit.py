"""molfit.molfit, en_change and en_mole_pair over the COMMON mirror
(src/epsr26/library/molfit.py:51-541, 805-919, 1527-1578).  molfit covers the
whole-molecule branch (nisum < intra, no internal rotation groups, no tethering)."""

from __future__ import annotations

import numpy as np
import torch

from .. import _native
from .. import device as D
from ..common import clustercom, epsr16, gpu, molecom, ran1com

F32 = np.float32
last_stats = None


def molfit(fadds, faddbs, nisum, intra, rotfreq, ireset):
    """One MC cycle of whole-molecule rotations and translations over the cells, on the
    ledger of the last fmolec1.  Returns (fadds, faddbs, nisum)."""
    global last_stats
    if nisum >= intra:
        raise NotImplementedError("molfit: the intramolecular branch (nisum >= intra) is not part of this package")
    if not gpu.has_ledger():
        raise RuntimeError("molfit: no pair-energy ledger (call fmolec1 first)")
    c = gpu.ctx()
    b = gpu.upload_coords()
    led = gpu.ledger()
    nmol, nmt = clustercom.nmol, clustercom.nmoltype
    mc = D.MCState(c, nmol, nmt, molecom.simt, stepr=molecom.stepr, stepm=molecom.stepm, rejrate=epsr16.rejrate,
                   ireset=int(ireset), wholerot=clustercom.wholerot[1:nmt + 1].copy(),
                   wholetrans=clustercom.wholetrans[1:nmt + 1].copy())
    mc.nstepreject.copy_(torch.from_numpy(np.ascontiguousarray(clustercom.nstepreject[1:nmol + 1])))
    mc.fadds = F32(fadds)
    st = ran1com.state()
    D.molfit(c, b, gpu.tables(), led, mc, st)
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
    nisum = 1 if nisum >= intra else nisum + 1
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
