"""press on the GPU: the whole-box intermolecular energy and virial pressure
(src/epsr26/library/ato_manipulation_rev4.py:3901-4063), with the reference's return
values and its host bookkeeping.

The pair sums run in `epsr_press` (csrc/press.hip) over the ledger's partner lists; this
module adds what press does on the host side of those sums: abs_energy_ep (:4128-4157)
and the per-dataset energf/pressf normalisation (:4056-4061).  Three-body potentials and
coordination constraints are outside this package (energthreeb and prthreebout are 0).
"""

from __future__ import annotations

import numpy as np

F32 = np.float32


def abs_energy_ep(pcoefp):
    """The largest (max - min) range of a pair's EP coefficients 2..ncoeffp, each range
    starting from 0 (ato_manipulation_rev4.py:4128-4157).  pcoefp: [npairs, ncoeffp + 1]
    with column 0 unused and column 1 the pair's multiplier."""
    energabs = F32(0.0)
    p = np.asarray(pcoefp, dtype=F32)
    for ip in range(p.shape[0]):
        lo = hi = F32(0.0)
        for term in p[ip, 2:]:
            lo = min(lo, term)
            hi = max(hi, term)
        energabs = max(energabs, F32(hi - lo))
    return F32(energabs)


def normalise_dataset_terms(energf, pressf, wsum, nmol, volume):
    """energf(if) /= nmol wsum(if); pressf(if) = ((-pressf/3) 100) / (6.0222 volume wsum)
    in float32, in place (ato_manipulation_rev4.py:4056-4061)."""
    for i in range(len(energf)):
        energf[i] = F32(energf[i]) / (F32(nmol) * F32(wsum[i]))
        pressf[i] = ((-F32(pressf[i]) / F32(3.0)) * F32(100.0)) / (F32(6.0222) * F32(volume) * F32(wsum[i]))


def press(ledger, tables, npt1, simt, volume, pcoefp=None, energf=None, pressf=None, wsum=None, nmol=None):
    """The reference's press: returns (energy, pressc, energabs, energthreeb, prthreebout)
    and normalises energf/pressf in place when given."""
    energy, pressc, _ = ledger.press(tables, npt1, simt, volume)
    energabs = abs_energy_ep(pcoefp) if pcoefp is not None else F32(0.0)
    if energf is not None:
        normalise_dataset_terms(energf, pressf, wsum, nmol if nmol is not None else ledger.box.nmol, volume)
    return energy, pressc, energabs, F32(0.0), F32(0.0)
