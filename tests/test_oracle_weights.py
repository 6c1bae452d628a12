"""The oracle's invert_cwt, pinned by what its Monte Carlo minimises: the distance of the
cross array [cwt; (1 - feedback) I] . cwtt from the unit matrix, whose minimum over cwtt is
the Moore-Penrose pseudo-inverse with the value nfile (the rank deficit of the stacked
system).  The walk must end there."""

import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as po

F32 = np.float32


@pytest.mark.parametrize("nfile,npairs,seed", [(3, 3, 1), (3, 6, 2), (1, 3, 3), (4, 10, 4)])
def test_invert_cwt_reaches_the_pseudo_inverse(nfile, npairs, seed):
    rng = np.random.default_rng(seed)
    cwt = rng.uniform(0.05, 0.6, (nfile, npairs))
    nnew = nfile + npairs
    cwtt = np.zeros((npairs, nnew))
    carray = np.zeros(nnew)
    co = np.zeros((nnew, nnew))
    efacm = rng.uniform(0.5, 1.0, nfile).astype(F32)
    fb = F32(0.9)
    r = po.Ran1.seeded(-77)
    po.lib().ora_invert_cwt(cwt.ctypes.data_as(po.dp), cwtt.ctypes.data_as(po.dp), carray.ctypes.data_as(po.dp),
                            co.ctypes.data_as(po.dp), fb, po.Pf(efacm), nfile, npairs, C.byref(r))
    fb1 = float(F32(1.0) - fb)
    # cwt comes back scaled by feedback * efacm, cwtt scaled by the same factors per column
    A = np.vstack([cwt, fb1 * np.eye(npairs)])
    X = np.linalg.pinv(A)
    scale = np.concatenate([float(fb) * efacm.astype(np.float64), np.full(npairs, fb1)])
    Xmc = cwtt / scale[None, :]
    assert np.abs(Xmc - X).max() <= 1e-5 * max(1.0, np.abs(X).max())
    total = ((A @ Xmc - np.eye(nnew)) ** 2).sum()
    assert abs(total - nfile) < 1e-9
    # carray holds the last row of the cross array cwtt . [cwt; fb1 I] (close to a unit row)
    last = Xmc[npairs - 1] @ A
    np.testing.assert_allclose(carray[:npairs], last, atol=1e-6)
