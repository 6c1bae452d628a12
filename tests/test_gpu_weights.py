"""invert_cwt through the library (host code behind the C ABI) against the oracle: the
same ran1 state gives the same walk, so every output agrees bit for bit.  Marked gpu
because the library's context is created on a device."""

import ctypes as C

import numpy as np
import pytest

from epsr_python_translation_amd import pipeline as PL
from oracle import pyoracle as po

F32 = np.float32
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from epsr_python_translation_amd import device as D
    return D.Context(0)


@pytest.mark.parametrize("nfile,npairs,seed", [(3, 3, 1), (3, 21, 5)])
def test_invert_cwt_bit_exact(ctx, nfile, npairs, seed):
    from epsr_python_translation_amd import device as D
    print(f"invert_cwt parity nfile={nfile} npairs={npairs}", flush=True)
    rng = np.random.default_rng(seed)
    cwt = rng.uniform(0.0, 0.6, (nfile, npairs))
    nnew = nfile + npairs
    cwtt = rng.normal(0.0, 0.1, (npairs, nnew))
    efacm = rng.uniform(0.5, 1.0, nfile).astype(F32)
    ca, co = np.zeros(nnew), np.zeros((nnew, nnew))
    c1, t1 = cwt.copy(), cwtt.copy()
    r = po.Ran1.seeded(-5)
    it = po.lib().ora_invert_cwt(c1.ctypes.data_as(po.dp), t1.ctypes.data_as(po.dp), ca.ctypes.data_as(po.dp),
                                 co.ctypes.data_as(po.dp), F32(0.8), po.Pf(efacm), nfile, npairs, C.byref(r))
    st = D.ran1_state_array(-5)
    c2, t2 = cwt.copy(), cwtt.copy()
    carray, carrayold, it2 = PL.invert_cwt(ctx, c2, t2, F32(0.8), efacm, st)
    assert it2 == it
    np.testing.assert_array_equal(st, r.as_array())
    for a, b in ((c1, c2), (t1, t2), (ca, carray), (co, carrayold)):
        assert np.array_equal(a.view(np.int64), b.view(np.int64))
