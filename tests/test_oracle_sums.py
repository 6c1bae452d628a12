"""The oracle's running averages (sumsq, sumgr), pinned by what they must equal:
after n updates with the reference's factors fc1 = 1/n, fc2 = (n-1)/n, the average is
the mean of the n inputs and the error bar their population standard deviation; the
zero limit is the running sum of ycoh (g - 1) along r."""

import numpy as np

from epsr_python_translation_amd import pipeline as PL
from oracle import pyoracle as po

F32 = np.float32


def _run(npairs, npts, ld, seq, ycoh=None):
    avg = np.zeros(2 * npairs * ld, F32)
    err = np.zeros(2 * npairs * ld, F32)
    ent = np.zeros(npairs * ld, F32)
    L = po.lib()
    for n, x in enumerate(seq, start=1):
        fc1, fc2 = PL.sum_factors(n)
        if ycoh is None:
            L.ora_sumsq(npairs, npts, ld, po.Pf(x), fc1, fc2, po.Pf(avg), po.Pf(err))
        else:
            L.ora_sumgr(npairs, npts, ld, po.Pf(x), po.Pf(ycoh), fc1, fc2, po.Pf(avg), po.Pf(err), po.Pf(ent))
    return avg, err, ent


def test_running_average_is_mean_and_standard_deviation():
    rng = np.random.default_rng(1)
    npairs, nq, ld = 3, 50, 52
    seq = [rng.normal(1.0, 0.3, 2 * npairs * ld).astype(F32) for _ in range(7)]
    avg, err, _ = _run(npairs, nq, ld, seq)
    X = np.stack(seq).astype(np.float64).reshape(7, 2 * npairs, ld)[:, :, 1:nq + 1]
    a = avg.reshape(2 * npairs, ld)[:, 1:nq + 1]
    e = err.reshape(2 * npairs, ld)[:, 1:nq + 1]
    np.testing.assert_allclose(a, X.mean(0), rtol=2e-6, atol=1e-6)
    np.testing.assert_allclose(e, X.std(0), rtol=1e-3, atol=5e-6)
    # point 0 is not a Q point and stays untouched
    assert np.all(avg.reshape(2 * npairs, ld)[:, 0] == 0.0)


def test_constant_input_gives_no_error_bar():
    npairs, nq, ld = 2, 10, 11
    seq = [np.full(2 * npairs * ld, F32(0.75))] * 5
    avg, err, _ = _run(npairs, nq, ld, seq)
    v = avg.reshape(2 * npairs, ld)[:, 1:nq + 1]
    assert np.all(v == F32(0.75)) and np.all(err.reshape(2 * npairs, ld)[:, 1:nq + 1] <= 1e-6)


def test_sumgr_divides_by_ycoh_and_sums_the_zero_limit():
    rng = np.random.default_rng(2)
    npairs, nz, ld = 2, 40, 42
    ycoh = np.concatenate([[0.0, 0.0], rng.uniform(0.5, 2.0, ld - 2)]).astype(F32)
    ycoh[5] = 0.0                                  # a bin without counts in the uniform histogram
    g = rng.uniform(0.5, 1.5, (2 * npairs, ld)).astype(F32)
    wav = (g * ycoh[None, :]).astype(F32).reshape(-1)
    avg, err, ent = _run(npairs, nz, ld, [wav], ycoh=ycoh)
    a = avg.reshape(2 * npairs, ld)
    assert np.all(a[:, 1] == 0.0) and np.all(a[:, 5] == 0.0)      # r point 1 untouched, ycoh = 0 gives 0
    np.testing.assert_allclose(a[:, 2:nz + 1], (wav.reshape(2 * npairs, ld) / np.where(ycoh > 0, ycoh, 1))[:, 2:nz + 1]
                               * (ycoh[2:nz + 1] > 0), rtol=1e-6)
    inter = a[npairs:]
    want = np.cumsum(ycoh[None, 2:nz + 1].astype(np.float64) * (inter[:, 2:nz + 1] - 1.0), axis=1)
    e = ent.reshape(npairs, ld)
    assert np.all(e[:, 1] == 0.0)
    np.testing.assert_allclose(e[:, 2:nz + 1], want, rtol=1e-5, atol=1e-5)
