"""epsr_gensq_noalloc (gensq mode 1 with its scratch in the library's fixed device array)
against epsr_gensq mode 1 through the C ABI: bit for bit on two calls in a row, the second
smaller (the array keeps the first call's contents), plus the refusal of a size beyond the
array."""

import ctypes as C

import numpy as np
import pytest
import torch

F32 = np.float32
pytestmark = pytest.mark.gpu


def _case(ctx, rng, nq, npairs, npt1):
    ldw, ldf, ldq = npt1 + 1, npt1 + 3, nq + 2
    wav = rng.uniform(0.0, 2.0, 2 * npairs * ldw).astype(F32)
    wav[rng.uniform(size=wav.size) < 0.2] = 0.0
    ycoh = rng.uniform(0.5, 1.5, ldw).astype(F32)
    fo = rng.uniform(-1.0, 1.0, (nq + 1) * ldf).astype(F32)
    return dict(nq=nq, npairs=npairs, npt1=npt1, ldw=ldw, ldf=ldf, ldq=ldq, wav=ctx.tensor(wav),
                ycoh=ctx.tensor(ycoh), fo=ctx.tensor(fo))


def _run(ctx, c, noalloc):
    from epsr_python_translation_amd import device as D
    out = ctx.tensor(np.full(2 * c["npairs"] * c["ldq"], -7.0, F32))
    p = D._ptr
    if noalloc:
        rc = ctx.lib.epsr_gensq_noalloc(ctx.h, c["nq"], c["npairs"], c["npt1"], c["ldw"], p(c["wav"]), p(c["ycoh"]),
                                        p(c["fo"]), c["ldf"], None, p(out), c["ldq"])
    else:
        rc = ctx.lib.epsr_gensq(ctx.h, c["nq"], c["npairs"], c["npt1"], c["ldw"], p(c["wav"]), p(c["ycoh"]),
                                p(c["fo"]), c["ldf"], None, p(out), c["ldq"], 1)
    ctx.check(rc)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_noalloc_equals_gensq_on_two_calls_in_a_row():
    from epsr_python_translation_amd import device as D
    ctx = D.Context(0)
    rng = np.random.default_rng(5)
    # the 3200-atom box's shape (21 pairs, 600 Q points, 528 bins), then a smaller one
    for nq, npairs, npt1 in ((600, 21, 528), (200, 3, 300)):
        print(f"gensq no-alloc: nq={nq} npairs={npairs} npt1={npt1}", flush=True)
        c = _case(ctx, rng, nq, npairs, npt1)
        want = _run(ctx, c, False)
        got = _run(ctx, c, True)
        assert np.array_equal(got.view(np.int32), want.view(np.int32))
        assert np.all(np.isfinite(got.reshape(2, npairs, -1)[:, :, 1:nq + 1]))


def test_noalloc_refuses_a_size_beyond_its_array():
    from epsr_python_translation_amd import device as D
    ctx = D.Context(0)
    rng = np.random.default_rng(6)
    c = _case(ctx, rng, 9000, 40, 20)
    p = D._ptr
    out = ctx.zeros(2 * c["npairs"] * c["ldq"])
    rc = ctx.lib.epsr_gensq_noalloc(ctx.h, c["nq"], c["npairs"], c["npt1"], c["ldw"], p(c["wav"]), p(c["ycoh"]),
                                    p(c["fo"]), c["ldf"], None, p(out), c["ldq"])
    assert rc == -1
    assert b"scratch" in ctx.lib.epsr_last_error(ctx.h)
