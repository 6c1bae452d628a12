"""sumsq and sumgr on the GPU against the oracle, bit for bit, over several accumulation
steps (random inputs at the water20000 sizes, and the pipeline's own S(Q) and histograms)."""

import numpy as np
import pytest
import torch

from epsr_python_translation_amd import pipeline as PL
from oracle import pyoracle as po
from tests import helpers as H

F32 = np.float32
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from epsr_python_translation_amd import device as D
    return D.Context(0)


def test_running_averages_bit_exact(ctx):
    from epsr_python_translation_amd import device as D
    print("sumsq/sumgr parity: 3 pairs, 1200 Q, 2100 r, 4 steps", flush=True)
    rng = np.random.default_rng(9)
    npairs, nq, ldq, nz, ldw = 3, 1200, 1201, 2100, 2101
    ycoh = np.concatenate([[0.0, 0.0], rng.uniform(0.0, 3.0, ldw - 2)]).astype(F32)
    ycoh[rng.integers(2, ldw, 20)] = 0.0
    L = po.lib()
    a = [np.zeros(2 * npairs * ldq, F32) for _ in range(2)]
    b = [np.zeros(2 * npairs * ldw, F32) for _ in range(2)]
    ent = np.zeros(npairs * ldw, F32)
    ga = [ctx.zeros(2 * npairs * ldq) for _ in range(2)]
    gb = [ctx.zeros(2 * npairs * ldw) for _ in range(2)]
    gent = ctx.zeros(npairs * ldw)
    gy = ctx.tensor(ycoh)
    for n in range(1, 5):
        fc1, fc2 = PL.sum_factors(n)
        q = rng.normal(1.0, 0.5, 2 * npairs * ldq).astype(F32)
        w = (rng.uniform(0.0, 2.0, 2 * npairs * ldw) * np.tile(ycoh, 2 * npairs)).astype(F32)
        L.ora_sumsq(npairs, nq, ldq, po.Pf(q), fc1, fc2, po.Pf(a[0]), po.Pf(a[1]))
        L.ora_sumgr(npairs, nz, ldw, po.Pf(w), po.Pf(ycoh), fc1, fc2, po.Pf(b[0]), po.Pf(b[1]), po.Pf(ent))
        D.sumsq(ctx, npairs, nq, ldq, ctx.tensor(q), fc1, fc2, ga[0], ga[1])
        D.sumgr(ctx, npairs, nz, ldw, ctx.tensor(w), gy, fc1, fc2, gb[0], gb[1], gent)
        torch.cuda.synchronize()
        for ref, got in ((a[0], ga[0]), (a[1], ga[1]), (b[0], gb[0]), (b[1], gb[1]), (ent, gent)):
            assert np.array_equal(ref.view(np.int32), got.cpu().numpy().view(np.int32)), n


def test_refinement_accumulates_its_own_partials(ctx):
    print("Refinement.accumulate on c0", flush=True)
    s, g = H.system("c0"), H.grid("c0")
    R = PL.Refinement(ctx, s, idum=-3)
    R.build_tables()
    R.cells_and_ledger()
    L = po.lib()
    npairs = s.npairs
    soq = [np.zeros(2 * npairs * R.ldq, F32) for _ in range(2)]
    gor = [np.zeros(2 * npairs * R.ldw, F32) for _ in range(2)]
    ent = np.zeros(npairs * R.ldw, F32)
    for n in range(1, 3):
        R.molfit()
        R.atogen()
        R.gensq()
        R.accumulate(n)
        torch.cuda.synchronize()
        fc1, fc2 = PL.sum_factors(n)
        L.ora_sumsq(npairs, g.nq, R.ldq, po.Pf(R.sofq.cpu().numpy()), fc1, fc2, po.Pf(soq[0]), po.Pf(soq[1]))
        L.ora_sumgr(npairs, g.npt1, R.ldw, po.Pf(R.wav.cpu().numpy()), po.Pf(g.ycoh), fc1, fc2, po.Pf(gor[0]),
                    po.Pf(gor[1]), po.Pf(ent))
        assert np.array_equal(soq[0].view(np.int32), R.soq.cpu().numpy().view(np.int32))
        assert np.array_equal(soq[1].view(np.int32), R.sq2.cpu().numpy().view(np.int32))
        assert np.array_equal(gor[0].view(np.int32), R.gor.cpu().numpy().view(np.int32))
        assert np.array_equal(gor[1].view(np.int32), R.gr2.cpu().numpy().view(np.int32))
        assert np.array_equal(ent.view(np.int32), R.entsig.cpu().numpy().view(np.int32))
