"""Replicas of the SiO2 refinement on one GPU: each replica's iteration (ledger, MC, atogen,
gensq, stog, running averages, ptgnsq with the fit, the EP update and generate_ep) gives the
same coefficients whether it runs among other replicas on worker streams or alone, bit for
bit; with delayed-acceptance MC driven by a surrogate trained on the GPU, the fit to the data
(the program's R-factor) improves over the iterations."""

import numpy as np
import pytest
import torch

F32 = np.float32
pytestmark = pytest.mark.gpu


def test_replica_in_ensemble_equals_replica_alone():
    from epsr_python_translation_amd import _native
    from epsr_python_translation_amd import device as D
    from epsr_python_translation_amd import replicas as RP
    print("replicas: 3 SiO2 replicas x 2 iterations on 2 workers, then replica 1 alone", flush=True)
    ctx = D.Context(0)
    wts, data, fr = RP.synthetic_targets(ctx)
    ens = RP.ReplicaRefinement(nrep=3, nworkers=2, seed0=40, idum0=-4000)
    ens.set_targets(wts, data)
    res = [ens.iterate(sweeps=1, fit_mode=_native.FIT_NORMAL) for _ in range(2)]
    torch.cuda.synchronize()
    got = ens.boxes[1].pcoefp.copy()
    rf = [r[1][0] for r in res]
    print(f"  replica 1 r-factors {rf[0]} -> {rf[1]}", flush=True)
    del ens
    alone = RP.ReplicaRefinement(nrep=1, nworkers=1, seed0=41, idum0=-4001)
    alone.set_targets(wts, data)
    res1 = [alone.iterate(sweeps=1, fit_mode=_native.FIT_NORMAL) for _ in range(2)]
    torch.cuda.synchronize()
    assert np.array_equal(alone.boxes[0].pcoefp.view(np.int32), got.view(np.int32))
    assert np.array_equal(res1[1][0][0].view(np.int32), res[1][1][0].view(np.int32))
    assert np.all(np.isfinite(res[1][1][0])) and np.all(np.isfinite(res[1][1][1]))


def test_replica_iteration_with_the_monte_carlo_fit():
    """The same iteration with fit_expansion's exact Monte Carlo replay: it draws from the
    replica's ran1 stream, and two runs from the same state agree bit for bit."""
    from epsr_python_translation_amd import _native
    from epsr_python_translation_amd import device as D
    from epsr_python_translation_amd import replicas as RP
    print("replicas: Monte Carlo fit, 1 replica x 1 iteration, twice", flush=True)
    ctx = D.Context(0)
    wts, data, _ = RP.synthetic_targets(ctx)
    out = []
    for _ in range(2):
        ens = RP.ReplicaRefinement(nrep=1, nworkers=1, seed0=42, idum0=-4002)
        ens.set_targets(wts, data)
        st0 = ens.boxes[0].ran1.copy()
        rf, chi, _ = ens.iterate(sweeps=1, fit_mode=_native.FIT_MC)[0]
        torch.cuda.synchronize()
        assert not np.array_equal(ens.boxes[0].ran1, st0)
        assert np.all(np.isfinite(rf)) and np.all(np.isfinite(chi))
        out.append((ens.boxes[0].pcoefp.copy(), ens.boxes[0].ran1.copy(), chi))
        del ens
    assert np.array_equal(out[0][0].view(np.int32), out[1][0].view(np.int32))
    np.testing.assert_array_equal(out[0][1], out[1][1])


def test_refinement_with_delayed_acceptance_fits_the_data():
    """2 replicas x 6 iterations of 5 delayed-acceptance sweeps towards data from a relaxed
    configuration: the program's R-factor for the data (mean over datasets and replicas)
    falls below 0.7 of its first value, and every replica's moves keep being accepted."""
    from epsr_python_translation_amd import _native
    from epsr_python_translation_amd import device as D
    from epsr_python_translation_amd import replicas as RP
    print("replicas: targets from a relaxed SiO2 box, surrogate training", flush=True)
    ctx = D.Context(0)
    wts, data, _ = RP.synthetic_targets(ctx)
    ens = RP.ReplicaRefinement(nrep=2, nworkers=2, seed0=50, idum0=-5000)
    ens.set_targets(wts, data)
    losses = ens.train_surrogate(ctx, nsamples=100_000)
    print(f"  training losses {losses}", flush=True)
    assert losses[-1] < losses[0]
    rfd = []
    for it in range(6):
        res = ens.iterate(sweeps=5, fit_mode=_native.FIT_NORMAL, da=True)
        torch.cuda.synchronize()
        rfd.append(float(np.mean([r[2] for r in res])))
        acc = [R.mc.stats.accepted for R in ens.boxes]
        print(f"  iteration {it}: R-factor {rfd[-1]:.4g}, accepted in the last cycle {acc}", flush=True)
        assert all(a > 0 for a in acc)
        assert all(np.all(np.isfinite(r[0])) and np.all(np.isfinite(r[2])) for r in res)
    assert rfd[-1] < 0.7 * rfd[0], rfd


def test_batched_transforms_equal_each_box_alone():
    """gensq and stog of three SiO2 replicas (one grid) as one GEMM each give every box the
    S(Q) and g(r) of its own mode-1 transforms, bit for bit."""
    from epsr_python_translation_amd import device as D
    from epsr_python_translation_amd import pipeline as PL
    from epsr_python_translation_amd import synth
    print("batched gensq + stog: 3 SiO2 replicas against each alone", flush=True)
    ctx = D.Context(0)
    boxes = []
    for k in range(3):
        R = PL.Refinement(ctx, synth.config4(seed=60 + k), idum=-60 - k).build_tables()
        R.cells_and_ledger()
        R.molfit(ncycles=1)
        R.atogen()
        boxes.append(R)
    alone = []
    for R in boxes:
        R.gensq(mode=1)
        R.stog(mode=1)
        torch.cuda.synchronize()
        alone.append((R.sofq.cpu().numpy().copy(), R.gofr.cpu().numpy().copy()))
        R.sofq.fill_(np.nan)
        R.gofr.fill_(np.nan)
    torch.cuda.synchronize()
    PL.transforms_batch(ctx, boxes)
    torch.cuda.synchronize()
    for R, (sq, gr) in zip(boxes, alone):
        npq = R.s.npairs * R.ldq
        got_sq, got_gr = R.sofq.cpu().numpy(), R.gofr.cpu().numpy()
        # the S(Q) rows the transforms write (1..nq of every column) and every g(r) point
        rows = np.zeros(R.ldq, bool)
        rows[1:R.g.nq + 1] = True
        sel = np.tile(rows, 2 * R.s.npairs)
        assert np.array_equal(got_sq[sel].view(np.int32), sq[sel].view(np.int32))
        cols = np.zeros(R.ldg, bool)
        cols[1:R.nr + 1] = True
        selg = np.tile(cols, R.s.npairs)
        assert np.array_equal(got_gr[selg].view(np.int32), gr[selg].view(np.int32))
        assert np.isfinite(got_sq[sel]).all() and np.isfinite(got_gr[selg]).all()
