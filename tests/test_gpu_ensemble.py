"""Batched boxes on one GPU: launch shapes, HIP graph replay and the ensemble.

* Launch shapes (epsr_box_set_launch) change the time, never the result: MC cycles,
  atogen and the ledger are bit-identical across pair-kernel workgroup sizes and atogen
  workgroup targets.
* A graph of gensq + stog, replayed after eager work changed its inputs, writes what the
  eager calls write, bit for bit, on every replay (several replays, eager work between).
* An ensemble of boxes of uneven size on several streams gives every box the bits it
  gets when it runs alone."""

import numpy as np
import pytest
import torch

from tests import helpers as H

F32 = np.float32
pytestmark = pytest.mark.gpu


def _small_boxes(n=5):
    from epsr_python_translation_amd import synth
    boxes = synth.config2_boxes(nbox=16)
    boxes.sort(key=lambda s: s.natom)
    return boxes[:n]


def _state(R):
    torch.cuda.synchronize()
    out = {k: v.copy() for k, v in R.box.get_coords().items()}
    out["ergpair"] = R.ledger.export()["ergpair"].copy()
    out["ran1"] = R.ran1.copy()
    out["wav"] = R.wav.cpu().numpy().copy()
    return out


def _same(a, b):
    for k in a:
        assert np.array_equal(np.asarray(a[k]).view(np.int32), np.asarray(b[k]).view(np.int32)), k


def test_launch_shapes_do_not_change_results():
    from epsr_python_translation_amd import _native
    from epsr_python_translation_amd import device as D
    from epsr_python_translation_amd import pipeline as PL
    s = H.system("c0")
    outs = []
    for blk, wg in ((256, 256), (512, 1024), (1024, 4096)):
        print(f"launch shapes pair_block={blk} atogen_workgroups={wg}", flush=True)
        ctx = D.Context(0)
        R = PL.Refinement(ctx, s, idum=-55).build_tables()
        R.box.set_launch(_native.LAUNCH_PAIR_BLOCK, blk)
        R.box.set_launch(_native.LAUNCH_ATOGEN_WORKGROUPS, wg)
        R.cells_and_ledger()
        for _ in range(2):
            R.molfit()
        R.atogen(ngrsamples=s.nmol)
        R.cells_and_ledger()
        outs.append(_state(R))
    _same(outs[0], outs[1])
    _same(outs[0], outs[2])


def test_graph_replay_of_transforms_matches_eager():
    from epsr_python_translation_amd import device as D
    from epsr_python_translation_amd import pipeline as PL
    s = H.system("c0")
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        ctx = D.Context(0)
        gctx = D.Context(0)
        R = PL.Refinement(ctx, s, idum=-56).build_tables()
        R.cells_and_ledger()
        R.atogen(ngrsamples=s.nmol)
        R.gensq(mode=1, ctx=gctx)
        R.stog(mode=1, ctx=gctx)
        torch.cuda.synchronize()
        with D.Graph.capture(gctx) as g:
            R.gensq(mode=1, ctx=gctx)
            R.stog(mode=1, ctx=gctx)
        # the elements the transforms write: Q 1..nq and r 1..nr of every column
        sq = lambda a: a.reshape(2 * s.npairs, R.ldq)[:, 1:R.g.nq + 1]
        gr = lambda a: a.reshape(s.npairs, R.ldg)[:, 1:R.nr + 1]
        for rep in range(3):
            print(f"graph replay {rep}", flush=True)
            R.molfit()
            R.atogen(ngrsamples=s.nmol)
            R.gensq(mode=1, ctx=gctx)
            R.stog(mode=1, ctx=gctx)
            torch.cuda.synchronize()
            sofq, gofr = sq(R.sofq.cpu().numpy()).copy(), gr(R.gofr.cpu().numpy()).copy()
            R.sofq.fill_(float("nan"))
            R.gofr.fill_(float("nan"))
            torch.cuda.synchronize()
            g.launch()
            torch.cuda.synchronize()
            assert np.array_equal(sq(R.sofq.cpu().numpy()).view(np.int32), sofq.view(np.int32))
            assert np.array_equal(gr(R.gofr.cpu().numpy()).view(np.int32), gofr.view(np.int32))
        g.close()


def test_ensemble_boxes_match_boxes_alone():
    from epsr_python_translation_amd import device as D
    from epsr_python_translation_amd import ensemble as E
    from epsr_python_translation_amd import pipeline as PL
    systems = _small_boxes(5)
    print("ensemble of", [s.natom for s in systems], flush=True)
    ens = E.Ensemble(systems, nworkers=3, idum0=-300, tune=True)
    for _ in range(2):
        st = ens.step(sweeps=2, ngrsamples=0)
        print(f"  step: {st}", flush=True)
    torch.cuda.synchronize()
    for i, s in enumerate(systems):
        ctx = D.Context(0)
        R = PL.Refinement(ctx, s, idum=-300 - i).build_tables()
        R.cells_and_ledger()
        ref = ens.boxes[i]
        # replay what the ensemble did to this box: set-up ledger, the tuner's sweeps (on
        # the representative of its size class), then two steps of (ledger, 2 sweeps,
        # atogen) and the transforms
        n_tune = ens.tune_sweeps[i]
        tables = R.tables
        for _ in range(n_tune):
            R.molfit(tables)
        for _ in range(2):
            R.cells_and_ledger()
            tables = R.tables
            for _ in range(2):
                R.molfit(tables)
            R.atogen(ngrsamples=0)
        R.gensq(mode=1)
        R.stog(mode=1)
        _same(_state(R), _state(ref))
        torch.cuda.synchronize()
        assert np.array_equal(R.sofq.cpu().numpy().view(np.int32), ref.sofq.cpu().numpy().view(np.int32))
        assert np.array_equal(R.gofr.cpu().numpy().view(np.int32), ref.gofr.cpu().numpy().view(np.int32))

