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



def _large_boxes(n=3):
    from epsr_python_translation_amd import synth
    boxes = synth.config2_boxes(nbox=64)
    boxes.sort(key=lambda s: s.natom)
    return boxes[-n:]


def test_ensemble_largest_boxes_against_oracle():
    """The three largest boxes of the 64-box draw (about 6,000 molecules each), run together
    on three streams for one step -- ledger, one MC sweep, atogen with 1000 origin
    molecules -- each against the oracle from the same start and on the GPU's own tables:
    the ledger, coordinates, ran1 state, step sizes, counters and histograms, bit for bit."""
    import ctypes as C
    from epsr_python_translation_amd import ensemble as E
    from epsr_python_translation_amd import tables as T
    from oracle import pyoracle as po
    systems = _large_boxes(3)
    print("ensemble largest boxes", [s.natom for s in systems], flush=True)
    ens = E.Ensemble(systems, nworkers=3, idum0=-400, tune=False)
    torch.cuda.synchronize()
    before = []
    for R in ens.boxes:
        tabs = {k: getattr(R, k).cpu().numpy().copy() for k in ("xstar", "xavs", "xavs1", "ecoul", "ecoul1")}
        before.append(dict(ergpair=R.ledger.export()["ergpair"].copy(), tabs=tabs, wr=R.mc.wholerot.copy(),
                           wt=R.mc.wholetrans.copy(), ran1=R.ran1.copy(), rminex=R.rminex.cpu().numpy().copy()))
    st = ens.step(sweeps=1, ngrsamples=0)
    print(f"  step: {st}", flush=True)
    torch.cuda.synchronize()
    Lc, Lo = po.lib("cr"), po.lib("omp")
    for i, (s, R, b) in enumerate(zip(systems, ens.boxes, before)):
        print(f"  oracle box {i}: {s.nmol} molecules, {s.natom} atoms", flush=True)
        g = R.g
        t = b["tabs"]
        tt = po.Tables(R.ld, g.npt2, po.Pf(t["xstar"]), po.Pf(t["xavs"]), po.Pf(t["xavs1"]), po.Pf(t["ecoul"]),
                       po.Pf(t["ecoul1"]))
        ob = po.OracleBox(s)
        cl = po.Cells()
        Lc.ora_cells_init(C.byref(cl), s.nmolcell, po.Pf(ob.xadd), po.Pf(ob.yadd), po.Pf(ob.zadd), s.nmol,
                          T.sepmax(s, g))
        led = po.Ledger()
        ft, fb, f64 = C.c_float(), C.c_float(), C.c_double()
        ei = np.ascontiguousarray(s.eintra, F32)
        Lc.ora_fmolec1(C.byref(ob.box), C.byref(ob.centres), C.byref(tt), C.byref(cl), po.Pf(ei), C.byref(led),
                       C.byref(ft), C.byref(fb), C.byref(f64))
        erg0 = np.ctypeslib.as_array(led.ergpair, shape=(led.nergpair,)).copy()
        assert np.array_equal(erg0.view(np.int32), b["ergpair"].view(np.int32)), f"box {i} ledger"
        wr, wt = b["wr"].copy(), b["wt"].copy()
        nwh = np.zeros(2 * s.nmoltype, np.int32)
        nmv = np.zeros(2 * s.nmoltype, np.int32)
        nsr = np.zeros(s.nmol, np.int32)
        mcs = po.MCState(po.Pf(wr), po.Pf(wt), po.Pi(nwh), po.Pi(nmv), po.Pi(nsr), s.simt, F32(1), F32(1),
                         F32(s.rejrate), 0)
        rng = po.Ran1.from_array(b["ran1"])
        fadds, nacc = C.c_float(0.0), C.c_int()
        Lc.ora_molfit(C.byref(ob.box), po.Pf(ob.xa), po.Pf(ob.ya), po.Pf(ob.za), C.byref(ob.centres), C.byref(tt),
                      C.byref(cl), C.byref(led), C.byref(mcs), C.byref(rng), C.byref(fadds), C.byref(nacc))
        ldw = g.npt1 + 1
        wav = np.zeros(2 * s.npairs * ldw, F32)
        potex, potexmax, cs = np.zeros(s.npairs, F32), np.zeros(s.npairs, F32), np.zeros(s.ntype, F32)
        ngr = Lo.ora_atogen(C.byref(ob.box), C.byref(ob.centres), g.npt1, ldw, po.Pf(g.ycoh),
                            po.Pi(np.arange(s.npairs, dtype=np.int32)), po.Pf(b["rminex"]), 0, C.byref(rng),
                            po.Pf(wav), po.Pf(potex), po.Pf(potexmax), po.Pf(cs), None)
        assert ngr == min(1000, s.nmol)
        co = R.box.get_coords()
        for k in ("xa", "ya", "za", "xadd", "yadd", "zadd"):
            assert np.array_equal(co[k].view(np.int32), getattr(ob, k).view(np.int32)), f"box {i} {k}"
        np.testing.assert_array_equal(R.ran1, rng.as_array())
        assert np.array_equal(R.mc.wholerot.view(np.int32), wr.view(np.int32))
        assert np.array_equal(R.mc.wholetrans.view(np.int32), wt.view(np.int32))
        np.testing.assert_array_equal(R.mc.nwhole, nwh)
        np.testing.assert_array_equal(R.mc.nstepreject.cpu().numpy(), nsr)
        erg1 = np.ctypeslib.as_array(led.ergpair, shape=(led.nergpair,))
        assert np.array_equal(R.ledger.export()["ergpair"].view(np.int32), erg1.view(np.int32)), f"box {i} ledger"
        sl = np.s_[:, :, 1:g.npt1 + 1]
        got = R.wav.cpu().numpy().reshape(2, s.npairs, ldw)[sl]
        assert np.array_equal(got.view(np.int32), wav.reshape(2, s.npairs, ldw)[sl].view(np.int32)), f"box {i} wav"
