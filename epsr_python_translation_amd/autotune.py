"""Launch-shape autotuner for a box: picks, by timing on the box itself, the resident
workgroups of atogen's persistent pair kernel and the workgroup size of the pair-energy kernels
(epsr_box_set_launch).  Results never depend on these shapes, only the time does.  The
choice is cached per size class (atoms rounded to a quarter power of two, type pairs),
so an ensemble of similar boxes tunes once per class."""

from __future__ import annotations

import math
import time

import numpy as np
import torch

from . import _native

ATOGEN_WORKGROUPS = (512, 768, 1024, 2048)
PAIR_BLOCKS = (256, 512, 1024)

_CACHE: dict = {}


def size_class(system):
    return (round(4 * math.log2(max(2, system.natom))), int(system.npairs))


def _time(fn, repeats):
    ts = []
    for _ in range(repeats):
        torch.cuda.synchronize()
        t = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t)
    return float(np.median(ts))


def tune(R, ngrsamples=None, repeats=3, sweeps=2):
    """Times the candidates on Refinement R and applies the fastest.  atogen runs with
    R's ran1 state restored after every call; the MC candidates are timed on sweeps of
    R's own chain (they count as equilibration).  Returns the chosen shapes."""
    key = size_class(R.s)
    R.tune_sweeps = 0
    if key in _CACHE:
        apply(R, _CACHE[key])
        return _CACHE[key]
    ngr = R.s.nmol if ngrsamples is None else ngrsamples
    st = R.ran1.copy()
    best = {}
    res = []
    for wg in ATOGEN_WORKGROUPS:
        R.box.set_launch(_native.LAUNCH_ATOGEN_WORKGROUPS, wg)

        def run():
            R.ran1[:] = st
            R.atogen(ngrsamples=ngr)
        run()
        res.append((_time(run, repeats), wg))
    R.ran1[:] = st
    best["atogen_workgroups"] = min(res)[1]
    R.box.set_launch(_native.LAUNCH_ATOGEN_WORKGROUPS, best["atogen_workgroups"])
    if R.ledger.built:
        res = []
        tables = R.tables
        for blk in PAIR_BLOCKS:
            R.box.set_launch(_native.LAUNCH_PAIR_BLOCK, blk)
            R.molfit(tables)
            res.append((_time(lambda: [R.molfit(tables) for _ in range(sweeps)], 1), blk))
            R.tune_sweeps += 1 + sweeps
        best["pair_block"] = min(res)[1]
        R.box.set_launch(_native.LAUNCH_PAIR_BLOCK, best["pair_block"])
    _CACHE[key] = best
    return best


def apply(R, shapes):
    if "atogen_workgroups" in shapes:
        R.box.set_launch(_native.LAUNCH_ATOGEN_WORKGROUPS, shapes["atogen_workgroups"])
    if "pair_block" in shapes:
        R.box.set_launch(_native.LAUNCH_PAIR_BLOCK, shapes["pair_block"])
