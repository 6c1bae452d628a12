"""Many independent boxes of uneven size on one GPU.

An :class:`Ensemble` spreads its boxes over workers; each worker owns a HIP stream (and a
library context on it) and a host thread, so the kernels of different boxes run side by
side on the GPU while every box keeps the exact per-box semantics of
:class:`pipeline.Refinement` (the MC chain, atogen and the transforms give the same bits
as when the box runs alone).  Launch shapes come from the autotuner, once per box size
class.  The transform stage of a worker's boxes (gensq and stog, no host reads) is
captured into one HIP graph when the ensemble is built and replayed by every step: one
launch per worker instead of four per box.  A box's MC cycles of a step run back to back in
one call (the step sizes adapt on the device; the host waits once per box and step), and
atogen selects its origins with ran1 on the host, so both run as eager calls.
"""

from __future__ import annotations

import threading
import time

import numpy as np
import torch

from . import autotune
from . import device as D
from . import pipeline as PL


class _Worker:
    def __init__(self, device, timing):
        self.stream = torch.cuda.Stream(device=device)
        with torch.cuda.stream(self.stream):
            self.ctx = D.Context(device, timing=timing)
            # the transform stage runs on a context of its own: its scratch is sized by the
            # first eager pass and never moves, so the captured graph's pointers stay valid
            self.gctx = D.Context(device)
        self.boxes: list[PL.Refinement] = []
        self.graph: D.Graph | None = None

    def run(self, fn):
        with torch.cuda.stream(self.stream):
            fn()


class Ensemble:
    def __init__(self, systems, nworkers=8, device=0, idum0=-101, tune=True, timing=False):
        self.systems = list(systems)
        self.workers = [_Worker(device, timing) for _ in range(max(1, min(nworkers, len(self.systems))))]
        # largest boxes first, each to the least loaded worker
        load = [0] * len(self.workers)
        order = sorted(range(len(self.systems)), key=lambda i: -self.systems[i].natom)
        self.owner = [None] * len(self.systems)
        for i in order:
            w = int(np.argmin(load))
            load[w] += self.systems[i].natom
            self.owner[i] = w
        self.boxes: list[PL.Refinement | None] = [None] * len(self.systems)
        for i, s in enumerate(self.systems):
            wk = self.workers[self.owner[i]]
            with torch.cuda.stream(wk.stream):
                R = PL.Refinement(wk.ctx, s, idum=idum0 - i).build_tables()
            wk.boxes.append(R)
            self.boxes[i] = R
        torch.cuda.synchronize()
        self.totals = np.zeros((len(self.systems), 2))   # FP64 ftest, ftestb per box
        self._parallel(lambda wk: [self._ledger(R) for R in wk.boxes])
        self.shapes = {}
        if tune:
            # one representative per size class, tuned alone on an otherwise idle GPU
            seen = {}
            for R in self.boxes:
                key = autotune.size_class(R.s)
                if key not in seen:
                    wk = self.workers[self.owner[self.boxes.index(R)]]
                    wk.run(lambda R=R: autotune.tune(R, sweeps=2))
                    seen[key] = autotune._CACHE[key]
            for R in self.boxes:
                autotune.apply(R, seen[autotune.size_class(R.s)])
                self.shapes[R.s.name] = seen[autotune.size_class(R.s)]
            torch.cuda.synchronize()
        # MC cycles the tuner ran on each box (its size class's representative)
        self.tune_sweeps = [getattr(R, "tune_sweeps", 0) for R in self.boxes]
        # the transform graphs: an eager pass sizes every graph context's scratch, then the
        # workers' passes are captured one at a time, with no other work in flight (a call
        # that synchronises the device from another thread would invalidate a capture)
        self._parallel(self._transforms)
        for wk in self.workers:
            wk.gctx.check(wk.gctx.lib.epsr_ctx_synchronize(wk.gctx.h))
            with torch.cuda.stream(wk.stream):
                with D.Graph.capture(wk.gctx) as g:
                    self._transforms(wk)
            wk.graph = g
        torch.cuda.synchronize()

    def _ledger(self, R):
        f64, fb64 = R.cells_and_ledger()
        self.totals[self.boxes.index(R)] = (f64, fb64)

    def _parallel(self, fn):
        errors = []

        def body(wk):
            try:
                wk.run(lambda: fn(wk))
                wk.stream.synchronize()
            except Exception as e:   # re-raised in the caller's thread
                errors.append(e)
        threads = [threading.Thread(target=body, args=(wk,)) for wk in self.workers]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        if errors:
            raise errors[0]

    def _transforms(self, wk):
        for R in wk.boxes:
            R.gensq(mode=1, ctx=wk.gctx)
            R.stog(mode=1, ctx=wk.gctx)

    def step(self, sweeps=5, ngrsamples=0):
        """One step of every box: the ledger (FP64 energy totals), `sweeps` MC cycles,
        atogen (ngrsamples as the reference: <= 0 means 1000 origins), gensq and stog."""
        stats = {}

        def work(wk):
            acc = mov = 0
            for R in wk.boxes:
                self._ledger(R)
                R.molfit(R.tables, ncycles=sweeps)
                acc += R.mc.stats.accepted_total
                mov += R.mc.stats.moves * R.mc.stats.cycles
                R.atogen(ngrsamples=ngrsamples)
            wk.graph.launch()
            stats[id(wk)] = (acc, mov)
        t = time.perf_counter()
        self._parallel(work)
        dt = time.perf_counter() - t
        acc = sum(a for a, _ in stats.values())
        mov = sum(m for _, m in stats.values())
        atoms = sum(s.natom for s in self.systems)
        return {"seconds": dt, "accepted": acc, "moves": mov, "atom_moves": sweeps * atoms}
