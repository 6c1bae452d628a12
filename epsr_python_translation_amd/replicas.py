"""Independent refinements of one system on one GPU (the SiO2 workload: 16 replicas of
8000 Si + 16,000 O, three weighted S(Q) targets, a 40-coefficient EP per pair).

Every replica is a full `pipeline.Refinement` with its own seed and ran1 stream; one
iteration of a replica is what the reference's runepsr17 loop does between potential
refinements, restricted to this package's path: the ledger, `sweeps` MC cycles (plain, back
to back on the GPU, or two-stage with a learned surrogate of the energy change), atogen,
gensq, stog, the running averages (sumsq/sumgr), the fit to the data the program reports
(its R-factor per dataset) and ptgnsq (difference from the data, the Poisson-basis fit,
the EP update and generate_ep).  Replicas run on the ensemble's worker streams and threads,
so one replica's serial host steps overlap the others' kernels; a replica's results do not
depend on which worker runs it or on the other replicas.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _native
from . import formats as FM
from . import pipeline as PL
from . import surrogate as SG
from . import synth
from .ensemble import Ensemble

F32 = np.float32

# coherent neutron scattering lengths (fm) for the two elements; the datasets vary the O
# value between two isotopes by an isotope fraction, as a substitution series does
B_SI = 4.1491
B_O_PAIR = (5.803, -3.0)


def target_weights(type_names, fractions):
    """One neutron weights file per dataset (formats.WtsFile): w_ab = b_a b_b / 100 with
    b_O mixed by the dataset's isotope fraction."""
    out = []
    for x in fractions:
        b = {"Si": B_SI, "O": float(x) * B_O_PAIR[0] + (1.0 - float(x)) * B_O_PAIR[1]}
        pairs = []
        for i, a in enumerate(type_names):
            for c in type_names[i:]:
                w = F32(b[a] * b[c] / 100.0)
                pairs.append((FM.fa(a, 3), FM.fa(c, 3), w, w))
        out.append(FM.WtsFile(rhofac=F32(1.0), pairs=pairs))
    return out


def synthetic_targets(ctx, seed=4, target_seed=999, ndata=3, relax_sweeps=100):
    """Isotope fractions U(0, 1) (seeded), the weights they give, and the data: the weighted
    intermolecular S(Q) of another SiO2 configuration (target_seed) after relax_sweeps MC
    sweeps under the reference potential, so the refinement has a structure to move
    towards.  Returns (wts, data [ndata, nq + 1], fractions)."""
    rng = np.random.default_rng(seed)
    fractions = rng.uniform(0.0, 1.0, ndata)
    s = synth.config4(seed=target_seed)
    wts = target_weights(s.type_names, fractions)
    R = PL.Refinement(ctx, s, idum=-target_seed).build_tables()
    R.cells_and_ledger()
    left = relax_sweeps
    while left > 0:
        n = min(left, 10)
        R.molfit(ncycles=n)
        R.cells_and_ledger()   # cells and ledger follow the moved atoms
        left -= n
    R.atogen()
    R.gensq()
    torch.cuda.synchronize()
    ldq, npairs = R.ldq, s.npairs
    soq = R.sofq.cpu().numpy().reshape(2, npairs, ldq)[1]
    st = np.zeros(35, np.int32)
    st[1] = -1
    w = PL.get_wts(None, s, wts, F32(0.9), np.ones(ndata, F32), st, recompute=False)["cwt"]
    data = np.zeros((ndata, ldq), F32)
    for i in range(ndata):
        data[i] = (w[i][:, None].astype(np.float64) * soq.astype(np.float64)).sum(0).astype(F32)
    del R
    return wts, data, fractions


class ReplicaRefinement(Ensemble):
    """nrep replicas of config4 (seeds seed0 .. seed0 + nrep - 1) on `nworkers` workers."""

    def __init__(self, nrep=16, nworkers=8, seed0=40, idum0=-4000, tune=False, feedback=0.9):
        systems = [synth.config4(seed=seed0 + k) for k in range(nrep)]
        super().__init__(systems, nworkers=nworkers, idum0=idum0, tune=tune)
        self.feedback = F32(feedback)
        self.nsum = 0
        self.weights = None
        self.net = None
        self.da = None

    def train_surrogate(self, ctx, nsamples=200_000, train_seed=777, epochs=3, batch=4096, lr=2e-3, seed=1):
        """The learned energy-change surrogate for delayed acceptance: plain GPU MC on a
        separate SiO2 box (train_seed; 10 adapting sweeps first) records (descriptor, exact
        later-partner dE) pairs, the network trains on them on the GPU, and every replica
        gets a two-stage cycle driven by it (stage-1 uniforms seeded per replica).  Returns
        the training losses per epoch."""
        s = synth.config4(seed=train_seed)
        R = PL.Refinement(ctx, s, idum=-train_seed).build_tables()
        R.cells_and_ledger()
        R.molfit(ncycles=10)
        self.spec = SG.DescriptorSpec()
        ldx = self.spec.width_for(s.npairs)
        nsweep = max(1, int(np.ceil(nsamples / s.nmol)))
        X, T = SG.collect(ctx, R.box, R.tables, R.ledger, R.mc, R.ran1, self.spec, nsweep, ldx)
        self.net = SG.Network(ctx, ldx, seed=seed)
        hist = self.net.train(X, T, ldx, epochs=epochs, batch=batch, lr=lr, tclip=25.0 * float(s.simt))
        torch.cuda.synchronize()
        del X, T, R
        self.da = [SG.DelayedAcceptance(self.net, self.spec, seed=1000 + 7919 * i) for i in range(len(self.boxes))]
        return hist

    def set_targets(self, wts, data):
        """Weights (formats.WtsFile per dataset) and data [ndata, nq + 1]; the inverted weights
        matrix is formed once per replica (get_wts with ireset = 1) from its ran1 stream."""
        self.data = np.asarray(data, F32)
        self.weights = []
        for R in self.boxes:
            ctx = R.ctx
            self.weights.append(PL.get_wts(ctx, R.s, wts, self.feedback, np.ones(len(wts), F32), R.ran1))

    def iterate(self, sweeps=1, fit_mode=_native.FIT_NORMAL, da=False, batch_transforms=True):
        """One iteration of every replica; returns per replica (the EP fit's r-factors per
        dataset, chi^2, the fit to the data per dataset).  da: MC cycles with delayed
        acceptance (train_surrogate first), else plain cycles back to back.
        batch_transforms: gensq and stog of all replicas as one GEMM each between the MC
        and the fit phases (the replicas share their grids; results are the same bits as
        each replica's own transforms)."""
        assert self.weights is not None, "set_targets first"
        assert not da or self.da is not None, "train_surrogate first"
        self.nsum += 1
        out = [None] * len(self.boxes)

        def sample(R):
            i = self.boxes.index(R)
            self._ledger(R)
            tables = R.tables
            if da:
                for _ in range(sweeps):
                    self.da[i].sweep(R.ctx, R.box, tables, R.ledger, R.mc, R.ran1)
            else:
                R.molfit(tables, ncycles=sweeps)
            R.atogen()

        def refine(R):
            i = self.boxes.index(R)
            R.accumulate(self.nsum)
            w = self.weights[i]
            rfacd, _ = R.rfactor_data(self.data, w)
            rf, chi = R.ptgnsq(self.data, w["cwt"], w["cwtpot"], fit_mode=fit_mode)
            out[i] = (rf, chi, rfacd)

        if batch_transforms:
            self._parallel(lambda wk: [sample(R) for R in wk.boxes])
            wk0 = self.workers[0]
            wk0.run(lambda: PL.transforms_batch(wk0.ctx, self.boxes))
            self._parallel(lambda wk: [refine(R) for R in wk.boxes])
        else:
            def work(wk):
                for R in wk.boxes:
                    sample(R)
                    R.gensq()
                    R.stog()
                    refine(R)
            self._parallel(work)
        return out
