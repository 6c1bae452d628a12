"""Independent refinements of one system on one GPU (the SiO2 workload: 16 replicas of
8000 Si + 16,000 O, three weighted S(Q) targets, a 40-coefficient EP per pair).

Every replica is a full `pipeline.Refinement` with its own seed and ran1 stream; one
iteration of a replica is what the reference's runepsr17 loop does between potential
refinements, restricted to this package's path: the ledger, `sweeps` MC cycles, atogen,
gensq, the running averages (sumsq/sumgr) and ptgnsq (difference from the data, the
Poisson-basis fit, the EP update and generate_ep).  Replicas run on the ensemble's worker
streams and threads, so one replica's serial host steps overlap the others' kernels; a
replica's results do not depend on which worker runs it or on the other replicas.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _native
from . import formats as FM
from . import pipeline as PL
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


def synthetic_targets(ctx, seed=4, target_seed=999, ndata=3):
    """Isotope fractions U(0, 1) (seeded), the weights they give, and the data: the weighted
    intermolecular S(Q) of another SiO2 configuration (target_seed), so the refinement has
    a structure to move towards.  Returns (wts, data [ndata, nq + 1], fractions)."""
    rng = np.random.default_rng(seed)
    fractions = rng.uniform(0.0, 1.0, ndata)
    s = synth.config4(seed=target_seed)
    wts = target_weights(s.type_names, fractions)
    R = PL.Refinement(ctx, s, idum=-target_seed).build_tables()
    R.cells_and_ledger()
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

    def set_targets(self, wts, data):
        """Weights (formats.WtsFile per dataset) and data [ndata, nq + 1]; the inverted weights
        matrix is formed once per replica (get_wts with ireset = 1) from its ran1 stream."""
        self.data = np.asarray(data, F32)
        self.weights = []
        for R in self.boxes:
            ctx = R.ctx
            self.weights.append(PL.get_wts(ctx, R.s, wts, self.feedback, np.ones(len(wts), F32), R.ran1))

    def iterate(self, sweeps=1, fit_mode=_native.FIT_NORMAL):
        """One iteration of every replica; returns per replica (r-factors, chi^2) and the
        attempted moves."""
        assert self.weights is not None, "set_targets first"
        self.nsum += 1
        out = [None] * len(self.boxes)

        def work(wk):
            for R in wk.boxes:
                i = self.boxes.index(R)
                self._ledger(R)
                tables = R.tables
                for _ in range(sweeps):
                    R.molfit(tables)
                R.atogen()
                R.gensq()
                R.accumulate(self.nsum)
                w = self.weights[i]
                out[i] = R.ptgnsq(self.data, w["cwt"], w["cwtpot"], fit_mode=fit_mode)
        self._parallel(work)
        return out
