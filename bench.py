"""Benchmark of the EPSR refinement hot loop on one MI355X.

Workload (default): the 60,000-atom box of 20,000 rigid 3-site waters (L = 84.3 A,
0.02 A bins to 42 A, 1200 Q points), all inputs resident in HBM.  One step is the core
of one EPSR iteration on the box:
  cells + pair-energy ledger (fmolec1)
  -> 10 MC sweeps (molfit: 200,000 rigid-molecule moves, 600,000 atom-moves)
  -> atogen over every molecule pair (1.8e9 distances)
  -> gensq (FP32 MFMA GEMM 1200 x 2100 x 6) -> stog of the intermolecular partials.
The headline metric is atom-moves/s over the whole step; pair-distances/s over the same
step is reported beside it.  A second, smaller workload (200 DMSO + 400 water, 3200
atoms, one sweep per step) gives the MC rate of the reference's test case.

Prints one JSON line.  --gpus N > 1 is rejected: this build runs on one GPU.
"""

from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

F32 = np.float32
FP32_PEAK_TFLOPS = 157.3
OPS_PER_DISTANCE = 20      # 3 sub + 3 round + 3 cvt + 3 sub + 3 mul + 2 add + sqrt + div + round
OPS_PER_LOOKUP = 30        # distance (18) + 2 x (div, trunc, sub, 2 interp) + scale + add


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def cpu_baseline(system, seconds_target=10.0):
    """The oracle (C restatement of the reference's molfit/en_change, one core) timed on
    a bounded sample of the same workload: MC sweeps of the same box, scaled to
    atom-moves/s."""
    import ctypes as C

    from epsr_python_translation_amd import tables as T
    from oracle import pyoracle as po

    s = system
    g = T.make_grid(s)
    L = po.lib("glibc")
    ob = po.OracleBox(s)
    # tables from the oracle itself (the CPU path end to end)
    ld = g.ld
    wfac = np.array([0.0] + [L.ora_truncate(F32(g.x[j]), F32(s.rminpt), F32(s.rmaxpt)) for j in range(1, g.npt + 1)],
                    F32)
    wfacc = np.array([0.0] + [L.ora_truncatec(F32(g.x[j]), F32(s.rmaxpt)) for j in range(1, g.npt + 1)], F32)
    wfac[g.npt] = wfacc[g.npt] = 0.0
    z = lambda: np.zeros(s.npairs * ld, F32)
    epot, ecoul, epot1, ecoul1, xavs, xavs1, xt, xt1 = (z() for _ in range(8))
    qr = np.zeros(s.npairs, F32)
    L.ora_generate_rp(g.npt, g.npt2, po.Pf(g.x), po.Pf(g.xstar), po.Pf(wfac), po.Pf(wfacc), s.npairs, po.Pf(s.ecor),
                      po.Pf(s.dcor), po.Pf(s.qcor), po.Pf(qr), po.Pf(qr), F32(s.power), F32(1.0), None, None,
                      F32(0.3), F32(0.0), po.Pf(epot), po.Pf(ecoul), po.Pf(epot1), po.Pf(ecoul1), ld)
    pc = np.ascontiguousarray(s.pcoefp.astype(F32))
    ea = C.c_float()
    L.ora_generate_ep(g.npt, g.npt2, po.Pf(g.x), po.Pf(g.xstar), po.Pf(wfac), s.npairs,
                      po.Pi(np.arange(s.npairs, dtype=np.int32)), s.nterms, po.Pf(pc.reshape(-1)), s.nterms + 1,
                      F32(s.pdmin), F32(s.pdstep), F32(s.rmaxpt), F32(s.psigma2), F32(s.rho), F32(0), F32(s.dmax),
                      F32(10), F32(0), F32(0), po.Pf(ecoul), po.Pf(xavs), po.Pf(xavs1), po.Pf(xt), po.Pf(xt1), ld,
                      C.byref(ea))
    tt = po.Tables(ld, g.npt2, po.Pf(g.xstar), po.Pf(xavs), po.Pf(xavs1), po.Pf(ecoul), po.Pf(ecoul1))
    cl = po.Cells()
    L.ora_cells_init(C.byref(cl), s.nmolcell, po.Pf(ob.xadd), po.Pf(ob.yadd), po.Pf(ob.zadd), s.nmol,
                     T.sepmax(s, g))
    led = po.Ledger()
    ft, fb, f64 = C.c_float(), C.c_float(), C.c_double()
    L.ora_fmolec1(C.byref(ob.box), C.byref(ob.centres), C.byref(tt), C.byref(cl),
                  po.Pf(np.ascontiguousarray(s.eintra, F32)), C.byref(led), C.byref(ft), C.byref(fb), C.byref(f64))
    wr = np.full(s.nmoltype, 0.05, F32)
    wt = np.full(s.nmoltype, F32(0.1) / F32(s.dmax), F32)
    nwh = np.zeros(2 * s.nmoltype, np.int32)
    nmv = np.zeros(2 * s.nmoltype, np.int32)
    nsr = np.zeros(s.nmol, np.int32)
    mc = po.MCState(po.Pf(wr), po.Pf(wt), po.Pi(nwh), po.Pi(nmv), po.Pi(nsr), s.simt, F32(1), F32(1), F32(0.75), 0)
    rng = po.Ran1.seeded(-5)
    fadds, nacc = C.c_float(0.0), C.c_int()
    sweeps = 0
    t0 = time.perf_counter()
    while True:
        L.ora_molfit(C.byref(ob.box), po.Pf(ob.xa), po.Pf(ob.ya), po.Pf(ob.za), C.byref(ob.centres), C.byref(tt),
                     C.byref(cl), C.byref(led), C.byref(mc), C.byref(rng), C.byref(fadds), C.byref(nacc))
        sweeps += 1
        dt = time.perf_counter() - t0
        if dt > seconds_target or sweeps >= 50:
            break
    rate = sweeps * s.natom / dt
    return {"value": rate, "unit": "atom-moves/s", "cores": 1, "kind": "port",
            "sample": f"oracle molfit (C restatement of the reference MC cycle, -O2 -fno-tree-vectorize) on the "
                      f"same box: {sweeps} sweep(s) of {s.nmol} molecule moves in {dt:.2f} s; "
                      f"CPU {cpu_model()}, nproc {os.cpu_count()}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--sweeps", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (see DESIGN.md, sharding)",
              file=sys.stderr)
        sys.exit(2)

    import torch

    from epsr_python_translation_amd import device as D
    from epsr_python_translation_amd import pipeline as PL
    from epsr_python_translation_amd import synth

    s = synth.config1()
    ctx = D.Context(0, timing=True)
    R = PL.Refinement(ctx, s, idum=-2024).build_tables()
    g = R.g
    nr = g.npt1
    r = ctx.tensor(g.x[:nr + 1])
    ldg = nr + 1
    gofr = ctx.zeros(s.npairs * ldg)
    nqmin = ctx.tensor(np.full(s.npairs, 2, np.int32))
    nqmax = ctx.tensor(np.full(s.npairs, g.nq, np.int32))
    xwt = ctx.tensor(np.zeros(s.npairs, F32))
    nat = (s.nal - s.naf + 1).astype(np.int64)
    intra_pairs = int((nat * (nat - 1) // 2).sum())
    inter_pairs = (s.natom * (s.natom - 1) // 2) - intra_pairs
    pairs_per_step = inter_pairs + intra_pairs
    moves_per_step = args.sweeps * s.nmol
    atom_moves_per_step = args.sweeps * s.natom
    ktime = {}

    def add(name, v):
        if v >= 0:
            ktime.setdefault(name, []).append(v)

    def step(record):
        R.cells_and_ledger()
        if record:
            add("fmolec1_pair", ctx.kernel_ms("fmolec1_pair"))
        for _ in range(args.sweeps):
            R.molfit()
            if record:
                add("mc_pair", ctx.kernel_ms("mc_pair"))
                add("mc_resolve", ctx.kernel_ms("mc_resolve"))
        R.atogen(ngrsamples=s.nmol)
        if record:
            add("atogen_inter", ctx.kernel_ms("atogen_inter"))
        R.gensq(mode=1)
        if record:
            add("gensq", ctx.kernel_ms("gensq"))
        D.stog(ctx, s.rho, R.q, R.sofq[s.npairs * R.ldq:], R.ldq, r, gofr, ldg, nqmin, nqmax, nr, s.npairs, xwt,
               mode=1)
        if record:
            add("stog", ctx.kernel_ms("stog"))

    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    acc0 = R.mc.stats.accepted
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1000.0 * elapsed / args.steps
    ktot = {k: float(np.sum(v)) / args.steps for k, v in ktime.items()}
    # dominant kernel family per step
    dom = max(ktot, key=ktot.get)
    if dom == "atogen_inter":
        k_ms = float(np.mean(ktime[dom]))
        achieved = OPS_PER_DISTANCE * inter_pairs / (k_ms * 1e-3) / 1e12
        work = f"{OPS_PER_DISTANCE} fp32 ops x {inter_pairs} distances per launch"
    else:
        k_ms = float(np.mean(ktime[dom]))
        achieved = None
        work = "serial Metropolis chain (latency bound)"
    roof = {"bound": "mfma", "kernel": dom, "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": (achieved / FP32_PEAK_TFLOPS) if achieved else None, "traffic": None, "kernel_ms": k_ms,
            "work": work, "per_step_kernel_ms": ktot}
    out = {
        "metric": "atom-moves/s",
        "value": atom_moves_per_step * args.steps / elapsed,
        "unit": "atom-moves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded rigid-water box)",
        "config": {"workload": f"water20000: 60,000 atoms, L=84.3 A; step = fmolec1 + {args.sweeps} MC sweeps + "
                               "atogen over all 1.8e9 pairs + gensq + stog", "atoms": s.natom, "molecules": s.nmol,
                   "sweeps_per_step": args.sweeps, "bins": g.npt1, "q_points": g.nq},
        "pair_distances_per_s": pairs_per_step * args.steps / elapsed,
        "molecule_moves_per_s": moves_per_step * args.steps / elapsed,
        "accepted_fraction": (R.mc.stats.accepted / max(1, R.mc.stats.moves)),
        "roofline": roof,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(s)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
