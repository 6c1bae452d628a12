"""Benchmark of the EPSR refinement hot loop on one MI355X.

Workload (default): the 60,000-atom box of 20,000 rigid 3-site waters (L = 84.3 A,
0.02 A bins to 42 A, 1200 Q points).  One step is one pass of the hot path over the
box, inputs resident in HBM:
  MC sweeps (molfit) -> atogen over every molecule pair (1.8e9 distances)
  -> gensq (FP32 MFMA GEMM) -> stog of the intermolecular partials (FP32 MFMA GEMM).
Prints one JSON line (the driver's contract).  --gpus N > 1 is rejected: this build
runs on one GPU.
"""

from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

F32 = np.float32
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
OPS_PER_DISTANCE = 20      # 3 sub + 3 round + 3 cvt + 3 sub + 3 mul + 2 add + sqrt + div + round


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def cpu_baseline(system, grid, seconds_target=10.0):
    """The oracle (the C restatement of the reference loops, one core) timed on a
    bounded sample of the same workload: atogen with a sampled subset of origin
    molecules, scaled to pair-distances/s."""
    import ctypes as C

    from oracle import pyoracle as po

    s, g = system, grid
    ob = po.OracleBox(s)
    ldw = g.npt1 + 1
    wav = np.zeros(2 * s.npairs * ldw, F32)
    pot = np.zeros(s.npairs, F32)
    potm = np.zeros(s.npairs, F32)
    ir = np.arange(s.npairs, dtype=np.int32)
    rm = np.zeros(s.npairs, F32)
    L = po.lib("glibc")
    ngr = 50
    while True:
        rng = po.Ran1.seeded(-11)
        cs = np.zeros(s.ntype, F32)
        t0 = time.perf_counter()
        n = L.ora_atogen(C.byref(ob.box), C.byref(ob.centres), g.npt1, ldw, po.Pf(g.ycoh), po.Pi(ir), po.Pf(rm), ngr,
                         C.byref(rng), po.Pf(wav), po.Pf(pot), po.Pf(potm), po.Pf(cs), None)
        dt = time.perf_counter() - t0
        if dt > 2.0 or ngr >= s.nmol:
            break
        ngr = min(s.nmol, int(ngr * max(2.0, seconds_target / max(dt, 1e-3))))
    # origins sampled: csampled atoms of each type, each against every atom not in
    # its own molecule, plus all intramolecular pairs
    n_origin_atoms = float(cs.sum())
    nat_per_mol = (s.nal - s.naf + 1)
    pairs = n_origin_atoms * (s.natom - float(np.mean(nat_per_mol))) + float((nat_per_mol * (nat_per_mol - 1) // 2).sum())
    return {"value": pairs / dt, "unit": "pair-distances/s", "cores": 1, "kind": "port",
            "sample": f"oracle atogen on the same 60,000-atom box, {n} sampled origin molecules "
                      f"({pairs:.3g} distances, {dt:.2f} s), CPU {cpu_model()}, nproc {os.cpu_count()}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (see DESIGN.md, sharding)",
              file=sys.stderr)
        sys.exit(2)

    import torch

    from epsr_python_translation_amd import device as D
    from epsr_python_translation_amd import synth, tables

    s = synth.config1()
    g = tables.make_grid(s)
    ctx = D.Context(0, timing=True)
    box = D.DeviceBox.from_system(ctx, s)
    ldw = g.npt1 + 1
    ycoh = ctx.tensor(g.ycoh)
    irefpcof = ctx.tensor(np.arange(s.npairs, dtype=np.int32))
    rminex = ctx.tensor(np.zeros(s.npairs, F32))
    fo, ldf = D.foqols(ctx, g, s.rho, s.rmaxgr)
    ldq = g.nq + 1
    sofq = ctx.zeros(2 * s.npairs * ldq)
    wav = ctx.empty(2 * s.npairs * ldw)
    q = ctx.tensor(g.q)
    nr = g.npt1
    r = ctx.tensor(g.x[:nr + 1])
    ldg = nr + 1
    gofr = ctx.zeros(s.npairs * ldg)
    nqmin = ctx.tensor(np.full(s.npairs, 2, np.int32))
    nqmax = ctx.tensor(np.full(s.npairs, g.nq, np.int32))
    xwt = ctx.tensor(np.zeros(s.npairs, F32))
    state = D.ran1_state_array(-2024)

    # distances per step: every unordered pair of atoms on different molecules, plus
    # the intramolecular pairs
    nat = (s.nal - s.naf + 1).astype(np.int64)
    intra_pairs = int((nat * (nat - 1) // 2).sum())
    inter_pairs = (s.natom * (s.natom - 1) // 2) - intra_pairs
    pairs_per_step = inter_pairs + intra_pairs

    def step():
        D.atogen(ctx, box, g.npt1, ldw, ycoh, irefpcof, rminex, s.nmol, state, wav=wav)
        D.gensq(ctx, g.nq, s.npairs, g.npt1, ldw, wav, ycoh, fo, ldf, sofq, ldq, mode=1)
        D.stog(ctx, s.rho, q, sofq[s.npairs * ldq:], ldq, r, gofr, ldg, nqmin, nqmax, nr, s.npairs, xwt, mode=1)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    atogen_ms = []
    for _ in range(args.steps):
        step()
        atogen_ms.append(ctx.kernel_ms("atogen_inter"))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1000.0 * elapsed / args.steps
    pair_rate = pairs_per_step * args.steps / elapsed
    k_ms = float(np.mean(atogen_ms))
    achieved = OPS_PER_DISTANCE * inter_pairs / (k_ms * 1e-3) / 1e12
    out = {
        "metric": "pair-distances/s",
        "value": pair_rate,
        "unit": "pair-distances/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded rigid-water box)",
        "config": {"workload": "water20000: 60,000 atoms, L=84.3 A, bins 0.02 A to 42 A (2100 x 3 partials), "
                               "1200 Q points, atogen over all 1.8e9 pairs + gensq + stog",
                   "atoms": s.natom, "molecules": s.nmol, "bins": g.npt1, "q_points": g.nq},
        "roofline": {"bound": "mfma", "kernel": "atogen_inter", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None,
                     "kernel_ms": k_ms, "work": f"{OPS_PER_DISTANCE} fp32 ops x {inter_pairs} distances"},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(s, g)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
