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

Prints one JSON line.  --gpus N > 1 is rejected: the library runs on one GPU.
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


def _oracle_setup(s, L):
    """The CPU path end to end on box s: tables (truncation, generate_rp, generate_ep), cells
    and the fmolec1 ledger, all from the oracle."""
    import ctypes as C

    from epsr_python_translation_amd import tables as T
    from oracle import pyoracle as po

    g = T.make_grid(s)
    ob = po.OracleBox(s)
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
    keep = (xavs, xavs1, ecoul, ecoul1)
    tt = po.Tables(ld, g.npt2, po.Pf(g.xstar), po.Pf(xavs), po.Pf(xavs1), po.Pf(ecoul), po.Pf(ecoul1))
    cl = po.Cells()
    L.ora_cells_init(C.byref(cl), s.nmolcell, po.Pf(ob.xadd), po.Pf(ob.yadd), po.Pf(ob.zadd), s.nmol,
                     T.sepmax(s, g))
    led = po.Ledger()
    ft, fb, f64 = C.c_float(), C.c_float(), C.c_double()
    L.ora_fmolec1(C.byref(ob.box), C.byref(ob.centres), C.byref(tt), C.byref(cl),
                  po.Pf(np.ascontiguousarray(s.eintra, F32)), C.byref(led), C.byref(ft), C.byref(fb), C.byref(f64))
    return dict(g=g, ob=ob, tt=tt, cl=cl, led=led, keep=keep)


def _cpu_mc_rate(s, L, seconds_target):
    """Oracle molfit sweeps of box s on one core until seconds_target: (atom-moves/s, sweeps, s)."""
    import ctypes as C

    from oracle import pyoracle as po

    o = _oracle_setup(s, L)
    ob = o["ob"]
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
        L.ora_molfit(C.byref(ob.box), po.Pf(ob.xa), po.Pf(ob.ya), po.Pf(ob.za), C.byref(ob.centres),
                     C.byref(o["tt"]), C.byref(o["cl"]), C.byref(o["led"]), C.byref(mc), C.byref(rng),
                     C.byref(fadds), C.byref(nacc))
        sweeps += 1
        dt = time.perf_counter() - t0
        if dt > seconds_target or sweeps >= 50:
            break
    return sweeps * s.natom / dt, sweeps, dt


def _cpu_pair_rate(s, flavour, norigin_mol):
    """Oracle atogen of box s with norigin_mol origin molecules (ran1-sampled, the reference's
    ngrsamples): (pair distances/s, pairs, s)."""
    import ctypes as C

    from epsr_python_translation_amd import tables as T
    from oracle import pyoracle as po

    g = T.make_grid(s)
    ob = po.OracleBox(s)
    ldw = g.npt1 + 1
    wav = np.zeros(2 * s.npairs * ldw, F32)
    px, pm, cs = np.zeros(s.npairs, F32), np.zeros(s.npairs, F32), np.zeros(s.ntype, F32)
    flags = np.zeros(s.nmol, np.int32)
    rng = po.Ran1.seeded(-9)
    t0 = time.perf_counter()
    po.lib(flavour).ora_atogen(C.byref(ob.box), C.byref(ob.centres), g.npt1, ldw, po.Pf(g.ycoh),
                               po.Pi(np.arange(s.npairs, dtype=np.int32)), po.Pf(np.zeros(s.npairs, F32)),
                               int(norigin_mol), C.byref(rng), po.Pf(wav), po.Pf(px), po.Pf(pm), po.Pf(cs),
                               po.Pi(flags))
    dt = time.perf_counter() - t0
    nat = (np.asarray(s.nal) - np.asarray(s.naf) + 1).astype(np.int64)
    sel = flags != 0
    pairs = int((nat[sel] * (s.natom - nat[sel])).sum())
    nintra = int((nat * (nat - 1) // 2).sum())
    return (pairs + nintra) / dt, pairs + nintra, dt


def cpu_baseline(system, seconds_target=8.0):
    """The oracle (C restatement of the reference's loops, -O2 -fno-tree-vectorize) timed on
    bounded samples of the workloads on this host:
      value   molfit sweeps of the same box on one core, atom-moves/s (the MC chain is
              serial: each move's acceptance depends on the previous ones, so it has no
              all-cores form);
      pair_distances_per_s   atogen of the same box with ran1-sampled origin molecules, on
              one core and on all cores (origin molecules over OpenMP threads);
      dmso_water_3200   molfit sweeps of the 200 DMSO + 400 water box on one core."""
    from epsr_python_translation_amd import synth
    from oracle import pyoracle as po

    s = system
    L = po.lib("glibc")
    rate, sweeps, dt = _cpu_mc_rate(s, L, seconds_target)
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    p1, n1, t1 = _cpu_pair_rate(s, "glibc", 100)
    pn, nn, tn = _cpu_pair_rate(s, "omp", 2000)
    s0 = synth.config0()
    r0, sw0, dt0 = _cpu_mc_rate(s0, L, 3.0)
    return {"value": rate, "unit": "atom-moves/s", "cores": 1, "kind": "port",
            "sample": f"oracle molfit (C restatement of the reference MC cycle, -O2 -fno-tree-vectorize) on the "
                      f"same box: {sweeps} sweep(s) of {s.nmol} molecule moves in {dt:.2f} s; "
                      f"CPU {cpu_model()}, nproc {os.cpu_count()}",
            "pair_distances_per_s": {
                "cores_1": p1, "cores_all": pn, "threads": threads,
                "sample": f"oracle atogen on the same box: 100 origin molecules ({n1:.3g} distances) in {t1:.2f} s "
                          f"on one core; 2000 origin molecules ({nn:.3g} distances) in {tn:.2f} s on {threads} "
                          f"OpenMP threads"},
            "dmso_water_3200": {
                "value": r0, "unit": "atom-moves/s", "cores": 1,
                "sample": f"oracle molfit on 200 DMSO + 400 water (3200 atoms): {sw0} sweep(s) in {dt0:.2f} s"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--sweeps", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-small-box", action="store_true")
    ap.add_argument("--no-da", action="store_true", help="skip the surrogate / delayed-acceptance leg")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the 64-box ensemble leg")
    ap.add_argument("--no-extras", action="store_true", help="skip press and the running averages after the step")
    ap.add_argument("--no-replicas", action="store_true", help="skip the 16-replica SiO2 refinement leg")
    ap.add_argument("--da-samples", type=float, default=2e6, help="training pairs for the surrogate")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first copy 256 MiB device-to-device (the known-size reference for PMC byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the library has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)

    import torch

    from epsr_python_translation_amd import device as D
    from epsr_python_translation_amd import pipeline as PL
    from epsr_python_translation_amd import synth

    if args.calibrate_copy:
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
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
    mc_bytes = []

    families = ("fmolec1_pair", "mc_pair", "mc_resolve", "atogen_inter", "gensq", "stog")

    def step(record):
        # every launch is timed by HIP events on its stream; they are read after the timed
        # steps, so that no read waits on the GPU inside the step
        R.cells_and_ledger()
        # the sweeps back to back on the GPU (one host wait)
        R.molfit(ncycles=args.sweeps)
        if record:
            mc_bytes.append(R.mc.stats.resolver_bytes)
        R.atogen(ngrsamples=s.nmol)
        R.gensq(mode=1)
        D.stog(ctx, s.rho, R.q, R.sofq[s.npairs * R.ldq:], R.ldq, r, gofr, ldg, nqmin, nqmax, nr, s.npairs, xwt,
               mode=1)

    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    for fam in families:   # the warm-up launches' events, read and dropped
        ctx.kernel_ms_sum(fam)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    for fam in families:
        t, n = ctx.kernel_ms_sum(fam)
        if n:
            ktime[fam] = [t / n] * n
    ms_per_step = 1000.0 * elapsed / args.steps
    ktot = {k: float(np.sum(v)) / args.steps for k, v in ktime.items()}
    kavg = {k: float(np.mean(v)) for k, v in ktime.items()}

    # algorithmic work per launch of each kernel family
    npair_entries = R.ledger.nergpair - int(np.count_nonzero(s.eintra))
    apm = float(np.mean(nat))                                     # atoms per molecule (3 for every water)
    lookups_per_sweep = 3.0 * npair_entries * apm * apm          # A for both slots + Bv for the later one
    # the pair kernels are vector-ALU work (no matrix shape): their peak is MI355X's FP32
    # vector rate, 157.3 TFLOP/s, which counts an FMA as two operations; without contraction
    # (the float32 contract of the path) one operation per lane per cycle gives half of it
    work = {
        "atogen_inter": ("valu", OPS_PER_DISTANCE * inter_pairs, "TFLOP/s",
                         f"{OPS_PER_DISTANCE} fp32 ops x {inter_pairs} distances"),
        "mc_pair": ("valu", OPS_PER_LOOKUP * lookups_per_sweep, "TFLOP/s",
                    f"{OPS_PER_LOOKUP} fp32 ops x {lookups_per_sweep:.3g} atom-pair lookups"),
        "gensq": ("mfma", 2.0 * g.nq * g.npt1 * 2 * s.npairs, "TFLOP/s", "2 nq npt1 2npairs flop"),
        "stog": ("mfma", 2.0 * nr * g.nq * s.npairs, "TFLOP/s", "2 nr nq npairs flop"),
        "mc_resolve": ("hbm", float(np.mean(mc_bytes)) if mc_bytes else 0.0, "GB/s",
                       "thresholds + in-block deltas + row offsets + pushed rows + bits, bytes per sweep; "
                       "a serial chain: latency-bound"),
    }
    kernels = {}
    for k, (bound, units, unit, what) in work.items():
        if k not in kavg:
            continue
        rate = units / (kavg[k] * 1e-3)
        achieved = rate / 1e12 if unit == "TFLOP/s" else rate / 1e9
        peak = FP32_PEAK_TFLOPS if unit == "TFLOP/s" else HBM_PEAK_GBS
        kernels[k] = {"bound": bound, "achieved": achieved, "peak": peak, "unit": unit, "frac": achieved / peak,
                      "ms": kavg[k], "work_per_launch": what}
        if bound == "valu":
            kernels[k]["peak_unfused"] = FP32_PEAK_TFLOPS / 2.0
            kernels[k]["frac_unfused"] = achieved / (FP32_PEAK_TFLOPS / 2.0)
    dom = max(ktot, key=ktot.get)
    kd = kernels.get(dom, {})
    traffic = _traffic(dom)
    roof = {"bound": kd.get("bound", "hbm"), "achieved": kd.get("achieved"), "peak": kd.get("peak"),
            "unit": kd.get("unit"), "frac": kd.get("frac"), "traffic": traffic, "kernel": dom,
            "frac_unfused": kd.get("frac_unfused"),
            "kernel_ms": kavg.get(dom), "work_per_launch": kd.get("work_per_launch"), "per_step_kernel_ms": ktot,
            "kernels": kernels}
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
    if not args.no_extras:
        out["iteration_extras"] = iteration_extras(R, s)
    if not args.no_small_box:
        out["dmso_water_3200"] = small_box_rate(D, PL, synth, ctx)
    if not args.no_da:
        out["surrogate_da"] = surrogate_leg(D, PL, synth, ctx, int(args.da_samples))
    if not args.no_ensemble:
        out["ensemble_64"] = ensemble_leg(synth)
    if not args.no_replicas:
        out["replicas_16"] = replicas_leg(ctx)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(s)
    print(json.dumps(out), flush=True)


HBM_PEAK_GBS = 8000.0
FORTRAN_ATOM_MOVES_PER_S = 1.45e4   # derived lower bound for the 3200-atom test case (5 cycles in <= 1.1 s)


def _traffic(kernel):
    """HBM bytes per launch of the kernel on this workload, from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (calibrated against a copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_water20000.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return t.get(kernel, {}).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def iteration_extras(R, s, reps=5):
    """The once-per-iteration steps around the EP refinement on the same water20000 box,
    after the timed steps (not part of `value`): press over the ledger (GPU pair sums, host
    running sums in the reference's order) and sumsq + sumgr (running averages).  Wall
    clock per call, synchronised, median of `reps`."""
    import torch
    torch.cuda.synchronize()
    tp = []
    for _ in range(reps + 1):
        t0 = time.perf_counter()
        e, p, *_ = R.press()
        tp.append(time.perf_counter() - t0)
    ta = []
    R.gensq()
    for n in range(1, reps + 2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        R.accumulate(n)
        torch.cuda.synchronize()
        ta.append(time.perf_counter() - t0)
    npairs, g = s.npairs, R.g
    elems = 2 * npairs * g.nq + 2 * npairs * (g.npt1 - 1)
    ms_a = float(np.median(ta[1:])) * 1e3
    return {"press": {"ms": float(np.median(tp[1:])) * 1e3, "energy_kj_mol": float(e), "pressure_kbar": float(p),
                      "ledger_entries": int(R.ledger.nergpair)},
            "sumsq_sumgr": {"ms": ms_a, "elements": elems, "bytes": 20 * elems,
                            "GB_per_s_wall": 20 * elems / (ms_a * 1e-3) / 1e9}}


def small_box_rate(D, PL, synth, ctx, sweeps=20):
    """MC sweeps of the 3200-atom DMSO/water box (600 molecules): atom-moves/s of molfit
    alone, beside the Fortran's derived rate on the same-size test case."""
    import torch

    s0 = synth.config0()
    R0 = PL.Refinement(ctx, s0, idum=-77).build_tables()
    R0.cells_and_ledger()
    R0.molfit(ncycles=3)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    R0.molfit(ncycles=sweeps)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    rate = sweeps * s0.natom / dt
    return {"workload": "200 DMSO + 400 water, 3200 atoms, L=31.7 A", "sweeps": sweeps,
            "ms_per_sweep": 1000.0 * dt / sweeps, "mc_atom_moves_per_s": rate,
            "x_fortran_derived": rate / FORTRAN_ATOM_MOVES_PER_S}


BF16_PEAK_TFLOPS = 2500.0


def ensemble_leg(synth, sweeps=5, steps=2, workers=8):
    """64 independent DMSO/water boxes (300-6000 molecules, log-uniform; DMSO fraction
    U(0, 0.5); 0.1 atoms/A^3): per step and box the ledger (FP64 energy totals), 5 MC
    sweeps, atogen (1000 origins), gensq and stog.  Boxes run on `workers` streams with
    autotuned launch shapes and the transform stage replayed from HIP graphs; the same
    work on one stream is timed beside it."""
    import torch

    from epsr_python_translation_amd import ensemble as E

    systems = synth.config2_boxes()
    atoms = sum(s.natom for s in systems)
    res = {"workload": "64 boxes, 300-6000 molecules (log-uniform), DMSO fraction U(0,0.5), 0.1 atoms/A^3",
           "boxes": len(systems), "atoms": atoms, "sweeps_per_step": sweeps}
    for nw in (workers, 1):
        t0 = time.perf_counter()
        ens = E.Ensemble(systems, nworkers=nw, idum0=-7000, tune=True)
        setup = time.perf_counter() - t0
        ens.step(sweeps=sweeps)   # warm-up; captures the transform graphs
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        acc = mov = 0
        for _ in range(steps):
            st = ens.step(sweeps=sweeps)
            acc += st["accepted"]
            mov += st["moves"]
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        res[f"workers_{nw}"] = {"setup_s": setup, "ms_per_step": 1e3 * dt,
                                "atom_moves_per_s": sweeps * atoms / dt, "accepted_fraction": acc / max(1, mov)}
        if nw == workers:
            shapes = {}
            for v in ens.shapes.values():
                k = f"atogen_wg={v.get('atogen_workgroups')},pair_block={v.get('pair_block')}"
                shapes[k] = shapes.get(k, 0) + 1
            res["tuned_shapes"] = shapes
            res["energy_totals_fp64_first_boxes"] = [float(x) for x in ens.totals[:4, 0]]
        del ens
    res["speedup_streams"] = res[f"workers_1"]["ms_per_step"] / res[f"workers_{workers}"]["ms_per_step"]
    return res


def replicas_leg(ctx, nrep=16, workers=8, iters=50, sweeps=5):
    """16 replicas of the SiO2 refinement (8000 Si + 16,000 O each, three weighted S(Q)
    targets from another configuration relaxed by 100 MC sweeps, 41-term EP per pair): the
    surrogate is trained once (plain MC on a separate SiO2 box), then per iteration and
    replica the ledger, `sweeps` delayed-acceptance MC cycles, atogen, gensq, stog, the
    running averages, the fit to the data (the program's R-factor) and ptgnsq (FP64
    normal-equation fit, EP update, generate_ep).  One warm-up iteration, then `iters`
    timed: the reference's schedule of 50."""
    import torch

    from epsr_python_translation_amd import _native
    from epsr_python_translation_amd import replicas as RP
    t0 = time.perf_counter()
    wts, data, fr = RP.synthetic_targets(ctx)
    ens = RP.ReplicaRefinement(nrep=nrep, nworkers=workers)
    ens.set_targets(wts, data)
    setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    losses = ens.train_surrogate(ctx)
    t_train = time.perf_counter() - t0

    def one():
        res = ens.iterate(sweeps=sweeps, fit_mode=_native.FIT_NORMAL, da=True)
        return float(np.mean([r[2] for r in res])), float(np.mean([r[0] for r in res]))
    first = one()
    rfd, rfe = [first[0]], [first[1]]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(iters):
        a, b = one()
        rfd.append(a)
        rfe.append(b)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / iters
    atoms = sum(s.natom for s in ens.systems)
    tb = transforms_batch_rate(ens.boxes)
    del ens
    pick = sorted(set([0, 1, 2, 5, 10, 20, 30, 40, len(rfd) - 1]))
    return {"workload": f"{nrep} replicas x SiO2 24,000 atoms, 3 datasets (isotope fractions "
                        f"{', '.join(f'{x:.2f}' for x in fr)}), {sweeps} delayed-acceptance MC sweeps per "
                        "iteration, stog included",
            "setup_s": setup, "surrogate_train_s": t_train, "surrogate_losses": losses,
            "iterations_timed": iters, "ms_per_iteration": 1e3 * dt, "atom_moves_per_s": sweeps * atoms / dt,
            "total_s_timed": iters * dt,
            "rfactor_data_by_iteration": {str(k): rfd[k] for k in pick},
            "rfactor_ep_fit_by_iteration": {str(k): rfe[k] for k in pick},
            "rfactor_definition": "rfactor_data: the program's 'R-factor for data file' (mean over Q of (data - "
                                  "weighted simulated S(Q))^2 on the running averages), mean over datasets and "
                                  "replicas; rfactor_ep_fit: ptgnsq's relative error of the EP fit to the "
                                  "difference",
            "transforms_batched": tb}


def transforms_batch_rate(boxes, reps=5):
    """gensq and stog of the replicas as one FP32 MFMA GEMM each (the iteration's path),
    timed with HIP events on a context of its own, beside the same transforms box by box."""
    import torch

    from epsr_python_translation_amd import device as D
    from epsr_python_translation_amd import pipeline as PL
    tctx = D.Context(0, timing=True)
    R0 = boxes[0]
    g, npairs, nb = R0.g, R0.s.npairs, len(boxes)
    gs, ss = [], []
    for _ in range(reps + 1):
        PL.transforms_batch(tctx, boxes)
        gs.append(tctx.kernel_ms("gensq"))
        ss.append(tctx.kernel_ms("stog"))
    gb, sb = float(np.median(gs[1:])), float(np.median(ss[1:]))
    one_g = one_s = 0.0
    for R in boxes:
        R.gensq(mode=1, ctx=tctx)
        one_g += tctx.kernel_ms("gensq")
        R.stog(mode=1, ctx=tctx)
        one_s += tctx.kernel_ms("stog")
    torch.cuda.synchronize()
    fg = 2.0 * g.nq * g.npt1 * 2 * npairs * nb
    fs = 2.0 * R0.nr * g.nq * npairs * nb
    return {"boxes": nb, "gensq": {"gemm": f"[{g.nq} x {g.npt1}] . [{g.npt1} x {2 * npairs * nb}]", "ms": gb,
                                   "tflops": fg / (gb * 1e-3) / 1e12,
                                   "frac_fp32_mfma": fg / (gb * 1e-3) / 1e12 / FP32_PEAK_TFLOPS,
                                   "box_by_box_ms": one_g},
            "stog": {"gemm": f"[{R0.nr} x {g.nq}] . [{g.nq} x {npairs * nb}]", "ms": sb,
                     "tflops": fs / (sb * 1e-3) / 1e12, "frac_fp32_mfma": fs / (sb * 1e-3) / 1e12 / FP32_PEAK_TFLOPS,
                     "box_by_box_ms": one_s},
            "note": "FP32 MFMA (v_mfma_f32_16x16x4_f32) against MI355X's 157.3 TFLOP/s; each box's outputs equal "
                    "its own transforms bit for bit"}


def surrogate_leg(D, PL, synth, ctx, nsamples, timed_sweeps=20, gr_sweeps=20):
    """Learned Delta-E surrogate + delayed acceptance on 1600 DMSO + 3200 water (25,600
    atoms): plain GPU MC records (descriptor, exact later-partner dE) pairs, the network
    trains on them on the GPU, then accepted moves/s of plain and delayed-acceptance
    cycles are timed with the step sizes frozen, and their intermolecular histograms are
    compared over segments plain | DA | plain against the plain | plain spread."""
    import torch

    from epsr_python_translation_amd import surrogate as SG

    s = synth.config3()
    R = PL.Refinement(ctx, s, idum=-3303).build_tables()
    R.cells_and_ledger()
    tables = R.tables
    for _ in range(10):
        R.molfit(tables)
    R.mc.params.ireset = 1   # step sizes fixed from here on
    spec = SG.DescriptorSpec()
    ldx = spec.width_for(s.npairs)
    nsweep = max(1, int(np.ceil(nsamples / s.nmol)))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    X, T = SG.collect(ctx, R.box, tables, R.ledger, R.mc, R.ran1, spec, nsweep, ldx)
    torch.cuda.synchronize()
    t_collect = time.perf_counter() - t0
    net = SG.Network(ctx, ldx, seed=1)
    tclip = 25.0 * float(s.simt)
    batch = 4096
    epochs = 3
    t0 = time.perf_counter()
    hist = net.train(X, T, ldx, epochs=epochs, batch=batch, lr=2e-3, tclip=tclip)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    train_ms = ctx.kernel_ms("mlp_train")
    tvar = float(T.clamp(-tclip, tclip).var())
    del X, T

    def run(da, n, timed=True):
        acc = mov = rej = 0
        torch.cuda.synchronize()
        t = time.perf_counter()
        fwd, desc, pairs, res = [], [], [], []
        for _ in range(n):
            if da is None:
                R.molfit(tables)
            else:
                da.sweep(ctx, R.box, tables, R.ledger, R.mc, R.ran1)
                fwd.append(ctx.kernel_ms("mlp_forward"))
                desc.append(ctx.kernel_ms("surr_descriptors"))
            pairs.append(ctx.kernel_ms("mc_pair"))
            res.append(ctx.kernel_ms("mc_resolve"))
            acc += R.mc.stats.accepted
            mov += R.mc.stats.moves
            rej += R.mc.stats.stage1_rejected
        torch.cuda.synchronize()
        kms = {"mc_pair": float(np.mean(pairs)), "mc_resolve": float(np.mean(res))}
        if desc:
            kms.update(surr_descriptors=float(np.mean(desc)), mlp_forward=float(np.mean(fwd)))
        return acc, mov, rej, time.perf_counter() - t, fwd, kms

    dam = SG.DelayedAcceptance(net, spec, seed=4242)
    run(None, 2)
    run(dam, 2)
    ap_, mp, _, tp, _, kp = run(None, timed_sweeps)
    ad, md, rd, td, fwd, kd = run(dam, timed_sweeps)

    def hist_segment(da):
        rows = []
        for _ in range(gr_sweeps):
            if da is None:
                R.molfit(tables)
            else:
                da.sweep(ctx, R.box, tables, R.ledger, R.mc, R.ran1)
            R.atogen(ngrsamples=s.nmol)
            rows.append(R.wav.view(2, s.npairs, R.ldw)[1].double().cpu().numpy().copy())
        return np.stack(rows)

    sA = hist_segment(None)
    sB = hist_segment(dam)
    sC = hist_segment(None)
    hA, hB, hC = sA.sum(0), sB.sum(0), sC.sum(0)
    dist = lambda a, b: float(np.sum((a - b) ** 2) / np.sum(((a + b) / 2) ** 2))

    def chi2(a, b):
        """chi^2 per degree of freedom of the difference of two segments' mean intermolecular
        g(r) histograms, each bin's variance of the mean from its sweep-to-sweep spread"""
        va = a.var(0, ddof=1) / a.shape[0]
        vb = b.var(0, ddof=1) / b.shape[0]
        v = va + vb
        ok = v > 0
        d = (a.mean(0) - b.mean(0))[ok]
        return float(np.sum(d * d / v[ok]) / max(1, int(ok.sum()))), int(ok.sum())

    c_da, dof = chi2(sB, np.concatenate([sA, sC]))
    c_pp, _ = chi2(sA, sC)
    plain_rate = ap_ / tp
    da_rate = ad / td
    flops_row = 2.0 * (ldx * 128 + 128 * 128 + 128)
    fwd_ms = float(np.mean(fwd)) if fwd else None
    fwd_tflops = (2 * s.nmol * flops_row / (fwd_ms * 1e-3) / 1e12) if fwd_ms else None
    train_flops = 3.0 * 2 * batch * flops_row
    return {
        "workload": "1600 DMSO + 3200 water, 25,600 atoms, L=63.5 A; step sizes frozen after 10 adapting sweeps",
        "descriptor": {"gaussians_per_type_pair": spec.ng, "rc_A": spec.rc, "width_A": spec.width,
                       "channels": s.npairs, "d": ldx},
        "network": "d -> 128 -> 128 -> 1, SiLU, BF16 weights/activations, FP32 accumulation",
        "training": {"pairs": nsweep * s.nmol, "collect_s": t_collect, "epochs": epochs, "batch": batch,
                     "train_s": t_train, "loss_first_epoch": hist[0], "loss_last_epoch": hist[-1],
                     "clamped_target_variance": tvar, "target_clamp": tclip,
                     "last_step_ms": train_ms,
                     "last_step_tflops_vs_fp32_mfma": (train_flops / (train_ms * 1e-3) / 1e12) if train_ms > 0 else None},
        "plain": {"sweeps": timed_sweeps, "ms_per_sweep": 1e3 * tp / timed_sweeps, "accepted_fraction": ap_ / mp,
                  "accepted_moves_per_s": plain_rate},
        "delayed_acceptance": {"sweeps": timed_sweeps, "ms_per_sweep": 1e3 * td / timed_sweeps,
                               "accepted_fraction": ad / md, "stage1_pass_fraction": 1.0 - rd / md,
                               "accepted_per_exact_evaluation": ad / max(1, md - rd),
                               "accepted_moves_per_s": da_rate,
                               "mlp_forward_ms": fwd_ms, "mlp_forward_tflops": fwd_tflops,
                               "mlp_forward_frac_bf16_peak": (fwd_tflops / BF16_PEAK_TFLOPS) if fwd_tflops else None},
        "da_over_plain_accepted_moves_per_s": da_rate / plain_rate,
        "plain_accepted_per_exact_evaluation": ap_ / mp,
        "histogram_distance": {"plain_vs_da": dist(hB, (hA + hC) / 2), "plain_vs_plain": dist(hA, hC),
                               "definition": "sum (a-b)^2 / sum ((a+b)/2)^2 over the intermolecular pair histograms "
                                             f"summed over {gr_sweeps} sweeps per segment"},
        "gr_chi2": {"da_vs_plain_per_dof": c_da, "plain_vs_plain_per_dof": c_pp, "dof": dof,
                    "definition": "sum over bins of (mean_a - mean_b)^2 / (var_a/n_a + var_b/n_b) / dof, the "
                                  "intermolecular g(r) histograms of segments plain | DA | plain, "
                                  f"{gr_sweeps} sweeps each (DA against both plain segments)"},
        "kernel_ms_per_sweep": {"plain": kp, "delayed_acceptance": kd},
    }


if __name__ == "__main__":
    main()
