#!/usr/bin/env python3
"""Benchmark: c-proj excess-functional + gradient evaluations per second.

Workload (the reference's 64^3 / mmax = 5 configuration): a 64^3 grid in a
30 A box, mmax = 5, molrotsymorder = 2 (6 x 11 x 5 = 330 orientations, 88
projections, 131 076 (q, -q) pairs), spce n0 = 0.0332891 A^-3 at 298.15 K,
synthetic c(q) for all 2002 nmax = 5 Luc tuples, xi = 1 + 0.1 U(-1, 1).
One step = one Solver::energy_cproj_mrso evaluation (F_exc and df) with xi
already resident in HBM.  value = evaluations/s of the MDFT_MIXED form; the
plain FP64 form is timed beside it, and |dF| and the gradient's relative L2
error of the mixed form are stated against it.  The CPU baseline is the FP64
oracle (C++/OpenMP restatement of the reference) on the host cores.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor figure
N_SIDE, BOX, MMAX, MRSO = 64, 30.0, 5, 2
TEMP, N0 = 298.15, 0.0332891
C_SEED, XI_SEED, XI_AMP = 2024, 12345, 0.1


def stage_bytes(N, no, np_, form):
    """Algorithmic HBM bytes of each stage of one evaluation: every stage reads
    its input once and writes its output once (delta_rho fused into the
    projection load, delta_f into the proj2angl epilogue)."""
    O = N * no * 8  # one FP64 orientation field
    P = N * np_ * (16 if form == "fp64" else 8)  # the projections
    return {
        "angl2proj": O + P,
        "fft_forward": 3 * 2 * P,  # z, y and x passes
        "oz": 2 * P,
        "fft_backward": 3 * 2 * P,
        "proj2angl_delta_f": P + 2 * O,
    }


def timed(solver, xi, df, ffd, steps, warmup):
    import torch

    for _ in range(warmup):
        solver.energy_cproj_async(xi, df, ffd)
    torch.cuda.synchronize()
    solver.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        solver.energy_cproj_async(xi, df, ffd)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms, cnt = solver.stage_ms()
    solver.set_timing(False)
    return (t1 - t0), {k: v / cnt for k, v in ms.items()}


def cpu_baseline(tab, n, L, rho0, pref, evals):
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    from mini_mdft_amd import rng

    threads = min(16, os.cpu_count() or 1)
    O.set_num_threads(threads)
    C = O.Conv(n, L, MMAX, MRSO, tab.tuples, tab.c, tab.dq)
    xi = rng.xi_field(XI_SEED, (n[0] * n[1] * n[2], 330), XI_AMP)
    C.energy_cproj(xi, rho0, pref)  # warm-up
    t0 = time.perf_counter()
    for _ in range(evals):
        ff, df = C.energy_cproj(xi, rho0, pref)
    dt = time.perf_counter() - t0
    return dict(value=evals / dt, unit="evals/s", cores=threads, kind="port",
                sample=f"{evals} full FP64 evaluations at 64^3 mmax=5 after 1 warm-up (oracle, OpenMP "
                       f"{threads} threads, {dt:.1f} s)"), ff, float(np.linalg.norm(df))


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-evals", type=int, default=2, help="timed CPU-baseline evaluations (0: skip)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs on one GPU: the c-proj hot path of one solute is a single-GPU workload here "
              "(multi-GPU sharding is a design note, not a measured configuration)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from mini_mdft_amd import FP64, MIXED, MIXED_X6, luc
    from mini_mdft_amd import mdft as M

    n, L = (N_SIDE,) * 3, (BOX,) * 3
    g = M.SpatialGrid(n, L)
    ag = M.AngularGrid(MMAX, MRSO)
    mp = M.OrientationProjectionMap(ag)
    tab = luc.synthetic_c(MMAX, MRSO, luc.nq_for(luc.qmax_of_grid(n, L), luc.DQ_TIP3P), seed=C_SEED)
    rho0, pref = M.rho0_of(N0, MRSO), M.prefactor_of(TEMP, MRSO, N0)
    N, no, np_ = g.N, ag.m_no, ag.m_np
    print(f"[bench] 64^3 mmax=5: N={N} no={no} np={np_}", file=sys.stderr, flush=True)

    dev = torch.device("cuda:0")
    xi = torch.empty((N, no), dtype=torch.float64, device=dev)
    M.fill_uniform(xi, XI_SEED, 1.0 - XI_AMP, 1.0 + XI_AMP)
    ffd = torch.zeros(1, dtype=torch.float64, device=dev)

    results = {}
    for form, prec in (("mixed", MIXED), ("mixed_x6", MIXED_X6), ("fp64", FP64)):
        conv = M.Convolution.from_table(tab, g, ag, mp, prec)
        solver = M.Solver(conv, rho0, pref, prec)
        df = torch.empty_like(xi)
        ff = solver.energy_cproj(xi, df)
        dt, stages = timed(solver, xi, df, ffd, args.steps, args.warmup)
        results[form] = dict(ff=ff, df=df, sec=dt, stages=stages)
        print(f"[bench] {form}: {args.steps / dt:.1f} evals/s, F={ff:.10g}, stages(ms)="
              f"{ {k: round(v, 4) for k, v in stages.items()} }", file=sys.stderr, flush=True)
        del solver, conv

    mix, f64, x6 = results["mixed"], results["fp64"], results["mixed_x6"]
    dF = mix["ff"] - f64["ff"]
    grad_rel = float(torch.linalg.norm(mix["df"] - f64["df"]) / torch.linalg.norm(f64["df"]))
    dF_x6 = x6["ff"] - f64["ff"]
    grad_rel_x6 = float(torch.linalg.norm(x6["df"] - f64["df"]) / torch.linalg.norm(f64["df"]))
    value = args.steps / mix["sec"]
    fp64_rate = args.steps / f64["sec"]

    by = stage_bytes(N, no, np_, "mixed")
    stages = {}
    for k, ms in mix["stages"].items():
        stages[k] = dict(ms=round(ms, 5), GBs=round(by[k] / (ms * 1e-3) / 1e9, 1))
    dom = max(mix["stages"], key=mix["stages"].get)
    ach = by[dom] / (mix["stages"][dom] * 1e-3) / 1e9
    roofline = dict(kernel=dom, bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(ach / HBM_PEAK_GBS, 4), traffic=None, algorithmic_bytes=by[dom])

    cpu = None
    if args.cpu_evals > 0:
        print("[bench] CPU baseline (FP64 oracle) ...", file=sys.stderr, flush=True)
        cpu, ff_cpu, _ = cpu_baseline(tab, n, L, rho0, pref, args.cpu_evals)
        cpu["fp64_gpu_vs_cpu_rel"] = abs(f64["ff"] - ff_cpu) / abs(ff_cpu)

    line = {
        "metric": "F+grad evals/s at 64^3 mmax=5 (mixed vs FP64) at |dF|<=0.01 kJ/mol; p99 solve",
        "value": round(value, 2),
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * mix["sec"] / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+bf16x3 (f64 reductions)",
        "data": "synthetic",
        "config": {"workload": "c-proj F_exc + gradient, 64^3 grid, 30 A box, mmax=5 (330 orientations, "
                               "88 projections, 131076 q-pairs), spce n0, synthetic c(q) nmax=5",
                   "grid": N_SIDE, "box_A": BOX, "mmax": MMAX, "mrso": MRSO},
        "fp64_evals_per_s": round(fp64_rate, 2),
        "speedup_mixed_vs_fp64": round(value / fp64_rate, 3),
        "F_fp64": f64["ff"],
        "dF_mixed_vs_fp64": dF,
        "grad_rel_l2_mixed_vs_fp64": grad_rel,
        "mixed_x6_evals_per_s": round(args.steps / x6["sec"], 2),
        "dF_mixed_x6_vs_fp64": dF_x6,
        "grad_rel_l2_mixed_x6_vs_fp64": grad_rel_x6,
        "stages_mixed": stages,
        "stages_fp64_ms": {k: round(v, 5) for k, v in f64["stages"].items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
