"""Benchmark: FP32 split-precision direct sum + Leapfrog on one MI355X.

Workload (one "step" = one Leapfrog KDK step whose force evaluation is the
hot path): Plummer sphere, N = 131072 (2^17), G = 1, equal masses, seed 2,
softening 1e-3, FP32 split-precision direct sum with Kahan-compensated tile
sums, FP64 positions/velocities, dt = 1e-3.  Inputs are resident in HBM
before the timed region.

Printed as one JSON line:
  value      directed interactions/s = steps * N (N-1) / wall time
  roofline   dominant kernel (direct_f32): 20 flops per interaction over its
             HIP-event-timed average launch, against the MI355X FP32 vector peak
  accuracy   relative error vs the FP64 GPU direct sum (4096 sampled targets)
  bh         Barnes-Hut (N = 2^20 two Plummer spheres, theta = 0.5, leaf 1):
             particles/s of build + walk and its error vs FP64 direct
  cpu_baseline  the reference C library (oracle/_ref) timed on this host,
             single thread (its pairwise path is serial), on a bounded sample
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
for p in (REPO, os.path.join(REPO, "gravity-simulator_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

N_DIRECT = 1 << 17
N_BH = 1 << 20
EPS = 1e-3
DT = 1e-3
FLOPS_PER_INTERACTION = 20.0
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


def cpu_baseline(x, m, sample_n):
    """Reference acceleration() (pairwise, serial) on the first sample_n particles."""
    from oracle import oracle as O

    xs, ms = np.ascontiguousarray(x[:sample_n]), np.ascontiguousarray(m[:sample_n])
    if O.ref_available():
        kind = "reference"
        fn = lambda: O.ref_acceleration(xs, ms, 1, G=1.0, eps=EPS)  # noqa: E731
    else:
        kind = "port"
        fn = lambda: O.pairwise(xs, ms, G=1.0, eps=EPS)  # noqa: E731
    t0 = time.perf_counter()
    fn()
    dt = time.perf_counter() - t0
    return {"value": sample_n * (sample_n - 1) / dt, "unit": "interactions/s", "cores": 1, "kind": kind,
            "sample": f"one FP64 pairwise force evaluation of the first {sample_n} particles of the "
                      f"N=2^17 Plummer input (eps=1e-3), {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-sample", type=int, default=65536)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-bh", action="store_true")
    ap.add_argument("--autotune", action="store_true",
                    help="run the FP32 shape autotuner first (default: the library's default shape)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "decomposition is described in DESIGN.md", file=sys.stderr)
        sys.exit(2)

    from grav_sim_amd import ic
    from grav_sim_amd import lib as L

    info = L.device_info()
    x, v, m = ic.plummer(N_DIRECT, 2)
    tune = None
    if args.autotune:
        tune = L.autotune_fp32(N_DIRECT)
    copy_gbs = L.copy_calibrate(1 << 28, 5)
    p = L.param(16, eps=EPS)
    sim = L.Sim(x, v, m, 1.0)
    sim.leapfrog_begin(p, DT)
    sim.leapfrog_steps(p, DT, args.warmup)
    L.synchronize()
    L.timing_enable(True)
    L.timing_reset()
    t0 = time.perf_counter()
    sim.leapfrog_steps(p, DT, args.steps)
    L.synchronize()
    wall = time.perf_counter() - t0
    ms, calls = L.timing_query("direct_f32")
    L.timing_enable(False)
    inter = float(N_DIRECT) * (N_DIRECT - 1)
    value = args.steps * inter / wall
    kavg = ms / max(calls, 1) / 1e3
    achieved = FLOPS_PER_INTERACTION * inter / kavg / 1e12
    med, p99, mx = sim.accel_error(p, samples=4096)
    sim.close()

    out = {
        "metric": "interactions/s + BH particles/s at fixed accel error vs FP64; plain vs chosen",
        "value": value,
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "FP32 split-precision direct sum + Leapfrog, Plummer N=131072, seed 2, "
                               "eps=1e-3, dt=1e-3", "N": N_DIRECT, "fp32_shape": tune},
        "roofline": {"bound": "fp32-valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None,
                     "kernel": "direct_f32", "kernel_avg_ms": kavg * 1e3,
                     "flops_per_launch": FLOPS_PER_INTERACTION * inter},
        "accuracy": {"vs": "FP64 GPU direct sum, 4096 sampled targets", "median_rel": med, "p99_rel": p99,
                     "max_rel": mx},
        "hbm_copy_gbs": copy_gbs,
        "device": {"name": info.name.decode(), "arch": info.arch.decode(), "cu": info.cu_count,
                   "clock_khz": info.clock_khz},
    }

    if not args.no_bh:
        xb, vb, mb = ic.two_spheres(N_BH, 4)
        pb = L.param(3, theta=0.5, leaf=1)
        simb = L.Sim(xb, vb, mb, 1.0)
        simb.acceleration(pb, copy=False)
        L.synchronize()
        L.timing_enable(True)
        L.timing_reset()
        reps = 5
        t0 = time.perf_counter()
        for _ in range(reps):
            simb.acceleration(pb, copy=False)
        L.synchronize()
        bwall = (time.perf_counter() - t0) / reps
        wk_ms, wk_calls = L.timing_query("bh_walk")
        bd_ms, bd_calls = L.timing_query("tree_build")
        so_ms, so_calls = L.timing_query("tree_sort")
        L.timing_enable(False)
        nint = L.last_walk_interactions()
        bmed, bp99, bmx = simb.accel_error(pb, samples=4096)
        simb.close()
        walk_s = wk_ms / max(wk_calls, 1) / 1e3
        out["bh"] = {
            "workload": "two Plummer spheres N=2^20 (seed 4), theta=0.5, leaf=1, monopole",
            "plain": {"particles_per_s": N_BH / bwall, "ms_per_eval": bwall * 1e3,
                      "walk_ms": walk_s * 1e3, "build_ms": bd_ms / max(bd_calls, 1),
                      "sort_ms": so_ms / max(so_calls, 1),
                      "interactions": nint,
                      "walk_tflops": FLOPS_PER_INTERACTION * nint / walk_s / 1e12,
                      "median_rel": bmed, "p99_rel": bp99, "max_rel": bmx},
        }

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(x, m, args.cpu_sample)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
