"""Benchmark: parcel-steps/s of the full LS-RK4 step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 128] [--no-cpu]

Workload: rising thermal, 128x128x128 grid, 8 parcels per cell
(16,777,216 parcels), fixed dt = 0.01, synthetic seeded inputs, state
resident in HBM.  One step = 5 stages of {par2grid, vor2vel,
vorticity_tendency, fused grid2par + dB/dt + push} + reflective BC.

Prints one JSON line (the driver's contract) with the step throughput,
p50/p99 step latency, the dominant kernel's roofline and the CPU baseline
(the FP64 oracle timed on the host cores on a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "parcel-steps/s per GPU (RK4 step) + p50/p99 step latency, batched vs plain"

# algorithmic bytes per parcel per stage (compulsory traffic of a minimal
# fused implementation): par2grid reads pos 24 + B 20 + V 4 + omega 12 + b 4
# + h 4; the fused grid2par + push reads 104 and writes 100.
BYTES_P2G = 68
BYTES_PUSH = 204


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(mem, max_threads=16):
    """One LS-RK4 step of the FP64 oracle on this host (bounded sample)."""
    from oracle import oracle as o
    from tests._util import oracle_state

    nthreads = min(max_threads, os.cpu_count() or 1)
    st = oracle_state(mem, nthreads=nthreads)
    t0 = time.perf_counter()
    st.ls_rk_step(0.0, 0.01)
    dt = time.perf_counter() - t0
    return {"value": mem.n / dt, "unit": "parcel-steps/s", "cores": nthreads, "kind": "port",
            "sample": f"1 RK4 step of the same {mem.nx}^3 member ({mem.n} parcels), FP64 oracle, "
                      f"OpenMP {nthreads} threads, {dt:.1f} s"}, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=2)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain device copy of 1 GiB (the reference kernel for scaling HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    from epic_amd import workloads as W
    from epic_amd.core import Member

    t0 = time.perf_counter()
    mem = W.rising_thermal(args.size)
    log(f"bench: generated {mem.name} with {mem.n} parcels in {time.perf_counter() - t0:.1f} s")
    gm = Member.from_workload(mem)
    dt = 0.01
    t = 0.0
    for _ in range(args.warmup):
        t = gm.ls_rk_step(t, dt)
    log(f"bench: {args.warmup} warm-up steps done")
    times = []
    for _ in range(args.steps):
        a = time.perf_counter()
        t = gm.ls_rk_step(t, dt)   # returns after the step's work has completed
        times.append(time.perf_counter() - a)
    times = np.array(times)
    ms = float(times.mean() * 1e3)
    log(f"bench: {args.steps} timed steps, mean {ms:.2f} ms")

    # per-kernel timing with HIP events on the member's stream
    gm.set_profiling(True)
    gm.timers(reset=True)
    for _ in range(args.profile_steps):
        t = gm.ls_rk_step(t, dt)
    tm = gm.timers(reset=True)
    gm.set_profiling(False)
    push_ms, push_calls = tm["push"]
    p2g_ms, p2g_calls = tm["par2grid"]
    phases = {k: round(v[0] / max(1, args.profile_steps), 4) for k, v in tm.items() if v[1]}
    push_avg = push_ms / max(1, push_calls)
    achieved = BYTES_PUSH * mem.n / (push_avg * 1e-3) / 1e9 if push_avg > 0 else 0.0
    roof = {"kernel": "k_g2p_tiles (fused grid2par + dB/dt + RK push)", "bound": "hbm", "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "avg_launch_ms": round(push_avg, 4), "bytes_per_parcel": BYTES_PUSH}
    line = {
        "metric": METRIC,
        "value": float(mem.n / (ms * 1e-3)),
        "unit": "parcel-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "p50_ms": float(np.percentile(times, 50) * 1e3),
        "p99_ms": float(np.percentile(times, 99) * 1e3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"rising thermal {mem.nx}x{mem.ny}x{mem.nz}, 8 parcels/cell ({mem.n} parcels), "
                               f"plain LS-RK4 step, dt=0.01; FP32 storage, FP64 positions/eigen/z-transforms",
                   "grid": [mem.nx, mem.ny, mem.nz], "parcels": mem.n, "form": "plain"},
        "phases_ms_per_step": phases,
        "roofline": roof,
    }
    if args.copy_calibration:
        import torch

        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("bench: copy calibration: 3 copies of 1 GiB")
    if not args.no_cpu:
        cb, cdt = cpu_baseline(mem)
        line["cpu_baseline"] = cb
        line["speedup_vs_cpu"] = round(cdt / (ms * 1e-3), 1)
        log(f"bench: cpu baseline {cb['value']:.3e} parcel-steps/s")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
