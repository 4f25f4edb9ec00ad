"""Benchmark: parcel-steps/s of the full LS-RK4 step on one MI355X, plain vs
batched vs served.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 128] [--no-cpu]
                    [--members 16] [--requests 16] [--forms plain,batched,served]

plain (the headline ``value``): rising thermal, 128x128x128 grid, 8 parcels
per cell (16,777,216 parcels), fixed dt = 0.01, synthetic seeded inputs,
state resident in HBM.  One step = 5 stages of {par2grid, vor2vel,
vorticity_tendency, fused grid2par + dB/dt + push} + reflective BC.

batched: an ensemble of members with grids drawn from {32,48,64,96}^2 x
{32,64}, 8 parcels per cell, each with its own seed, advanced by one batched
step (one launch per kernel over all members, one captured HIP graph)
against the same members stepped one after another with the plain step.

served: requests with grids drawn from {32,64,96,128}^2 x {32,64}, Poisson
arrivals, 5-20 steps each, admitted against a 200 GB HBM budget and
continuously batched; per-tick step latency p50/p99 and parcel-steps/s.

Prints one JSON line (the driver's contract) with the step throughput,
p50/p99 step latency, the dominant kernel's roofline and the CPU baseline
(the FP64 oracle timed on the host cores on a bounded sample).
"""
from __future__ import annotations

import argparse
import ctypes
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

# HBM traffic of k_g2p_tiles per launch on the 128^3 rising thermal, from
# rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, each scaled by what the same
# counter reads for a 1 GiB device-to-device copy (FETCH x 2.0, WRITE x 1.0):
# 6.336e9 B per launch over 16,777,216 parcels, averaged over the five stage
# launches (stages 1-4 also store the next stage's half axes and read
# buoyancy and humidity for the next par2grid's fixed-point maxima).
TRAFFIC_BYTES_PER_PARCEL = 377.7


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(mem, max_threads=16):
    """One LS-RK4 step of the FP64 oracle on this host (bounded sample)."""
    from tests._util import oracle_state

    nthreads = min(max_threads, os.cpu_count() or 1)
    st = oracle_state(mem, nthreads=nthreads)
    t0 = time.perf_counter()
    st.ls_rk_step(0.0, 0.01)
    dt = time.perf_counter() - t0
    return {"value": mem.n / dt, "unit": "parcel-steps/s", "cores": nthreads, "kind": "port",
            "sample": f"1 RK4 step of the same {mem.nx}^3 member ({mem.n} parcels), FP64 oracle, "
                      f"OpenMP {nthreads} threads, {dt:.1f} s"}, dt


def device_copy_calibration(nbytes=1 << 30, reps=3):
    """Plain device-to-device copies of a known size through the HIP runtime
    (the reference kernel that scales the HBM byte counters)."""
    hip = ctypes.CDLL("libamdhip64.so")
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(src, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    log(f"bench: copy calibration: {reps} device copies of {nbytes} bytes")


def ensemble_workloads(n, seed0, sizes_xy, sizes_z):
    from epic_amd import workloads as W

    rng = np.random.default_rng(seed0)
    out = []
    for m in range(n):
        nx, ny = (int(v) for v in rng.choice(sizes_xy, 2))
        nz = int(rng.choice(sizes_z))
        out.append(W.rising_thermal(nx, seed=seed0 + 1 + m, ny=ny, nz=nz))
    return out


def bench_batched(n_members, steps, seed0=20261202):
    from epic_amd.core import Batch, Member

    t0 = time.perf_counter()
    ws = ensemble_workloads(n_members, seed0, (32, 48, 64, 96), (32, 64))
    npar = sum(w.n for w in ws)
    log(f"bench: batched ensemble of {n_members} members, {npar} parcels, generated in "
        f"{time.perf_counter() - t0:.1f} s")
    dt = 0.01
    # plain: every member stepped alone, one after the other
    plain = [Member.from_workload(w) for w in ws]
    ts = [0.0] * n_members
    for i, m in enumerate(plain):
        ts[i] = m.ls_rk_step(ts[i], dt)   # warm-up (plan build + graph capture)
    a = time.perf_counter()
    for _ in range(steps):
        for i, m in enumerate(plain):
            ts[i] = m.ls_rk_step(ts[i], dt)
    plain_s = (time.perf_counter() - a) / steps
    for m in plain:
        m.close()
    # batched: the same members in one batch
    mem = [Member.from_workload(w) for w in ws]
    b = Batch(mem)
    tb = np.zeros(n_members)
    tb, _ = b.step(tb, dt)   # warm-up
    lat = []
    for _ in range(steps):
        a = time.perf_counter()
        tb, st = b.step(tb, dt)
        lat.append(time.perf_counter() - a)
        assert (st == 0).all()
    b.close()
    for m in mem:
        m.close()
    lat = np.array(lat)
    out = {
        "workload": f"{n_members} rising-thermal members, grids from {{32,48,64,96}}^2 x {{32,64}}, "
                    f"8 parcels/cell, seeds {seed0 + 1}..{seed0 + n_members}",
        "members": n_members, "parcels": npar, "steps": steps,
        "plain_parcel_steps_per_s": npar / plain_s,
        "batched_parcel_steps_per_s": npar / float(lat.mean()),
        "ratio": plain_s / float(lat.mean()),
        "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "plain_ms_per_ensemble_step": plain_s * 1e3,
    }
    log(f"bench: batched {out['batched_parcel_steps_per_s']:.3e} vs plain {out['plain_parcel_steps_per_s']:.3e} "
        f"parcel-steps/s (x{out['ratio']:.2f})")
    return out


def bench_served(n_requests, rate, seed0=20261203):
    from epic_amd.serve import Server, poisson_requests

    t0 = time.perf_counter()
    ws = ensemble_workloads(n_requests, seed0, (32, 64, 96, 128), (32, 64))
    log(f"bench: served: {n_requests} requests, {sum(w.n for w in ws)} parcels, generated in "
        f"{time.perf_counter() - t0:.1f} s")
    reqs = poisson_requests(ws, rate=rate, steps=(5, 20), dt=0.01, seed=seed0)
    budget = 200 * 10**9
    _, st = Server(budget, max_members=64).run(reqs)
    out = {
        "workload": f"{n_requests} requests, grids from {{32,64,96,128}}^2 x {{32,64}}, 8 parcels/cell, "
                    f"Poisson arrivals at {rate}/s, 5-20 RK4 steps each, 200 GB HBM budget, continuous batching",
        "requests": n_requests, "rate_per_s": rate,
        "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "ticks": st["ticks"],
        "parcel_steps_per_s": st["parcel_steps_per_s_busy"],
        "parcel_steps_per_s_wall": st["parcel_steps_per_s_wall"],
        "mean_members_per_tick": st["mean_members_per_tick"],
    }
    log(f"bench: served {out['parcel_steps_per_s']:.3e} parcel-steps/s, p50 {out['p50_ms']:.1f} ms, "
        f"p99 {out['p99_ms']:.1f} ms")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--moist", action="store_true",
                    help="the large moist member's statistics (humidity plume, 32 buoyancy blobs) instead of the "
                         "rising thermal; with --size 256: 134M parcels")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=2)
    ap.add_argument("--forms", default="plain,batched,served")
    ap.add_argument("--members", type=int, default=16)
    ap.add_argument("--requests", type=int, default=16)
    ap.add_argument("--rate", type=float, default=8.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run plain device copies of 1 GiB (the reference kernel for scaling HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    forms = set(args.forms.split(","))

    from epic_amd import workloads as W
    from epic_amd.core import Member

    t0 = time.perf_counter()
    gen_log = (lambda m: log(f"bench: generating: {m}")) if args.size > 128 else None
    mem = W.moist_member(args.size, log=gen_log) if args.moist else W.rising_thermal(args.size, log=gen_log)
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
    times = np.array(times) if times else np.full(1, np.nan)
    ms = float(times.mean() * 1e3)
    log(f"bench: {args.steps} timed steps, mean {ms:.2f} ms")

    # per-kernel timing with HIP events on the member's stream (eager launches)
    gm.set_profiling(True)
    gm.timers(reset=True)
    for _ in range(args.profile_steps):
        t = gm.ls_rk_step(t, dt)
    tm = gm.timers(reset=True)
    gm.set_profiling(False)
    push_ms, push_calls = tm["push"]
    phases = {k: round(v[0] / max(1, args.profile_steps), 4) for k, v in tm.items() if v[1]}
    push_avg = push_ms / max(1, push_calls)
    achieved = BYTES_PUSH * mem.n / (push_avg * 1e-3) / 1e9 if push_avg > 0 else 0.0
    # the counter figure belongs to the 128^3 rising thermal it was measured on
    measured = args.size == 128 and not args.moist
    traffic = TRAFFIC_BYTES_PER_PARCEL * mem.n if measured and TRAFFIC_BYTES_PER_PARCEL is not None else None
    roof = {"kernel": "k_g2p_tiles (fused grid2par + dB/dt + RK push)", "bound": "hbm", "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "avg_launch_ms": round(push_avg, 4), "bytes_per_parcel": BYTES_PUSH,
            "units_per_launch": mem.n}
    gm.close()
    if args.copy_calibration:
        device_copy_calibration()
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
        "config": {"workload": f"{'moist member' if args.moist else 'rising thermal'} {mem.nx}x{mem.ny}x{mem.nz}, "
                               f"8 parcels/cell ({mem.n} parcels), "
                               f"plain LS-RK4 step, dt=0.01; FP32 storage, FP64 positions/eigen/z-transforms",
                   "grid": [mem.nx, mem.ny, mem.nz], "parcels": mem.n, "form": "plain"},
        "phases_ms_per_step": phases,
        "roofline": roof,
    }
    if "batched" in forms and args.members > 0:
        line["batched"] = bench_batched(args.members, max(1, min(args.steps, 5)))
    if "served" in forms and args.requests > 0:
        line["served"] = bench_served(args.requests, args.rate)
    if not args.no_cpu:
        cb, cdt = cpu_baseline(mem)
        line["cpu_baseline"] = cb
        line["speedup_vs_cpu"] = round(cdt / (ms * 1e-3), 1)
        log(f"bench: cpu baseline {cb['value']:.3e} parcel-steps/s")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
