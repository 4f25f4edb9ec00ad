"""Benchmark: parcel-steps/s of the full LS-RK4 step on one MI355X, plain vs
batched vs served.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 128] [--no-cpu] [--eager]
                    [--forms plain,batched,served] [--members 64] [--ens-steps 20]
                    [--requests 12] [--rates 1,8,64] [--req-steps 50,500]

plain (the headline ``value``): rising thermal, 128x128x128 grid, 8 parcels
per cell (16,777,216 parcels), fixed dt = 0.01, synthetic seeded inputs,
state resident in HBM.  One step = 5 stages of {par2grid, vor2vel,
vorticity_tendency, fused grid2par + dB/dt + push} + reflective BC.  Step
1 from the generated state (its result gives the error table), W warm-up
steps, then K timed steps: replays of the captured step graph, captured
with an event pair around every phase, so the per-kernel times of the
roofline are those of the timed steps themselves (device wall-clock stamps
around every phase inside the graph; --eager: the timed steps launch
eagerly between HIP event pairs, as the counter runs under profiles/ do).

batched: 64 members with grids drawn from {32,48,64,96}^2 x {32,64}, 8
parcels per cell, each with its own seed, tile z chunks from the autotuner
(once per member shape), 20 steps of one batched step (one launch per
kernel over all members, one captured HIP graph) against the same members
stepped one after another with the plain step; each member's max-relative
error after its first batched step against the FP64 oracle.

served: requests with grids drawn from {32,64,96,128}^2 x {32,64}, 50-500
steps each, Poisson arrivals at each rate of the sweep, admitted against a
200 GB HBM budget and continuously batched, each member merging and
splitting its parcels after every step (EPIC's run loop); per-tick step
latency p50/p99 and parcel-steps/s (busy and wall clock).

cpu_baseline: the FP64 oracle (a C restatement of the reference's algorithm)
on the host cores: step 1 of the same member is the warm-up and gives the
GPU's first step its per-field error table, step 2 is timed.

Prints one JSON line (the driver's contract).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import re
import sys
import time
import traceback

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
# HBM traffic per launch from rocprofv3 FETCH_SIZE / WRITE_SIZE counter runs
# on the 128^3 rising thermal (the first step from the generated state, the
# same step this bench's event timing covers), scaled by the plain-copy
# factors recorded in the same file
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "hbm_traffic_rising_thermal128.txt")

DEFAULT_SEED = 20261202
# the served runs take EPIC's run-loop parcel operations (merge, then split)
# after every step of a member, as its run loop does: long runs need the
# splits (with them every 10 steps, ellipsoids outgrow their neighbour cells
# after about 300 steps)
PARCEL_OPS_EVERY = 1


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_threads():
    """The host cores this process may use: its CPU affinity, capped by
    OMP_NUM_THREADS where set (a GPU box grants each GPU a share)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def traffic_from_summary(kernel, path=TRAFFIC_SUMMARY):
    """Bytes per launch of `kernel` from the committed counter summary
    (lines "# traffic <kernel> <bytes per launch> <bytes per parcel>")."""
    try:
        with open(path) as fh:
            for line in fh:
                m = re.match(r"#\s*traffic\s+(\S+)\s+([0-9.eE+]+)\s+([0-9.eE+]+)", line)
                if m and m.group(1) == kernel:
                    return float(m.group(2)), float(m.group(3))
    except OSError:
        pass
    return None, None


def maxrel(a, ref):
    a = np.asarray(a, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    den = np.abs(ref).max()
    return float(np.abs(a - ref).max() / den) if den > 0 else float(np.abs(a).max())


FIELDS = ("volg", "vortg", "tbuoyg", "dbuoyg", "humg", "velog", "velgradg", "vtend")
ATTRS = ("vorticity", "B", "volume", "buoyancy", "humidity")


def gpu_snapshot(gm):
    """Parcels and the last stage's fields of a member (host FP64)."""
    snap = {"parcels": gm.get_parcels()}
    for f in FIELDS + ("nparg", "nsparg"):
        snap[f] = gm.get_field(f)
    return snap


def error_table(mem, snap, st):
    """max|gpu - fp64| / max|fp64| per field and attribute after the same
    step from the same inputs (positions relative to the cell size)."""
    nx, ny = mem.nx, mem.ny
    sl = (Ellipsis, slice(1, nx + 1), slice(1, ny + 1), slice(None))
    err = {}
    p = snap["parcels"]
    err["position/dx"] = float(np.abs(p["position"] - st.position).max() / max(mem.dx))
    for k in ATTRS:
        err[k] = maxrel(p[k], getattr(st, k))
    for f in FIELDS:
        err[f] = maxrel(snap[f][sl], getattr(st, f)[sl])
    # integer count differences after the step (positions differ from the
    # FP64 ones by ~1e-8 dx, so a parcel that close to a cell face may count
    # in the neighbouring cell)
    diff = int(max(np.abs(snap[f][sl].astype(np.int64) - getattr(st, f)[sl].astype(np.int64)).sum()
                   for f in ("nparg",)))
    return {k: float(f"{v:.3e}") for k, v in err.items()}, diff


def counts_match_host(mem, gm):
    """nparg / nsparg of a par2grid against a host count of the same
    parcels with the reference's get_index (fields.f90:165-172) and the
    count folds of parcel_interpl.f90:201-205, 239-243: True when equal."""
    p = gm.get_parcels()
    nx, ny, nz = mem.nx, mem.ny, mem.nz
    lo = np.asarray(mem.lower, dtype=np.float64)
    dxi = 1.0 / np.asarray(mem.dx, dtype=np.float64)
    ijk = np.floor((p["position"] - lo) * dxi).astype(np.int64)
    i, j, k = np.mod(ijk[:, 0], nx), np.mod(ijk[:, 1], ny), ijk[:, 2]
    if (k < -1).any() or (k > nz).any():
        return False
    key = ((k + 1) * ny + j) * nx + i
    vmin = mem.vcell / 20.0
    cnt = np.bincount(key, minlength=(nz + 2) * ny * nx).reshape(nz + 2, ny, nx)
    scnt = np.bincount(key[p["volume"] <= vmin], minlength=(nz + 2) * ny * nx).reshape(nz + 2, ny, nx)
    ok = True
    for f, c in (("nparg", cnt), ("nsparg", scnt)):
        want = np.zeros((nz + 3, ny, nx), dtype=np.int64)   # planes -1..nz+1
        want[:nz + 2] = c
        want[1] += c[0]        # plane 0 takes plane -1
        want[nz] += c[nz + 1]  # plane nz-1 takes plane nz
        a = gm.get_field(f)   # (nx+3, ny+3, nz+3), [ix+1, iy+1, iz+1]
        got = a.reshape(a.shape[-3:])[1:nx + 1, 1:ny + 1, :].transpose(2, 1, 0)
        ok = ok and np.array_equal(got.astype(np.int64), want)
    return bool(ok)


def cpu_baseline(mem, snap):
    """Two LS-RK4 steps of the FP64 oracle on this host: step 1 (warm-up)
    gives the GPU's first step its error table, step 2 is timed."""
    from tests._util import oracle_state

    nthreads = host_threads()
    st = oracle_state(mem, nthreads=nthreads)
    log(f"bench: cpu baseline: oracle state of {mem.n} parcels, {nthreads} threads")
    t0 = time.perf_counter()
    st.ls_rk_step(0.0, 0.01)
    t_warm = time.perf_counter() - t0
    log(f"bench: cpu baseline: step 1 (warm-up) in {t_warm:.1f} s")
    errors, exact = error_table(mem, snap, st) if snap is not None else (None, None)
    t0 = time.perf_counter()
    st.ls_rk_step(0.01, 0.01)
    dt = time.perf_counter() - t0
    cb = {"value": mem.n / dt, "unit": "parcel-steps/s", "cores": nthreads, "kind": "port",
          "nproc": os.cpu_count(), "cpu_model": cpu_model(),
          "sample": f"step 2 of the same {mem.nx}x{mem.ny}x{mem.nz} member ({mem.n} parcels), FP64 oracle "
                    f"(C, OpenMP {nthreads} threads, static schedule), {dt:.1f} s; step 1 ({t_warm:.1f} s) "
                    f"was the warm-up"}
    return cb, dt, errors, exact


def device_copy_calibration(nbytes=1 << 30, reps=3):
    """Plain device-to-device copies of a known size through the HIP runtime
    (the reference kernel that scales the HBM byte counters)."""
    hip = ctypes.CDLL("libamdhip64.so")
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    hip.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
    hip.hipMemset.argtypes = [vp, ctypes.c_int, sz]
    hip.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
    hip.hipFree.argtypes = [vp]
    src, dst = vp(), vp()
    assert hip.hipMalloc(ctypes.byref(src), nbytes) == 0
    assert hip.hipMalloc(ctypes.byref(dst), nbytes) == 0
    assert hip.hipMemset(src, 1, nbytes) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(dst, src, nbytes, 3) == 0   # hipMemcpyDeviceToDevice
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


def bench_batched(n_members, steps, errors=True, seed0=DEFAULT_SEED):
    from epic_amd.core import Batch, Member
    from tests._util import oracle_state

    t0 = time.perf_counter()
    ws = ensemble_workloads(n_members, seed0, (32, 48, 64, 96), (32, 64))
    npar = sum(w.n for w in ws)
    log(f"bench: batched ensemble of {n_members} members, {npar} parcels, generated in "
        f"{time.perf_counter() - t0:.1f} s")
    dt = 0.01
    # tile z chunks from the autotuner, once per grid and parcel count, for
    # both the plain and the batched members
    t1 = time.perf_counter()
    tiles = {}
    for w in ws:
        key = (w.nx, w.ny, w.nz, w.n)
        if key not in tiles:
            m = Member.from_workload(w)
            tiles[key] = m.autotune()
            m.close()
    log(f"bench: batched: autotuned {len(tiles)} member shapes in {time.perf_counter() - t1:.1f} s")

    def tuned(w):
        m = Member.from_workload(w)
        m.set_tiling(*tiles[(w.nx, w.ny, w.nz, w.n)])
        return m

    # plain: every member stepped alone, one after the other
    plain = [tuned(w) for w in ws]
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
    log(f"bench: batched: plain reference {plain_s * 1e3:.1f} ms per ensemble step")
    # batched: the same members in one batch
    mem = [tuned(w) for w in ws]
    b = Batch(mem)
    tb = np.zeros(n_members)
    tb, st = b.step(tb, dt)   # warm-up, step 1
    assert (st == 0).all(), st
    member_err = None
    if errors:
        t1 = time.perf_counter()
        member_err = []
        for w, m in zip(ws, mem):
            got = m.get_parcels()
            o = oracle_state(w, nthreads=host_threads())
            o.ls_rk_step(0.0, dt)
            e = max([float(np.abs(got["position"] - o.position).max() / max(w.dx))] +
                    [maxrel(got[k], getattr(o, k)) for k in ATTRS])
            member_err.append(float(f"{e:.3e}"))
            del o
        log(f"bench: batched: per-member errors after step 1 in {time.perf_counter() - t1:.1f} s, "
            f"max {max(member_err):.2e}")
    lat = []
    for _ in range(steps):
        a = time.perf_counter()
        tb, st = b.step(tb, dt)
        lat.append(time.perf_counter() - a)
        assert (st == 0).all(), st
    b.close()
    for m in mem:
        m.close()
    lat = np.array(lat)
    out = {
        "workload": f"{n_members} rising-thermal members, grids from {{32,48,64,96}}^2 x {{32,64}}, "
                    f"8 parcels/cell, seeds {seed0 + 1}..{seed0 + n_members}, {steps} steps, "
                    f"tile z chunks from the autotuner",
        "members": n_members, "parcels": npar, "steps": steps,
        "tilings": {f"{k[0]}x{k[1]}x{k[2]}": list(v) for k, v in sorted(tiles.items())},
        "plain_parcel_steps_per_s": npar / plain_s,
        "batched_parcel_steps_per_s": npar / float(lat.mean()),
        "ratio": plain_s / float(lat.mean()),
        "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "plain_ms_per_ensemble_step": plain_s * 1e3,
    }
    if member_err is not None:
        out["max_rel_error_vs_fp64_step1"] = {"per_member": member_err, "max": max(member_err),
                                              "fields": "position/dx, " + ", ".join(ATTRS)}
    log(f"bench: batched {out['batched_parcel_steps_per_s']:.3e} vs plain {out['plain_parcel_steps_per_s']:.3e} "
        f"parcel-steps/s (x{out['ratio']:.2f})")
    return out


def bench_served(n_requests, rates, req_steps, seed0=DEFAULT_SEED + 1):
    from epic_amd.serve import Server, poisson_requests

    t0 = time.perf_counter()
    ws = ensemble_workloads(n_requests, seed0, (32, 64, 96, 128), (32, 64))
    log(f"bench: served: {n_requests} requests, {sum(w.n for w in ws)} parcels, generated in "
        f"{time.perf_counter() - t0:.1f} s")
    budget = 200 * 10**9
    sweep = []
    for rate in rates:
        reqs = poisson_requests(ws, rate=rate, steps=req_steps, dt=0.01, seed=seed0)
        results, st = Server(budget, max_members=64, parcel_ops_every=PARCEL_OPS_EVERY).run(reqs)
        failed = sum(1 for r in results.values() if r.status != 0)
        failures = [f"request {r.rid} after {r.steps_done} steps: {r.message}" for r in results.values() if r.status != 0]
        sweep.append({"rate_per_s": rate, "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "ticks": st["ticks"],
                      "parcel_steps": st["parcel_steps"],
                      "parcel_steps_per_s_busy": st["parcel_steps_per_s_busy"],
                      "parcel_steps_per_s_wall": st["parcel_steps_per_s_wall"],
                      "wall_s": st["wall_s"], "mean_members_per_tick": st["mean_members_per_tick"],
                      "parcel_ops_ms_p50": st["ops_ms_p50"], "parcel_ops_ms_mean": st["ops_ms_mean"],
                      "queue_wait_s_p50": st["queue_wait_s_p50"], "failed": failed, "failures": failures[:4]})
        log(f"bench: served at {rate}/s: {st['parcel_steps_per_s_busy']:.3e} parcel-steps/s busy, "
            f"{st['parcel_steps_per_s_wall']:.3e} wall, p50 {st['p50_ms']:.1f} ms, p99 {st['p99_ms']:.1f} ms, "
            f"{st['mean_members_per_tick']:.1f} members per tick, parcel operations {st['ops_ms_mean']:.1f} ms per tick")
    top = max(sweep, key=lambda s: s["rate_per_s"])
    return {
        "workload": f"{n_requests} requests, grids from {{32,64,96,128}}^2 x {{32,64}}, 8 parcels/cell, "
                    f"{req_steps[0]}-{req_steps[1]} RK4 steps each, Poisson arrivals, 200 GB HBM budget, "
                    f"continuous batching, parcel merge + split after every {PARCEL_OPS_EVERY} step(s) of a member",
        "requests": n_requests, "rates_per_s": list(rates),
        "p50_ms": top["p50_ms"], "p99_ms": top["p99_ms"],
        "parcel_steps_per_s": top["parcel_steps_per_s_busy"],
        "parcel_steps_per_s_wall": top["parcel_steps_per_s_wall"],
        "sweep": sweep,
    }


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
    ap.add_argument("--eager", action="store_true",
                    help="the timed steps launch their kernels eagerly between event pairs instead of replaying the "
                         "captured step graph (for counter collection, which is not run on graph replays)")
    ap.add_argument("--forms", default="plain,batched,served")
    ap.add_argument("--members", type=int, default=64)
    ap.add_argument("--ens-steps", type=int, default=20)
    ap.add_argument("--no-member-errors", action="store_true")
    ap.add_argument("--requests", type=int, default=12)
    ap.add_argument("--rates", default="1,8,64")
    ap.add_argument("--req-steps", default="50,500")
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
    want_err = not args.no_cpu

    # step 1 from the generated state: its result gives the error table
    # against the FP64 oracle's step 1
    t = gm.ls_rk_step(t, dt)
    snap = gpu_snapshot(gm) if want_err else None
    for _ in range(args.warmup):
        t = gm.ls_rk_step(t, dt)
    log(f"bench: step 1 and {args.warmup} warm-up steps done")
    # the timed steps: graph replays whose capture holds device wall-clock
    # stamps around every phase (or eager launches between event pairs with
    # --eager), so the per-kernel times come from these very steps
    gm.set_profiling(True, graph=not args.eager)
    gm.timers(reset=True)
    times = []
    for _ in range(args.steps):
        a = time.perf_counter()
        t = gm.ls_rk_step(t, dt)   # returns after the step's work has completed
        times.append(time.perf_counter() - a)
    tm = gm.timers(reset=True)
    gm.set_profiling(False)
    times = np.array(times) if times else np.full(1, np.nan)
    ms = float(times.mean() * 1e3)
    nprof = max(1, args.steps)
    phases = {k: round(v[0] / nprof, 4) for k, v in tm.items() if v[1]}
    log(f"bench: {args.steps} timed steps, mean {ms:.2f} ms; per step {phases}")
    # cell indexing and counts on the same positions: a par2grid of the
    # final parcels against a host recount of them
    counts_exact = None
    if want_err:
        gm.par2grid()
        counts_exact = counts_match_host(mem, gm)
        log(f"bench: nparg/nsparg of the final parcels equal the host recount: {counts_exact}")
    gm.close()

    # roofline of the tile kernels from their event times over the timed
    # steps (average launch)
    measured = args.size == 128 and not args.moist
    kern = {}
    for name, timer, nbytes in (("k_g2p_tiles", "push", BYTES_PUSH), ("k_p2g_tiles", "p2g_tiles", BYTES_P2G)):
        tot, calls = tm.get(timer, (0.0, 0))
        if not calls:
            continue
        avg = tot / calls
        achieved = nbytes * mem.n / (avg * 1e-3) / 1e9
        traffic, per_parcel = traffic_from_summary(name) if measured else (None, None)
        kern[name] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                      "traffic_bytes_per_parcel": per_parcel, "avg_launch_ms": round(avg, 4), "launches": calls,
                      "total_ms_per_step": round(tot / nprof, 4), "bytes_per_parcel": nbytes,
                      "units_per_launch": mem.n}
    dominant = max(kern, key=lambda k: kern[k]["total_ms_per_step"]) if kern else None
    roof = dict(kern[dominant]) if dominant else {}
    if dominant:
        roof["kernel"] = dominant
        roof["measured_on"] = ("the timed steps: device wall-clock stamps around every phase captured into the "
                               "replayed step graph on the member's stream" +
                               (" (eager launches between HIP events)" if args.eager else "") +
                               "; traffic from rocprofv3 counters (" + os.path.relpath(TRAFFIC_SUMMARY, ROOT) + ")")
        roof["limiter"] = "memory latency and VALU issue (per-kernel counters in the profiles), below the HBM bound"
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
        "roofline_kernels": kern,
    }
    # a failure of the batched or served form is reported in its entry (and
    # on stderr with its traceback); the plain form's line is still printed
    def form(name, fn):
        try:
            line[name] = fn()
        except Exception as err:   # noqa: BLE001
            traceback.print_exc()
            log(f"bench: {name} form failed: {err}")
            line[name] = {"error": f"{type(err).__name__}: {err}"}

    if "batched" in forms and args.members > 0:
        form("batched", lambda: bench_batched(args.members, max(1, args.ens_steps), errors=not args.no_member_errors))
    if "served" in forms and args.requests > 0:
        rates = [float(r) for r in args.rates.split(",") if r]
        lo, hi = (int(v) for v in args.req_steps.split(","))
        form("served", lambda: bench_served(args.requests, rates, (lo, hi)))
    if not args.no_cpu:
        cb, cdt, errors, exact = cpu_baseline(mem, snap)
        line["cpu_baseline"] = cb
        line["speedup_vs_cpu"] = round(cdt / (ms * 1e-3), 1)
        if errors is not None:
            line["max_rel_error_vs_fp64_step1"] = errors
            line["nparg_total_abs_diff_vs_fp64_step1"] = exact
    if counts_exact is not None:
        line["counts_bit_exact_same_positions"] = counts_exact
        log(f"bench: cpu baseline {cb['value']:.3e} parcel-steps/s; step-1 errors {errors}")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
