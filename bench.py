#!/usr/bin/env python3
"""Throughput of the batched DOP853 integrator on one MI355X.

Workload (one "step" = one integration of the whole batch, inputs resident in
HBM): 1,048,576 orbits in the Milky-Way model (bulge + Miyamoto-Nagai disk +
NFW halo), exponential-disk initial conditions, DOP853 with
rtol = atol = 1e-12, 1,001 dense outputs over t = 0..50, FP64; the
(1001, 1048576, 6) time-major output (50.4 GB) is written to HBM every step.

Prints one JSON line: accepted particle-steps/s of the chosen schedule, the
plain (one lane per orbit) figure beside it, the roofline of the integrator
kernel (algorithmic FP64 FLOP/s from the per-orbit step statistics, over the
kernel's HIP-event duration), the same workload at rtol = 1e-7 in FP64 and in
the mixed mode (FP32 stage forces) with its energy and position errors, the
p50/p99 latency of the serving loop under a Poisson stream of tidal-stream
requests, a ragged batch of mixed jobs, and the CPU oracle timed on a strided
sample of the same orbits.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO))

METRIC = "accepted particle-steps/s, 1M MW orbits DOP853, plain vs chosen; p50/p99 latency"
UNIT = "accepted particle-steps/s"
FP64_PEAK_TFLOPS = 78.6  # MI355X vector / matrix FP64


def parse_args():
    p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--n", type=int, default=1 << 20, help="orbits (default 1,048,576)")
    p.add_argument("--schedules",
                   default="plain,plain:index,persistent:index,persistent:auto:256,persistent,tuned",
                   help="comma list of schedule[:ordering[:block]] to time; schedule plain|"
                        "persistent|tuned (the autotuner's choice), ordering auto|index|cost "
                        "(auto: cost order from 16,384 orbits, either schedule), block the "
                        "threads per block (default: the library's, 512 -- packed dense output "
                        "-- for persistent launches, 256 for plain ones)")
    p.add_argument("--cpu-stride", type=int, default=512,
                   help="CPU baseline integrates every k-th orbit")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--serving-seconds", type=float, default=8.0,
                   help="length of the Poisson serving run (0 = skip)")
    p.add_argument("--serving-rate", type=float, default=8.0,
                   help="the light offered load, requests per second")
    p.add_argument("--ragged-jobs", type=int, default=4096,
                   help="jobs of the ragged mix to run, in HBM-sized batches (0 = skip)")
    p.add_argument("--no-precision", action="store_true",
                   help="skip the rtol = 1e-7 FP64 / mixed-precision leg")
    p.add_argument("--no-origins", action="store_true",
                   help="skip the moving-origin leg")
    return p.parse_args()


TRAFFIC_PROFILE = REPO / "profiles" / "dop853_mw_1M_traffic.md"


def profiled_traffic(n: int, schedule: str, tuned: dict | None):
    """HBM bytes per launch of the default persistent 1M-orbit integrator
    (cost order, 512-thread blocks with packed dense output, direct output),
    from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (a counter
    run cannot share a process with the timed loop); None for any other
    configuration."""
    same_kernel = schedule == "persistent" or (
        tuned is not None and tuned["schedule"] == 1 and tuned["block_size"] in (0, 512)
        and tuned["ordering"] == 2 and tuned.get("staging", 1) in (0, 1))
    if n != 1 << 20 or not same_kernel or not TRAFFIC_PROFILE.exists():
        return None
    for line in TRAFFIC_PROFILE.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":", 1)[1])
    return None


# Offered loads of the serving leg: a light one, and the heaviest the host
# link sustains with every trajectory copied back to host memory (results of
# 1e3-1e6 stars x 101 outputs average 0.7 GB a request; the link carries
# about 55 GB/s into pinned memory, so FP64 at rtol 1e-12 saturates it near
# 64 requests/s and the mixed mode at rtol 1e-7, with fewer steps per output,
# sooner in GPU terms).
SERVING_LOADS = (
    ("fp64", 1e-12, 48.0),
    ("mixed", 1e-7, 64.0),
)
SERVING_KEEP = ("rate_hz", "requests", "stars_mean", "p50_s", "p99_s", "max_s",
                "queue_wait_p50_s", "queue_wait_p99_s", "throughput", "wall_s",
                "chunk_target_ms", "gpu_busy", "copy_out_gb", "copy_out_gbps",
                "copy_out_busy", "client_waits", "chunk_launches")


def serving_leg(seconds: float, light_hz: float):
    """Poisson tidal-stream requests through drift_server at two offered loads
    per precision (FP64 at rtol 1e-12, the mixed mode at rtol 1e-7).  Results
    land in one pinned host arena.  The chunk-target autotuner runs first on a
    2 s prefix of the heavy load's stream and its choice serves both loads."""
    from astrodrift_amd import _native, loadgen

    arena = loadgen.PinnedArena(64 << 30)
    out = {"pinned_arena_gb": 64 * 2**30 / 1e9}
    for name, rtol, heavy_hz in SERVING_LOADS:
        prec = _native.PRECISION_MIXED if name == "mixed" else _native.PRECISION_FP64
        tuned = loadgen.tune_chunk_target(heavy_hz, rtol=rtol, precision=prec, arena=arena)
        leg = {"rtol": rtol, "tuner": tuned}
        for load, hz in (("light", light_hz), ("heavy", heavy_hz)):
            r = loadgen.run_poisson(rate_hz=hz, duration_s=seconds, rtol=rtol, precision=prec,
                                    chunk_target_ms=tuned["chunk_target_ms"], arena=arena)
            leg[load] = {k: r[k] for k in SERVING_KEEP}
            print(f"[bench] serving {name} {load}: {leg[load]}", file=sys.stderr, flush=True)
        out[name] = leg
    del arena
    return out


def ragged_leg(njobs: int):
    """The first `njobs` jobs of the ragged mix through drift_ragged_device,
    inputs and outputs resident in HBM.  Jobs whose outputs together exceed
    70% of the free HBM run as consecutive batches (the full 4,096-job mix
    writes 371 GB).  Each batch runs its DOP853 jobs and then its DOPR54 jobs
    as two calls (the library makes one persistent launch per method either
    way), so each method's time and its algorithmic FP64 FLOP/s (from the
    per-orbit statistics) are reported beside the whole mix's rate."""
    import numpy as np
    import torch

    from astrodrift_amd import _native, flops, workloads

    jobs = workloads.ragged_mix(njobs)
    pots = list(workloads.RAGGED_POTENTIALS)
    dev = torch.device("cuda:0")
    # the earlier legs' session caches up to 51 GB (the staged-output buffer
    # of the rtol 1e-7 launches): release it so the batches see the whole HBM
    _native.release_session(0)
    torch.cuda.empty_cache()
    budget = 0.7 * torch.cuda.mem_get_info(dev)[0]
    batches, cur, cur_bytes = [], [], 0
    for j in jobs:
        b = 48 * j.n * j.nt + 64 * j.n + 8 * j.nt
        if cur and cur_bytes + b > budget:
            batches.append(cur)
            cur, cur_bytes = [], 0
        cur.append(j)
        cur_bytes += b
    batches.append(cur)
    acc = rej = bad = 0
    seconds = []
    methods = {m: {"seconds": 0.0, "accepted": 0, "flop": 0.0, "jobs": 0}
               for m in ("DOP853", "DOPR54")}
    tuned = None
    for batch in batches:
        keep, descs = [], []
        for j in batch:
            s0 = torch.from_numpy(j.state0).to(dev)
            tt = torch.from_numpy(j.times).to(dev)
            out = torch.empty((j.nt, j.n, 6), dtype=torch.float64, device=dev)
            st = torch.zeros((j.n, 4), dtype=torch.int32, device=dev)
            keep.append((s0, tt, out, st))
            lt = math.log(j.rtol)
            method = _native.METHOD_DOP853 if j.method == "DOP853" else _native.METHOD_DOPR54
            descs.append(_native.DeviceJob(method, j.potential, j.n, j.nt, lt, lt, s0.data_ptr(),
                                           tt.data_ptr(), out.data_ptr(), st.data_ptr(),
                                           span_hint=float(j.times[-1] - j.times[0])))
        torch.cuda.synchronize()
        if tuned is None:
            # block size per method from a 5% prefix of the first batch's grids,
            # outside the timed region; later launches use it
            tuned = _native.tune_ragged(descs, pots)
            print(f"[bench] ragged tuned blocks: {tuned}", file=sys.stderr, flush=True)
        batch_s = 0.0
        for name in ("DOP853", "DOPR54"):
            sub = [d for d, j in zip(descs, batch) if j.method == name]
            if not sub:
                continue
            t0 = time.perf_counter()
            _native.ragged_device(sub, pots)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            methods[name]["seconds"] += dt
            batch_s += dt
        seconds.append(batch_s)
        for j, (_, _, _, st) in zip(batch, keep):
            s = st.cpu().numpy()
            a = int(s[:, 0].astype(np.int64).sum())
            acc += a
            rej += int(s[:, 1].astype(np.int64).sum())
            bad += int(np.count_nonzero(s[:, 3] & ~0xFF))
            m = methods[j.method]
            m["accepted"] += a
            m["jobs"] += 1
            spec = pots[j.potential]
            m["flop"] += (flops.dop853_flops(s, j.nt, spec) if j.method == "DOP853"
                          else flops.dopr54_flops(s, spec))
        print(f"[bench] ragged batch of {len(batch)} jobs: {seconds[-1]:.3f} s", file=sys.stderr,
              flush=True)
        del keep, descs
        torch.cuda.empty_cache()
    n = sum(j.n for j in jobs)
    out_bytes = sum(48 * j.n * j.nt for j in jobs)
    wall = sum(seconds)
    total_flop = sum(m["flop"] for m in methods.values())
    per_method = {}
    for name, m in methods.items():
        if m["seconds"] > 0:
            tf = m["flop"] / m["seconds"] / 1e12
            per_method[name] = {"jobs": m["jobs"], "seconds": round(m["seconds"], 3),
                                "accepted": m["accepted"], "value": m["accepted"] / m["seconds"],
                                "tflops": tf, "frac": tf / FP64_PEAK_TFLOPS}
    achieved = total_flop / wall / 1e12
    return {
        "jobs": njobs, "orbits": n, "output_gb": round(out_bytes / 1e9, 2),
        "dop853_jobs": sum(j.method == "DOP853" for j in jobs),
        "batches": [len(b) for b in batches], "batch_seconds": [round(x, 3) for x in seconds],
        "accepted": acc, "rejected": rej, "flagged_orbits": bad,
        "seconds": round(wall, 3), "value": acc / wall, "unit": UNIT, "tuned_blocks": tuned,
        "per_method": per_method,
        # algorithmic FP64 work of the whole mix (flops.py, from the per-orbit
        # statistics) over the summed call times, against the FP64 vector peak
        "roofline": {"bound": "fp64_valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                     "flop": total_flop},
    }


def kepler_leg(steps: int, warmup: int, stream):
    """The reference's CPU-runnable case: 1,024 Kepler orbits (the unit galpy
    Kepler force), DOPR54 rtol = atol = 1e-10, ts = (0, 20 pi, 101), on the
    GPU (the faster of the plain and persistent schedules, inputs resident).
    The cpu_baseline leg times the same case on the oracle."""
    import numpy as np
    import torch

    from astrodrift_amd import _native, workloads

    k = workloads.kepler_disk()
    dev = torch.device("cuda:0")
    s0 = torch.from_numpy(k.state0).to(dev)
    tt = torch.from_numpy(k.times).to(dev)
    out = torch.empty((k.nt, k.n, 6), dtype=torch.float64, device=dev)
    stats = torch.zeros((k.n, 4), dtype=torch.int32, device=dev)
    lt = math.log(k.rtol)
    runs = {}
    for name, sched in (("plain", _native.SCHEDULE_PLAIN),
                        ("persistent", _native.SCHEDULE_PERSISTENT)):
        def launch():
            _native.launch_device(_native.METHOD_DOPR54, s0.data_ptr(), tt.data_ptr(),
                                  out.data_ptr(), k.n, k.nt, lt, lt, None, stats.data_ptr(),
                                  schedule=sched, stream=stream.cuda_stream)

        for _ in range(warmup):
            launch()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            launch()
        torch.cuda.synchronize()
        runs[name] = (time.perf_counter() - t0) / steps
    acc = int(stats.cpu().numpy()[:, 0].astype(np.int64).sum())
    best = min(runs, key=runs.get)
    return {"workload": "kepler_disk_dopr54: 1024 Kepler orbits, DOPR54 rtol=atol=1e-10, "
                        "ts=(0, 20 pi, 101)",
            "schedule": best, "ms": {n_: round(v * 1e3, 3) for n_, v in runs.items()},
            "accepted": acc, "value": acc / runs[best], "unit": UNIT}


def moving_origin_leg(steps: int, warmup: int, n: int, stream, static_ms: float):
    """The 1M-orbit Milky-Way DOP853 workload (rtol = atol = 1e-12) with moving
    origins added to the static force: Plummer spheres (amp 1e-3, b 0.05) on
    tilted circular tracks of radius 1 and period 9, spread in phase,
    tabulated as quintic Hermite segments (200 over t = 0..52).  One origin,
    then eight (launches with two or more run in time windows with the tracks'
    rows in LDS).  Reports each launch time beside the static one."""
    import numpy as np
    import torch

    from astrodrift_amd import _native, workloads

    dev = torch.device("cuda:0")
    w = workloads.mw_disk(n=n)
    grid = np.linspace(0.0, 52.0, 201)
    om = 2.0 * math.pi / 9.0
    s0 = torch.from_numpy(w.state0).to(dev)
    tt = torch.from_numpy(w.times).to(dev)
    out = torch.empty((w.nt, w.n, 6), dtype=torch.float64, device=dev)
    stats = torch.zeros((w.n, 4), dtype=torch.int32, device=dev)
    lt = math.log(w.rtol)
    legs = {}
    for k in (1, 8):
        th = om * grid[:, None] + np.linspace(0.0, 2.0 * math.pi, k, endpoint=False)[None, :]
        x = np.stack([np.cos(th), np.sin(th), 0.1 * np.sin(th)], axis=-1)
        v = np.stack([-om * np.sin(th), om * np.cos(th), 0.1 * om * np.cos(th)], axis=-1)
        track = _native.Origins.hermite(np.concatenate([x, v], axis=-1), -om * om * x, grid,
                                        (1e-3, 0.05))
        with _native.BoundPotential(w.potential.with_origins(track)) as pot:
            def launch():
                _native.launch_device(_native.METHOD_DOP853, s0.data_ptr(), tt.data_ptr(),
                                      out.data_ptr(), w.n, w.nt, lt, lt, pot, stats.data_ptr(),
                                      schedule=_native.SCHEDULE_PERSISTENT,
                                      stream=stream.cuda_stream)

            for _ in range(warmup):
                launch()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                launch()
            torch.cuda.synchronize()
            sec = (time.perf_counter() - t0) / steps
        st = stats.cpu().numpy()
        acc = int(st[:, 0].astype(np.int64).sum())
        legs[k] = {"origins": k, "ms": round(sec * 1e3, 3),
                   "ratio_to_static": round(sec * 1e3 / static_ms, 3), "accepted": acc,
                   "value": acc / sec, "bad_status": int(np.count_nonzero(st[:, 3])),
                   "unit": UNIT}
        print(f"[bench] moving origins x{k}: {legs[k]}", file=sys.stderr, flush=True)
    del out
    torch.cuda.empty_cache()
    return {"workload": "the 1M MW DOP853 launch plus moving Plummer origins (200 track "
                        "segments each)", "static_ms": round(static_ms, 3),
            "one_origin": legs[1], "eight_origins": legs[8],
            "eight_over_one": round(legs[8]["ms"] / legs[1]["ms"], 3),
            # the one-origin figures at the top level, as before
            "ms": legs[1]["ms"], "ratio": legs[1]["ratio_to_static"], "value": legs[1]["value"],
            "unit": UNIT}


def precision_leg(steps: int, warmup: int, n: int, stream, subset: int = 4096):
    """The 1M-orbit Milky-Way DOP853 workload at rtol = atol = 1e-7: FP64 and
    the mixed mode (FP32 stage forces, FP64 state / error norm / step
    control), each with the plain and the persistent (cost-ordered) schedule.
    Reports accepted particle-steps/s per run, the mixed persistent run over
    the FP64 plain one, |dE/E| at t = 50 for every orbit (closed-form
    potential), and position error at t = 50 against an FP64 rtol = 1e-13 run
    of a strided subset."""
    import numpy as np
    import torch

    from astrodrift_amd import _native, flops, workloads

    dev = torch.device("cuda:0")
    w = workloads.mw_disk(n=n, rtol=1e-7, atol=1e-7)
    s0 = torch.from_numpy(w.state0).to(dev)
    tt = torch.from_numpy(w.times).to(dev)
    out = torch.empty((w.nt, w.n, 6), dtype=torch.float64, device=dev)
    stats = torch.zeros((w.n, 4), dtype=torch.int32, device=dev)
    lt = math.log(1e-7)
    runs, finals = {}, {}
    for prec_name, prec in (("fp64", _native.PRECISION_FP64), ("mixed", _native.PRECISION_MIXED)):
        for sched_name, sched, staging, ordering in (
                ("plain", _native.SCHEDULE_PLAIN, _native.STAGING_DIRECT, _native.ORDER_AUTO),
                ("plain_index", _native.SCHEDULE_PLAIN, _native.STAGING_DIRECT,
                 _native.ORDER_INDEX),
                ("persistent", _native.SCHEDULE_PERSISTENT, _native.STAGING_DIRECT,
                 _native.ORDER_AUTO),
                ("persistent_staged", _native.SCHEDULE_PERSISTENT, _native.STAGING_STAGED,
                 _native.ORDER_AUTO),
                ("tuned", _native.SCHEDULE_TUNED, _native.STAGING_AUTO, _native.ORDER_AUTO)):
            if prec_name == "mixed" and sched_name == "plain_index":
                continue
            if sched == _native.SCHEDULE_TUNED:
                # the autotuner's choice for this precision (its trials run on a
                # prefix of the grid, outside the timed loop)
                choice = _native.tune_device(_native.METHOD_DOP853, s0.data_ptr(), tt.data_ptr(),
                                             w.n, w.nt, lt, lt, w.potential, precision=prec,
                                             stream=stream.cuda_stream)
                print(f"[bench] precision {prec_name} tuned choice: {choice}", file=sys.stderr,
                      flush=True)

            def launch():
                _native.launch_device(_native.METHOD_DOP853, s0.data_ptr(), tt.data_ptr(),
                                      out.data_ptr(), w.n, w.nt, lt, lt, w.potential,
                                      stats.data_ptr(), precision=prec, schedule=sched,
                                      staging=staging, ordering=ordering,
                                      stream=stream.cuda_stream)

            for _ in range(warmup):
                launch()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                launch()
            torch.cuda.synchronize()
            sec = (time.perf_counter() - t0) / steps
            st = stats.cpu().numpy()
            acc = int(st[:, 0].astype(np.int64).sum())
            key = f"{prec_name}_{sched_name}"
            # algorithmic FLOPs of the reference algorithm (flops.py) over the
            # launch time; the mixed mode does the force part in FP32, so its
            # fraction of the FP64 peak is an algorithmic rate, not a unit's
            tf = flops.dop853_flops(st, w.nt, w.potential) / sec / 1e12
            runs[key] = {"ms": round(sec * 1e3, 3), "accepted": acc,
                         "value": acc / sec, "bad_status": int(np.count_nonzero(st[:, 3])),
                         "tflops": round(tf, 3), "frac": round(tf / FP64_PEAK_TFLOPS, 4)}
            if sched_name == "persistent_staged":
                finals[prec_name] = out[-1].cpu().numpy()
            print(f"[bench] precision {key}: {runs[key]}", file=sys.stderr, flush=True)
    del out
    torch.cuda.empty_cache()
    e0 = workloads.mw_energy(w.potential, w.state0)
    energy = {}
    for k, fin in finals.items():
        de = np.abs((workloads.mw_energy(w.potential, fin) - e0) / e0)
        energy[k] = {"median": float(np.median(de)), "p99": float(np.quantile(de, 0.99)),
                     "max": float(de.max())}
    # reference trajectories of a strided subset at rtol = atol = 1e-13
    idx = np.arange(0, w.n, max(1, w.n // subset))
    ss = torch.from_numpy(np.ascontiguousarray(w.state0[idx])).to(dev)
    so = torch.empty((w.nt, len(idx), 6), dtype=torch.float64, device=dev)
    _native.launch_device(_native.METHOD_DOP853, ss.data_ptr(), tt.data_ptr(), so.data_ptr(),
                          len(idx), w.nt, math.log(1e-13), math.log(1e-13), w.potential,
                          schedule=_native.SCHEDULE_PERSISTENT, stream=stream.cuda_stream)
    torch.cuda.synchronize()
    truth = so[-1].cpu().numpy()
    position = {}
    for k, fin in finals.items():
        err = np.linalg.norm(fin[idx, :3] - truth[:, :3], axis=1) / \
            np.linalg.norm(truth[:, :3], axis=1)
        position[k] = {"median": float(np.median(err)), "p99": float(np.quantile(err, 0.99)),
                       "max": float(err.max())}
    best_mixed = max(runs[f"mixed_{k}"]["value"] for k in ("persistent", "persistent_staged", "tuned"))
    best_fp64 = max(runs[f"fp64_{k}"]["value"] for k in ("persistent", "persistent_staged", "tuned"))
    best_key = max((f"mixed_{k}" for k in ("persistent", "persistent_staged", "tuned")),
                   key=lambda k: runs[k]["value"])
    return {
        "workload": f"mw_disk_dop853: {w.n} orbits, DOP853 rtol=atol=1e-7, {w.nt} outputs over "
                    "t=0..50; mixed = FP32 stage forces, FP64 state/error norm/step control; "
                    "staged = orbit-blocked samples + tiled transpose to the time-major output",
        "runs": runs,
        # plain = one lane per orbit in cost order (the default from 16,384
        # orbits); plain_index = one lane per orbit in input order, the
        # reference's launch
        "mixed_persistent_over_fp64_plain": best_mixed / runs["fp64_plain"]["value"],
        "mixed_persistent_over_fp64_plain_index": best_mixed / runs["fp64_plain_index"]["value"],
        "mixed_persistent_over_fp64_persistent": best_mixed / best_fp64,
        "energy_rel_err_t50": energy,
        "position_rel_err_t50_vs_fp64_rtol1e-13": position,
        "position_subset": len(idx),
        # the fastest mixed run: algorithmic FLOP/s against the FP64 vector
        # peak; at this tolerance a step emits about three samples, so the
        # launch is bound by its sample writes as much as by arithmetic
        "roofline": {"bound": "fp64_valu", "run": best_key,
                     "achieved": runs[best_key]["tflops"], "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": runs[best_key]["frac"]},
    }


def cpu_baseline(w, stride: int):
    """The oracle (the reference CPU algorithm restated) on orbits 0, stride,
    2*stride, ...: one thread (as the reference engine), then all host cores."""
    sys.path.insert(0, str(REPO / "tests"))
    import numpy as np

    import oracle_lib

    tol_r, tol_a = math.log(w.rtol), math.log(w.atol)
    idx = np.arange(0, w.n, stride)
    y0 = np.ascontiguousarray(w.state0[idx])
    t0 = time.perf_counter()
    _, st1 = oracle_lib.batch(1, y0, w.times, tol_r, tol_a, w.potential, nthreads=1,
                              want_out=False)
    one = time.perf_counter() - t0
    # the host cores this process may use: its affinity set, capped by the
    # thread share the environment grants it (OMP_NUM_THREADS, 16 if unset)
    affinity = len(os.sched_getaffinity(0))
    share = int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
    cores = max(1, min(affinity, share))
    t0 = time.perf_counter()
    _, stn = oracle_lib.batch(1, y0, w.times, tol_r, tol_a, w.potential, nthreads=cores,
                              want_out=False)
    many = time.perf_counter() - t0
    acc = int(st1["accepted"].sum())
    # the reference's CPU-runnable case whole: 1,024 Kepler orbits, DOPR54
    from astrodrift_amd import workloads

    k = workloads.kepler_disk()
    lk = math.log(k.rtol)
    t0 = time.perf_counter()
    _, kst = oracle_lib.batch(0, k.state0, k.times, lk, lk, None, nthreads=1, want_out=False)
    kep = time.perf_counter() - t0
    return {
        "kepler_dopr54": {"value": int(kst["accepted"].sum()) / kep, "cores": 1,
                          "seconds": round(kep, 3),
                          "sample": "all 1024 orbits of kepler_disk_dopr54"},
        "value": acc / one,
        "unit": UNIT,
        "cores": 1,
        "kind": "port",
        "sample": f"{len(idx)} orbits (every {stride}th of the {w.n}), DOP853 rtol=atol=1e-12, "
                  f"t=0..50, {acc} accepted steps, {one:.1f} s on 1 thread",
        "all_cores": {"value": int(stn["accepted"].sum()) / many, "cores": cores,
                      "affinity_cores": affinity, "nproc": os.cpu_count(),
                      "seconds": round(many, 2)},
    }


def main() -> int:
    args = parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE={world}); "
              "the orbits are independent and a multi-GPU run would shard them by range with no "
              "communication, which bench.py does not launch.", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from astrodrift_amd import _native, flops, workloads

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    w = workloads.mw_disk(n=args.n)
    s0 = torch.from_numpy(w.state0).to(dev)
    tt = torch.from_numpy(w.times).to(dev)
    out = torch.empty((w.nt, w.n, 6), dtype=torch.float64, device=dev)
    stats = torch.zeros((w.n, 4), dtype=torch.int32, device=dev)
    tol_r, tol_a = math.log(w.rtol), math.log(w.atol)
    # a dedicated stream: the kernels and the timing events share it
    stream = torch.cuda.Stream()

    sched_ids = {"plain": _native.SCHEDULE_PLAIN, "persistent": _native.SCHEDULE_PERSISTENT,
                 "tuned": _native.SCHEDULE_TUNED}
    order_ids = {"auto": _native.ORDER_AUTO, "index": _native.ORDER_INDEX,
                 "cost": _native.ORDER_COST}
    results = {}
    for name in [s.strip() for s in args.schedules.split(",") if s.strip()]:
        # "schedule" or "schedule:ordering"; the default ordering (auto) sorts
        # the orbits by estimated cost; "plain:index" is the reference's
        # launch: one lane per orbit, in input order
        sched_name, order_name, block_name = (name.split(":") + ["", ""])[:3]
        schedule = sched_ids[sched_name]
        ordering = order_ids[order_name or "auto"]
        block = int(block_name or 0)
        tuned = None
        if schedule == _native.SCHEDULE_TUNED:
            # untimed: the autotuner times its candidates on the first 5% of
            # the time grid and stores its choice for SCHEDULE_TUNED launches
            tuned = _native.tune_device(_native.METHOD_DOP853, s0.data_ptr(), tt.data_ptr(), w.n,
                                        w.nt, tol_r, tol_a, w.potential,
                                        stream=stream.cuda_stream)
            print(f"[bench] autotuner choice: {tuned}", file=sys.stderr, flush=True)

        def launch():
            _native.launch_device(_native.METHOD_DOP853, s0.data_ptr(), tt.data_ptr(),
                                  out.data_ptr(), w.n, w.nt, tol_r, tol_a, w.potential,
                                  stats.data_ptr(), schedule=schedule, ordering=ordering,
                                  block_size=block, stream=stream.cuda_stream)

        for _ in range(args.warmup):
            launch()
        torch.cuda.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        wall0 = time.perf_counter()
        ev0.record(stream)
        for _ in range(args.steps):
            launch()
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - wall0
        kernel_s = ev0.elapsed_time(ev1) / 1e3 / args.steps
        st = stats.cpu().numpy()
        accepted = int(st[:, 0].astype(np.int64).sum())
        flop = flops.dop853_flops(st, w.nt, w.potential)
        if schedule == _native.SCHEDULE_TUNED:
            shape = (None, tuned["block_size"] or None)
        else:
            shape = _native.launch_shape(_native.METHOD_DOP853, w.n, w.potential,
                                         schedule=schedule, block_size=block)
        results[name] = {
            "grid": shape[0],
            "block": shape[1],
            "ms_per_step": wall / args.steps * 1e3,
            "kernel_ms": kernel_s * 1e3,
            "accepted": accepted,
            "rejected": int(st[:, 1].astype(np.int64).sum()),
            "dense_steps": int(st[:, 2].astype(np.int64).sum()),
            "bad_status": int(np.count_nonzero(st[:, 3])),
            "value": accepted / (wall / args.steps),
            "tflops": flop / kernel_s / 1e12,
            "flop_per_launch": flop,
        }
        if tuned is not None:
            results[name]["tuned_options"] = tuned
        print(f"[bench] {name}: {results[name]}", file=sys.stderr, flush=True)

    chosen = max(results, key=lambda k: results[k]["value"])
    r = results[chosen]
    line = {
        "metric": METRIC,
        "value": r["value"],
        "unit": UNIT,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"mw_disk_dop853: {w.n} Milky-Way orbits (bulge+MN disk+NFW halo), "
                        "exponential-disk ICs seed 0x100000, DOP853 rtol=atol=1e-12, "
                        f"{w.nt} outputs over t=0..50, FP64, (nt, n, 6) output in HBM",
            "schedule": chosen,
            "plain_value": results.get("plain", {}).get("value"),
            "schedules": results,
        },
        "roofline": {
            # compute-bound against the FP64 peak; the path has no matrix
            # product, so the units doing the work are the FP64 vector ALUs
            # (MI355X vector FP64 peak 78.6 TFLOP/s); no MFMA is involved
            "bound": "fp64_valu",
            "compute_unit": "fp64 vector ALU",
            "achieved": r["tflops"],
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": r["tflops"] / FP64_PEAK_TFLOPS,
            "traffic": profiled_traffic(w.n, chosen, r.get("tuned_options")),
            "traffic_source": str(TRAFFIC_PROFILE.relative_to(REPO)),
        },
    }
    del out  # release the 50 GB batch output before the other legs
    torch.cuda.empty_cache()
    if not args.no_precision:
        line["precision"] = precision_leg(args.steps, args.warmup, args.n, stream)
    line["kepler_dopr54"] = kepler_leg(args.steps, args.warmup, stream)
    print(f"[bench] kepler dopr54: {line['kepler_dopr54']}", file=sys.stderr, flush=True)
    if not args.no_origins:
        line["moving_origin"] = moving_origin_leg(args.steps, args.warmup, args.n, stream,
                                                  r["ms_per_step"])
        print(f"[bench] moving origin: eight over one {line['moving_origin']['eight_over_one']}",
              file=sys.stderr, flush=True)
    if args.serving_seconds > 0:
        line["serving"] = serving_leg(args.serving_seconds, args.serving_rate)
    if args.ragged_jobs > 0:
        line["ragged"] = ragged_leg(args.ragged_jobs)
        print(f"[bench] ragged: {line['ragged']}", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(w, args.cpu_stride)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
