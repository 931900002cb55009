"""Poisson load for the serving loop (tidal-stream requests).

Requests arrive as a Poisson process at ``rate_hz``; each is a tidal stream
of N stars (N log-uniform in [n_min, n_max]) integrated with DOP853 in the
Milky-Way model.  Reports per-request latency (submission to results in host
memory) percentiles and the accepted particle-steps per second the server
sustained over the run.
"""

from __future__ import annotations

import math
import time

import numpy as np

from . import _native, serving, workloads


class PinnedArena:
    """Receive buffers for results in pinned (page-locked) host memory, which
    the device copies into at the host link's rate (pageable memory takes
    about 9 GB/s on the MI355X box).  One allocation, carved first-fit into
    per-request ranges that are returned when the request has been read."""

    def __init__(self, nbytes: int):
        import torch

        self._buf = torch.empty(nbytes // 8, dtype=torch.float64, pin_memory=True)
        self._words = self._buf.numpy()
        self._free = [(0, self._words.size)]  # (start, length) in words, sorted

    def take(self, shape) -> tuple[np.ndarray, tuple[int, int]] | None:
        need = int(np.prod(shape))
        need += (-need) % 2  # keep every range 16-byte aligned
        for k, (start, length) in enumerate(self._free):
            if length >= need:
                self._free[k:k + 1] = [(start + need, length - need)] if length > need else []
                return self._words[start:start + int(np.prod(shape))].reshape(shape), (start, need)
        return None

    def give(self, rng: tuple[int, int]) -> None:
        self._free.append(rng)
        self._free.sort()
        merged = []
        for start, length in self._free:
            if merged and merged[-1][0] + merged[-1][1] == start:
                merged[-1] = (merged[-1][0], merged[-1][1] + length)
            else:
                merged.append((start, length))
        self._free = merged


def run_poisson(rate_hz: float = 10.0, duration_s: float = 8.0, n_min: int = 1000,
                n_max: int = 1_000_000, seed: int = 5, rtol: float = 1e-10,
                precision: int = _native.PRECISION_FP64, chunk_samples: int = 0,
                chunk_target_ms: float = 0.0, hbm_budget_bytes: int = 0,
                device: int = 0, arena: PinnedArena | None = None) -> dict:
    """chunk_samples = 0: the server adapts its chunk length to chunk_target_ms.
    With an arena, results land in its pinned ranges (a request whose result
    does not fit the arena's free space waits at the client until earlier
    ones are read; that wait is outside its latency and counted in
    client_waits)."""
    rng = np.random.default_rng(seed)
    arrivals = []
    t = 0.0
    while True:
        t += rng.exponential(1.0 / rate_hz)
        if t > duration_s:
            break
        arrivals.append(t)
    sizes = np.exp(rng.uniform(math.log(n_min), math.log(n_max), len(arrivals))).astype(int)
    loads = [workloads.tidal_stream(int(nn), rng, rtol=rtol, atol=rtol) for nn in sizes]
    infos = []
    with serving.Server(device=device, chunk_samples=chunk_samples, precision=precision,
                        hbm_budget_bytes=hbm_budget_bytes,
                        chunk_target_ms=chunk_target_ms) as srv:
        # warm the kernels and the table cache outside the measured stream
        warm = workloads.tidal_stream(256, np.random.default_rng(0), span=1.0, rtol=rtol, atol=rtol)
        srv.wait(srv.submit(1, warm.state0, warm.times, rtol, rtol, warm.potential))
        start = time.perf_counter()
        pending = []
        held = {}  # ticket -> arena range
        client_waits = 0

        def reap(block: bool) -> None:
            nonlocal pending
            still = []
            for r in pending:
                info = srv.wait(r) if block else srv.query(r)
                if info.state in (serving.DONE, serving.FAILED):
                    infos.append(info if block else srv.wait(r))
                    r.out = None
                    if r.ticket in held:
                        arena.give(held.pop(r.ticket))
                else:
                    still.append(r)
            pending = still

        for at, w in zip(arrivals, loads):
            delay = start + at - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            out, rng = None, None
            if arena is not None:
                got = arena.take((w.nt, w.n, 6))
                while got is None and pending:
                    client_waits += 1
                    srv.wait(pending[0])
                    reap(False)
                    got = arena.take((w.nt, w.n, 6))
                if got is not None:
                    out, rng = got
            r = srv.submit(1, w.state0, w.times, rtol, rtol, w.potential, out=out)
            if rng is not None:
                held[r.ticket] = rng
            pending.append(r)
            # release finished requests' host buffers as we go
            reap(False)
        reap(True)
        wall = time.perf_counter() - start
        chunks = srv.stats()
        busy = srv.busy()
    lat = np.array([i.complete_s - i.submit_s for i in infos])
    wait = np.array([i.admit_s - i.submit_s for i in infos])
    first = min(i.submit_s for i in infos)
    last = max(i.complete_s for i in infos)
    steps = sum(int(i.accepted_steps) for i in infos)
    return {
        "rate_hz": rate_hz,
        "requests": len(infos),
        "stars_mean": float(sizes.mean()) if len(sizes) else 0.0,
        "p50_s": float(np.percentile(lat, 50)),
        "p99_s": float(np.percentile(lat, 99)),
        "max_s": float(lat.max()),
        "queue_wait_p50_s": float(np.percentile(wait, 50)),
        "queue_wait_p99_s": float(np.percentile(wait, 99)),
        "accepted_steps": steps,
        "throughput": steps / (last - first),
        "wall_s": wall,
        "chunk_samples": chunk_samples if chunk_samples > 0 else "adaptive",
        "chunk_target_ms": (chunk_target_ms or 10.0) if chunk_samples == 0 else None,
        "rtol": rtol,
        "precision": "mixed" if precision == _native.PRECISION_MIXED else "fp64",
        "chunk_launches": chunks,
        # the GPU's share of the measured stream (chunk launches over the wall
        # time from the first submission) and the host link's rate for results
        "gpu_busy": busy["busy_s"] / (last - first),
        "copy_out_gb": busy["copy_bytes"] / 1e9,
        "copy_out_gbps": busy["copy_bytes"] / busy["copy_s"] / 1e9 if busy["copy_s"] else None,
        "copy_out_busy": busy["copy_s"] / (last - first),
        "pinned_results": arena is not None,
        "client_waits": client_waits,
    }


def tune_chunk_target(rate_hz: float, prefix_s: float = 2.0,
                      candidates=(5.0, 10.0, 20.0, 40.0), **kw) -> dict:
    """The serving loop's chunk-length autotuner: replays the first prefix_s
    seconds of the Poisson stream (same seed, so the same requests) on a fresh
    server once per candidate chunk target, and chooses the target with the
    lowest p99 latency (ties: the higher throughput).  Short chunks admit
    queued requests sooner; long chunks pay the pause/resume at their
    boundaries less often."""
    trials = {}
    for c in candidates:
        r = run_poisson(rate_hz=rate_hz, duration_s=prefix_s, chunk_target_ms=c, **kw)
        trials[c] = {"p99_s": r["p99_s"], "p50_s": r["p50_s"], "throughput": r["throughput"]}
    best = min(candidates, key=lambda c: (round(trials[c]["p99_s"], 3), -trials[c]["throughput"]))
    return {"chunk_target_ms": best, "prefix_s": prefix_s, "rate_hz": rate_hz,
            "trials": {str(c): v for c, v in trials.items()}}
