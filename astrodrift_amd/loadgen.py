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


def run_poisson(rate_hz: float = 10.0, duration_s: float = 8.0, n_min: int = 1000,
                n_max: int = 1_000_000, seed: int = 5, rtol: float = 1e-10,
                precision: int = _native.PRECISION_FP64, chunk_samples: int = 0,
                chunk_target_ms: float = 0.0, hbm_budget_bytes: int = 0,
                device: int = 0) -> dict:
    """chunk_samples = 0: the server adapts its chunk length to chunk_target_ms."""
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
        for at, w in zip(arrivals, loads):
            delay = start + at - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            pending.append(srv.submit(1, w.state0, w.times, rtol, rtol, w.potential))
            # release finished requests' host buffers as we go
            still = []
            for r in pending:
                info = srv.query(r)
                if info.state in (serving.DONE, serving.FAILED):
                    infos.append(srv.wait(r))
                    r.out = None
                else:
                    still.append(r)
            pending = still
        for r in pending:
            infos.append(srv.wait(r))
            r.out = None
        wall = time.perf_counter() - start
        chunks = srv.stats()
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
    }
