"""The continuous-batching server: every result is the same alone and in the
mix (bit for bit: chunking saves and restores the complete integrator state,
and orbits never interact), agrees with the oracle, and requests submitted
while others run are admitted and overlap with them."""

from __future__ import annotations

import math
import time

import numpy as np
import pytest

from astrodrift_amd import _native, serving, workloads

import oracle_lib

pytestmark = pytest.mark.gpu

MW = _native.PotentialSpec.milky_way()


def stream_ics(n, seed):
    """A tidal-stream-like request: Plummer-sphere offsets (scale 0.002,
    dispersion 0.005) around an eccentric Milky-Way progenitor orbit."""
    rng = np.random.default_rng(seed)
    R = rng.uniform(0.5, 2.0)
    vc = workloads.mw_vcirc(MW, np.array([R]))[0]
    prog = np.array([R, 0.0, 0.0, 0.0, rng.uniform(0.7, 1.0) * vc, 0.0])
    # Plummer positions: r = b / sqrt(u^(-2/3) - 1)
    u = rng.uniform(1e-6, 0.999, n)
    r = 0.002 / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    y0 = np.tile(prog, (n, 1))
    y0[:, :3] += np.minimum(r, 0.05)[:, None] * d
    y0[:, 3:] += rng.normal(0.0, 0.005, (n, 3))
    return y0


def test_alone_equals_mixed_and_matches_oracle(gpu):
    specs = [(0, 96, 21, 8.0, 1e-10, None), (1, 160, 31, 12.0, 1e-11, MW),
             (1, 64, 41, 5.0, 1e-9, _native.PotentialSpec.plummer(1.0, 0.1)),
             (0, 200, 11, 6.0, 1e-10, MW)]
    reqs = []
    for k, (method, n, nt, span, tol, pot) in enumerate(specs):
        y0 = stream_ics(n, seed=100 + k) if pot is MW else \
            workloads.kepler_disk(n=n, seed=100 + k).state0
        reqs.append((method, y0, workloads.grid_times(0.0, span, nt), tol, pot))

    # each request alone
    alone = []
    with serving.Server(chunk_samples=4) as srv:
        for method, y0, t, tol, pot in reqs:
            r = srv.submit(method, y0, t, tol, tol, pot)
            srv.wait(r, timeout_s=120)
            alone.append(r.out.copy())
    # all together, submitted back to back
    with serving.Server(chunk_samples=4) as srv:
        handles = [srv.submit(m, y0, t, tol, tol, pot) for m, y0, t, tol, pot in reqs]
        infos = [srv.wait(h, timeout_s=120) for h in handles]
        mixed = [h.out for h in handles]
        for h, info in zip(handles, infos):
            assert info.state == serving.DONE and info.chunks >= 1
            assert info.accepted_steps == int(h.stats["accepted"].sum())
    for (method, y0, t, tol, pot), a, m in zip(reqs, alone, mixed):
        assert np.array_equal(a, m)  # bit for bit
        ref, _ = oracle_lib.batch(method, y0, t, math.log(tol), math.log(tol), pot, nthreads=8)
        dp = np.linalg.norm(m[..., :3] - ref[..., :3], axis=-1) / np.linalg.norm(ref[..., :3], axis=-1)
        assert dp.max() < 1e-8, dp.max()


def test_chunking_does_not_change_results(gpu):
    """Chunks of 1, 7 and all 60 steps of the grid give the same bits, and the
    same bits as one eager drift_ragged_device launch of the request (the same
    kernel, no pauses, no graph).  The server captures a chunk's work into a
    hipGraph once and replays it for every further chunk of the same shape."""
    import torch

    w = workloads.mw_disk(n=300, seed=5, t_end=6.0, steps=61, rtol=1e-11, atol=1e-11)
    outs = []
    for chunk in (1, 7, 1000):
        with serving.Server(chunk_samples=chunk) as srv:
            r = srv.submit(1, w.state0, w.times, w.rtol, w.atol, w.potential)
            info = srv.wait(r, timeout_s=120)
            st = srv.stats()
            outs.append(r.out.copy())
        print(f"chunk {chunk}: {info.chunks} chunks, {st}")
        assert st["eager_chunks"] == 0
        assert st["graph_captures"] == 1 and st["graph_launches"] == info.chunks
        assert info.chunks == math.ceil(60 / min(chunk, 60))
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[1], outs[2])
    dev = torch.device("cuda:0")
    s0 = torch.from_numpy(w.state0).to(dev)
    tt = torch.from_numpy(w.times).to(dev)
    out = torch.empty((w.nt, w.n, 6), dtype=torch.float64, device=dev)
    job = _native.DeviceJob(_native.METHOD_DOP853, 0, w.n, w.nt, math.log(w.rtol),
                            math.log(w.atol), s0.data_ptr(), tt.data_ptr(), out.data_ptr())
    _native.ragged_device([job], [w.potential])
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), outs[0])


def test_requests_overlap_and_budget_admission(gpu):
    """A small HBM budget admits requests one or two at a time; later
    requests are queued, then admitted as earlier ones finish."""
    w = workloads.mw_disk(n=2000, seed=9, t_end=5.0, steps=51, rtol=1e-10, atol=1e-10)
    per = 8 * (6 * 2000 + 51 + 6 * 2000 * 51 + 20 * 2000) + 16 * 2000 + 256
    with serving.Server(chunk_samples=5, hbm_budget_bytes=int(2.5 * per)) as srv:
        hs = [srv.submit(1, w.state0, w.times, w.rtol, w.atol, w.potential) for _ in range(6)]
        infos = [srv.wait(h, timeout_s=120) for h in hs]
        st = srv.stats()
        busy = srv.busy()
    # the admitted set changed as requests came and went: captured anew each time
    print(st, busy)
    # every result copied once; chunk and copy time accounted
    assert busy["copy_bytes"] == 6 * 51 * 2000 * 6 * 8
    assert busy["busy_s"] > 0 and busy["copy_s"] > 0
    assert busy["busy_s"] < max(i.complete_s for i in infos) - min(i.submit_s for i in infos)
    assert st["eager_chunks"] == 0 and st["graph_captures"] >= 2
    assert st["graph_launches"] > st["graph_captures"]
    admits = sorted(i.admit_s for i in infos)
    completes = sorted(i.complete_s for i in infos)
    # at most two admitted before the first completion
    assert sum(a <= completes[0] for a in admits) <= 2
    for h in hs[1:]:
        assert np.array_equal(h.out, hs[0].out)


def test_dopr54_error_code_reaches_the_request(gpu):
    plunge = np.array([[1.0, 0.0, 0.0, -3.0, 0.0, 0.0]])
    with serving.Server() as srv:
        r = srv.submit(0, plunge, workloads.grid_times(0.0, 5.0, 64), 1e-13, 1e-13)
        with pytest.raises(_native.DriftNativeError):
            srv.wait(r, timeout_s=60)
        assert srv.query(r).status == _native.DRIFT_ERR_INTEGRATION


def test_mixed_precision_server_matches_one_shot_ragged(gpu):
    """The server in the mixed mode (FP32 stage forces, rtol 1e-7): chunked and
    graph-replayed, the same bits as one eager mixed ragged launch of the
    request, and within 1e-3 of the FP64 one (the mode's accuracy class)."""
    import torch

    w = workloads.mw_disk(n=500, seed=12, t_end=5.0, steps=41, rtol=1e-7, atol=1e-7)
    with serving.Server(chunk_samples=6, precision=_native.PRECISION_MIXED) as srv:
        r = srv.submit(1, w.state0, w.times, w.rtol, w.atol, w.potential)
        info = srv.wait(r, timeout_s=120)
        st = srv.stats()
    assert info.state == serving.DONE and st["graph_launches"] == info.chunks
    dev = torch.device("cuda:0")
    s0 = torch.from_numpy(w.state0).to(dev)
    tt = torch.from_numpy(w.times).to(dev)
    outs = {}
    for prec in (_native.PRECISION_MIXED, _native.PRECISION_FP64):
        out = torch.empty((w.nt, w.n, 6), dtype=torch.float64, device=dev)
        job = _native.DeviceJob(_native.METHOD_DOP853, 0, w.n, w.nt, math.log(w.rtol),
                                math.log(w.atol), s0.data_ptr(), tt.data_ptr(), out.data_ptr())
        _native.ragged_device([job], [w.potential], precision=prec)
        torch.cuda.synchronize()
        outs[prec] = out.cpu().numpy()
    assert np.array_equal(r.out, outs[_native.PRECISION_MIXED])
    ref = outs[_native.PRECISION_FP64]
    dp = np.linalg.norm(r.out[..., :3] - ref[..., :3], axis=-1) / np.linalg.norm(ref[..., :3],
                                                                                 axis=-1)
    print(f"mixed server vs FP64: worst position rel diff {dp.max():.2e}")
    assert 0.0 < dp.max() < 1e-3


def test_full_size_tidal_stream_requests_match_oracle(gpu):
    """Requests at the sizes of the serving load (loadgen's tidal streams:
    a Plummer sphere around an eccentric Milky-Way progenitor, DOP853 at rtol
    1e-10, 101 outputs): 100,000 stars over 100 time units, the longest span
    of the load, and 1,000,000 stars, its largest request, submitted together.
    The 100,000-star result equals the same request served alone, bit for
    bit, and 16 stars of each, strided over the request, agree with the
    oracle to the parity bar (1e-8 relative per sample)."""
    rng = np.random.default_rng(2024)
    small = workloads.tidal_stream(100_000, rng, span=100.0)
    large = workloads.tidal_stream(1_000_000, rng, span=60.0)
    print("serving full-size requests: 1e5 stars x 100 t.u. and 1e6 stars x 60 t.u.", flush=True)
    with serving.Server() as srv:
        hs = srv.submit(1, small.state0, small.times, small.rtol, small.atol, small.potential)
        hl = srv.submit(1, large.state0, large.times, large.rtol, large.atol, large.potential)
        infos = [srv.wait(h, timeout_s=300) for h in (hs, hl)]
    for h, info in zip((hs, hl), infos):
        assert info.state == serving.DONE
        assert info.accepted_steps == int(h.stats["accepted"].sum())
        assert not np.isnan(h.out).any()
        assert not np.any(h.stats["status"])
    with serving.Server() as srv:
        alone = srv.submit(1, small.state0, small.times, small.rtol, small.atol, small.potential)
        srv.wait(alone, timeout_s=300)
    assert np.array_equal(alone.out.view(np.uint64), hs.out.view(np.uint64))
    for w, h in ((small, hs), (large, hl)):
        idx = np.linspace(0, w.n - 1, 16).astype(np.int64)
        lt = math.log(w.rtol)
        ref, rst = oracle_lib.batch(1, np.ascontiguousarray(w.state0[idx]), w.times, lt, lt,
                                    w.potential, nthreads=8)
        got = h.out[:, idx]
        dp = np.linalg.norm(got[..., :3] - ref[..., :3], axis=-1) / \
            np.linalg.norm(ref[..., :3], axis=-1)
        dv = np.linalg.norm(got[..., 3:] - ref[..., 3:], axis=-1) / \
            np.linalg.norm(ref[..., 3:], axis=-1)
        same = float(np.mean(h.stats["accepted"][idx] == rst["accepted"]))
        print(f"  {w.n} stars: worst rel vs oracle {max(dp.max(), dv.max()):.3e}, identical "
              f"accepted counts on {same:.2f} of the checked stars", flush=True)
        assert dp.max() < 1e-8 and dv.max() < 1e-8


def test_poisson_load_into_pinned_arena(gpu):
    """The load generator's short run with results in a pinned host arena:
    every request finishes, every arena range comes back (the free list is
    one range again), and the copy accounting matches the results' bytes."""
    from astrodrift_amd import loadgen

    arena = loadgen.PinnedArena(8 << 30)
    total = arena._words.size
    print("poisson load: 16 requests/s for 1 s, 1e3-1e5 stars, pinned arena", flush=True)
    r = loadgen.run_poisson(rate_hz=16.0, duration_s=1.0, n_max=100_000, rtol=1e-10,
                            arena=arena)
    print({k: r[k] for k in ("requests", "p50_s", "p99_s", "gpu_busy", "copy_out_gbps")})
    assert r["requests"] >= 1 and r["client_waits"] == 0
    assert arena._free == [(0, total)]
    assert 0.0 < r["gpu_busy"] < 1.0 and r["copy_out_gb"] > 0
    assert r["p50_s"] <= r["p99_s"] <= r["max_s"]
