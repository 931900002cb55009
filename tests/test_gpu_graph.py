"""hipGraph capture of device-resident launches.

A persistent launch captured on a stream replays to the same bits as the
eager launch, with eager launches of the same session in between (they rotate
the work-queue counters the captured graph must not share).
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from astrodrift_amd import _native, workloads

pytestmark = pytest.mark.gpu


def _bits(a):
    return a.view(np.uint64)


def test_captured_launch_replays_match_eager(gpu):
    import torch

    dev = torch.device("cuda:0")
    w = workloads.mw_disk(n=20000, t_end=10.0, steps=101)
    s0 = torch.from_numpy(w.state0).to(dev)
    tt = torch.from_numpy(w.times).to(dev)
    out = torch.empty((w.nt, w.n, 6), dtype=torch.float64, device=dev)
    st = torch.zeros((w.n, 4), dtype=torch.int32, device=dev)
    lt = math.log(w.rtol)

    def launch(stream, ordering):
        _native.launch_device(_native.METHOD_DOP853, s0.data_ptr(), tt.data_ptr(), out.data_ptr(),
                              w.n, w.nt, lt, lt, w.potential, st.data_ptr(),
                              schedule=_native.SCHEDULE_PERSISTENT, ordering=ordering,
                              stream=stream)

    # eager reference; it also builds the MW table and the work-queue ring
    launch(torch.cuda.current_stream().cuda_stream, _native.ORDER_INDEX)
    torch.cuda.synchronize()
    ref, ref_st = out.cpu().numpy().copy(), st.cpu().numpy().copy()

    small = workloads.kepler_disk(n=64)
    k0 = torch.from_numpy(small.state0).to(dev)
    kt = torch.from_numpy(small.times).to(dev)
    kout = torch.empty((small.nt, small.n, 6), dtype=torch.float64, device=dev)

    cap = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=cap):
        launch(cap.cuda_stream, _native.ORDER_AUTO)  # AUTO is INDEX under capture
    for replay in range(3):
        out.fill_(float("nan"))
        st.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        print(f"replay {replay}: accepted={int(st[:, 0].sum())}", flush=True)
        assert np.array_equal(_bits(out.cpu().numpy()), _bits(ref)), replay
        assert np.array_equal(st.cpu().numpy(), ref_st), replay
        # a full turn of the queue ring of eager persistent launches
        for _ in range(70):
            _native.launch_device(_native.METHOD_DOPR54, k0.data_ptr(), kt.data_ptr(),
                                  kout.data_ptr(), small.n, small.nt, math.log(small.rtol),
                                  math.log(small.atol), schedule=_native.SCHEDULE_PERSISTENT)
        torch.cuda.synchronize()


def test_cost_ordering_is_refused_under_capture(gpu):
    import torch

    dev = torch.device("cuda:0")
    w = workloads.kepler_disk(n=20000)
    s0 = torch.from_numpy(w.state0).to(dev)
    tt = torch.from_numpy(w.times).to(dev)
    out = torch.empty((w.nt, w.n, 6), dtype=torch.float64, device=dev)
    lt = math.log(w.rtol)
    _native.launch_device(_native.METHOD_DOPR54, s0.data_ptr(), tt.data_ptr(), out.data_ptr(),
                          w.n, w.nt, lt, lt, schedule=_native.SCHEDULE_PERSISTENT)
    torch.cuda.synchronize()
    cap = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with pytest.raises(_native.DriftNativeError, match="cannot be captured"):
        with torch.cuda.graph(g, stream=cap):
            _native.launch_device(_native.METHOD_DOPR54, s0.data_ptr(), tt.data_ptr(),
                                  out.data_ptr(), w.n, w.nt, lt, lt,
                                  schedule=_native.SCHEDULE_PERSISTENT,
                                  ordering=_native.ORDER_COST, stream=cap.cuda_stream)
    torch.cuda.synchronize()
    # the session still launches normally afterwards
    _native.launch_device(_native.METHOD_DOPR54, s0.data_ptr(), tt.data_ptr(), out.data_ptr(),
                          w.n, w.nt, lt, lt, schedule=_native.SCHEDULE_PERSISTENT)
    torch.cuda.synchronize()
    assert not torch.isnan(out).any()


def test_staged_output_is_refused_and_tuned_falls_back_under_capture(gpu):
    """Staged output writes through the session's staging buffer, which a
    later eager launch may reallocate, so a captured launch refuses it; a
    TUNED launch whose stored choice is staged (and possibly cost-ordered)
    is captured with direct output and index order instead, and replays to
    the eager launch's bits."""
    import torch

    dev = torch.device("cuda:0")
    # rtol 1e-7 over a fine grid: several samples per step, where the tuner
    # always chooses staged output
    w = workloads.mw_disk(n=20000, t_end=10.0, steps=401, rtol=1e-7, atol=1e-7)
    s0 = torch.from_numpy(w.state0).to(dev)
    tt = torch.from_numpy(w.times).to(dev)
    out = torch.empty((w.nt, w.n, 6), dtype=torch.float64, device=dev)
    st = torch.zeros((w.n, 4), dtype=torch.int32, device=dev)
    lt = math.log(w.rtol)

    def launch(stream, **kw):
        _native.launch_device(_native.METHOD_DOP853, s0.data_ptr(), tt.data_ptr(), out.data_ptr(),
                              w.n, w.nt, lt, lt, w.potential, st.data_ptr(), stream=stream, **kw)

    choice = _native.tune_device(_native.METHOD_DOP853, s0.data_ptr(), tt.data_ptr(), w.n, w.nt,
                                 lt, lt, w.potential)
    print(f"tuner choice: {choice}", flush=True)
    assert choice["staging"] == _native.STAGING_STAGED
    # eager reference (staged or not, the bits are the same)
    launch(torch.cuda.current_stream().cuda_stream, schedule=_native.SCHEDULE_TUNED)
    torch.cuda.synchronize()
    ref, ref_st = out.cpu().numpy().copy(), st.cpu().numpy().copy()

    cap = torch.cuda.Stream()
    with pytest.raises(_native.DriftNativeError, match="cannot be captured"):
        with torch.cuda.graph(torch.cuda.CUDAGraph(), stream=cap):
            launch(cap.cuda_stream, schedule=_native.SCHEDULE_PERSISTENT,
                   ordering=_native.ORDER_INDEX, staging=_native.STAGING_STAGED)
    torch.cuda.synchronize()

    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=cap):
        launch(cap.cuda_stream, schedule=_native.SCHEDULE_TUNED)
    for replay in range(2):
        out.fill_(float("nan"))
        st.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(_bits(out.cpu().numpy()), _bits(ref)), replay
        assert np.array_equal(st.cpu().numpy(), ref_st), replay
        # an eager staged launch of a larger shape reallocates the staging
        # buffer between replays; the graph does not use it
        big = workloads.mw_disk(n=30000, t_end=10.0, steps=401, rtol=1e-7, atol=1e-7)
        b0 = torch.from_numpy(big.state0).to(dev)
        bout = torch.empty((big.nt, big.n, 6), dtype=torch.float64, device=dev)
        _native.launch_device(_native.METHOD_DOP853, b0.data_ptr(), tt.data_ptr(),
                              bout.data_ptr(), big.n, big.nt, lt, lt, big.potential,
                              schedule=_native.SCHEDULE_PERSISTENT,
                              staging=_native.STAGING_STAGED)
        torch.cuda.synchronize()
        del bout
