"""Staged output (DRIFT_STAGING_STAGED): samples go to an orbit-blocked
scratch buffer and a tiled transpose writes the time-major output.  The
output and the step statistics must equal the direct launch's bit for bit, for
both methods and both schedules, at orbit counts and output counts that do not
fill the transpose's tiles (64 orbits x 16 samples), including the smallest
shapes."""

from __future__ import annotations

import math

import numpy as np
import pytest

from astrodrift_amd import _native, workloads

pytestmark = pytest.mark.gpu


def _run(w, method, schedule, staging, precision=_native.PRECISION_FP64):
    import torch

    dev = torch.device("cuda:0")
    s0 = torch.from_numpy(np.ascontiguousarray(w.state0)).to(dev)
    tt = torch.from_numpy(w.times).to(dev)
    out = torch.full((w.nt, w.state0.shape[0], 6), -7.0, dtype=torch.float64, device=dev)
    st = torch.zeros((w.state0.shape[0], 4), dtype=torch.int32, device=dev)
    lt = math.log(w.rtol)
    _native.launch_device(method, s0.data_ptr(), tt.data_ptr(), out.data_ptr(), w.state0.shape[0],
                          w.nt, lt, lt, w.potential, st.data_ptr(), precision=precision,
                          schedule=schedule, staging=staging)
    torch.cuda.synchronize()
    return out.cpu().numpy(), st.cpu().numpy()


def _same(a, b):
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("method", [_native.METHOD_DOPR54, _native.METHOD_DOP853])
@pytest.mark.parametrize("schedule", [_native.SCHEDULE_PLAIN, _native.SCHEDULE_PERSISTENT])
@pytest.mark.parametrize("n,steps", [(1000, 203), (1, 2), (65, 17), (130, 16)])
def test_staged_equals_direct(gpu, method, schedule, n, steps):
    w = workloads.mw_disk(n=n, seed=3, t_end=8.0, steps=steps, rtol=1e-10, atol=1e-10)
    if method == _native.METHOD_DOPR54:
        k = workloads.kepler_disk(n=n)
        w = workloads.Workload("kepler", "DOPR54", k.state0, workloads.grid_times(0.0, 8.0, steps),
                               1e-10, 1e-10, None)
    d_out, d_st = _run(w, method, schedule, _native.STAGING_DIRECT)
    s_out, s_st = _run(w, method, schedule, _native.STAGING_STAGED)
    assert _same(d_out, s_out)
    assert np.array_equal(d_st, s_st)
    assert not np.any(d_out == -7.0)  # every record written


def test_staged_mixed_precision_and_reuse(gpu):
    """A larger staged launch after smaller ones (the scratch buffer grows),
    in the mixed mode, then a smaller one again."""
    for n, steps in [(300, 40), (5000, 301), (700, 99)]:
        w = workloads.mw_disk(n=n, seed=n, t_end=20.0, steps=steps, rtol=1e-7, atol=1e-7)
        d_out, d_st = _run(w, _native.METHOD_DOP853, _native.SCHEDULE_PERSISTENT,
                           _native.STAGING_DIRECT, _native.PRECISION_MIXED)
        s_out, s_st = _run(w, _native.METHOD_DOP853, _native.SCHEDULE_PERSISTENT,
                           _native.STAGING_STAGED, _native.PRECISION_MIXED)
        assert _same(d_out, s_out) and np.array_equal(d_st, s_st)


def test_ragged_refuses_staging(gpu):
    import torch

    w = workloads.kepler_disk(n=8)
    dev = torch.device("cuda:0")
    s0 = torch.from_numpy(w.state0).to(dev)
    tt = torch.from_numpy(w.times).to(dev)
    out = torch.empty((w.nt, w.n, 6), dtype=torch.float64, device=dev)
    st = torch.zeros((w.n, 4), dtype=torch.int32, device=dev)
    lt = math.log(w.rtol)
    job = _native.DeviceJob(_native.METHOD_DOP853, 0, w.n, w.nt, lt, lt, s0.data_ptr(),
                            tt.data_ptr(), out.data_ptr(), st.data_ptr())
    s = _native.session(0)
    import ctypes

    arr = (_native.CJob * 1)(job.to_c())
    pots = (_native.CPotential * 1)(_native.PotentialSpec.galpy_kepler().to_c())
    opts = _native.CLaunchOptions(0, 0, 0, 0, 0, _native.STAGING_STAGED)
    rc = s.lib.drift_ragged_device(s.handle, arr, 1, pots, 1, ctypes.byref(opts), None)
    assert rc == _native.DRIFT_ERR_UNSUPPORTED
    assert "staging" in _native.last_error()
