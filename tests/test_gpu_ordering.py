"""Cost ordering of the orbits (drift_launch_options.ordering).

The order in which lanes take orbits must not change any output: every
schedule/ordering combination gives bit-identical trajectories and step
statistics, including for orbits that start at a singular point.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import pytest

from astrodrift_amd import _native, workloads

pytestmark = pytest.mark.gpu


def _run(w, method, schedule, ordering, precision=_native.PRECISION_FP64, state0=None):
    import torch

    dev = torch.device("cuda:0")
    y0 = w.state0 if state0 is None else state0
    n = y0.shape[0]
    s0 = torch.from_numpy(np.ascontiguousarray(y0)).to(dev)
    tt = torch.from_numpy(w.times).to(dev)
    out = torch.full((w.nt, n, 6), float("nan"), dtype=torch.float64, device=dev)
    st = torch.zeros((n, 4), dtype=torch.int32, device=dev)
    _native.launch_device(method, s0.data_ptr(), tt.data_ptr(), out.data_ptr(), n, w.nt,
                          math.log(w.rtol), math.log(w.atol), w.potential, st.data_ptr(),
                          precision=precision, schedule=schedule, ordering=ordering)
    torch.cuda.synchronize()
    return out.cpu().numpy(), st.cpu().numpy()


def _same(a, b):
    # bitwise, NaN included
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


COMBOS = [
    (_native.SCHEDULE_PERSISTENT, _native.ORDER_COST),
    (_native.SCHEDULE_PERSISTENT, _native.ORDER_AUTO),
    (_native.SCHEDULE_PLAIN, _native.ORDER_COST),
    (_native.SCHEDULE_PLAIN, _native.ORDER_INDEX),
]


def test_mw_dop853_every_order_is_bit_identical(gpu):
    """20,000 Milky-Way orbits (above the AUTO threshold), t = 0..10, 201 outputs."""
    w = workloads.mw_disk(n=20000, t_end=10.0, steps=201)
    ref_out, ref_st = _run(w, _native.METHOD_DOP853, _native.SCHEDULE_PERSISTENT,
                           _native.ORDER_INDEX)
    assert not np.isnan(ref_out).any()
    for schedule, ordering in COMBOS:
        out, st = _run(w, _native.METHOD_DOP853, schedule, ordering)
        print(f"schedule={schedule} ordering={ordering}: accepted={int(st[:, 0].sum())}")
        assert _same(out, ref_out), (schedule, ordering)
        assert np.array_equal(st, ref_st), (schedule, ordering)


def test_three_wave_plain_kernel_is_bit_identical(gpu):
    """A plain FP64 Milky-Way launch large enough to fill the block slots twice
    at three waves per SIMD (400,000 orbits on 256 CUs) runs the three-wave
    plain kernel; its trajectories and statistics equal the persistent
    schedule's, bit for bit, in index and in cost order."""
    w = workloads.mw_disk(n=400000, t_end=1.0, steps=21)
    ref_out, ref_st = _run(w, _native.METHOD_DOP853, _native.SCHEDULE_PERSISTENT,
                           _native.ORDER_INDEX)
    assert not np.isnan(ref_out).any()
    for ordering in (_native.ORDER_INDEX, _native.ORDER_COST):
        out, st = _run(w, _native.METHOD_DOP853, _native.SCHEDULE_PLAIN, ordering)
        assert _same(out, ref_out), ordering
        assert np.array_equal(st, ref_st), ordering


def test_kepler_dopr54_cost_order_is_bit_identical(gpu):
    w = workloads.kepler_disk(n=20000)
    ref_out, ref_st = _run(w, _native.METHOD_DOPR54, _native.SCHEDULE_PLAIN, _native.ORDER_INDEX)
    for schedule, ordering in COMBOS[:3]:
        out, st = _run(w, _native.METHOD_DOPR54, schedule, ordering)
        assert _same(out, ref_out), (schedule, ordering)
        assert np.array_equal(st, ref_st), (schedule, ordering)


def test_mixed_precision_cost_order_is_bit_identical(gpu):
    w = workloads.mw_disk(n=20000, t_end=10.0, steps=101, rtol=1e-7, atol=1e-7)
    ref_out, ref_st = _run(w, _native.METHOD_DOP853, _native.SCHEDULE_PERSISTENT,
                           _native.ORDER_INDEX, precision=_native.PRECISION_MIXED)
    out, st = _run(w, _native.METHOD_DOP853, _native.SCHEDULE_PERSISTENT, _native.ORDER_COST,
                   precision=_native.PRECISION_MIXED)
    assert _same(out, ref_out)
    assert np.array_equal(st, ref_st)


def test_singular_and_free_starts_are_ordered_last_without_changing_results(gpu):
    """Orbits at the origin of the galpy Kepler force (NaN force) and orbits far
    out (tiny force) mixed into a normal batch."""
    w = workloads.kepler_disk(n=4096)
    w.potential = None  # the reference's unit galpy Kepler force
    y0 = w.state0.copy()
    y0[::97, :3] = 0.0         # at the singularity
    y0[5::101, :3] *= 1e6      # nearly free drifts
    ref_out, ref_st = _run(w, _native.METHOD_DOPR54, _native.SCHEDULE_PERSISTENT,
                           _native.ORDER_INDEX, state0=y0)
    out, st = _run(w, _native.METHOD_DOPR54, _native.SCHEDULE_PERSISTENT, _native.ORDER_COST,
                   state0=y0)
    assert _same(out, ref_out)
    assert np.array_equal(st, ref_st)


def test_bad_ordering_options_are_rejected(gpu):
    import torch

    w = workloads.kepler_disk(n=64)
    dev = torch.device("cuda:0")
    s0 = torch.from_numpy(w.state0).to(dev)
    tt = torch.from_numpy(w.times).to(dev)
    out = torch.empty((w.nt, w.n, 6), dtype=torch.float64, device=dev)
    with pytest.raises(_native.DriftNativeError, match="ordering"):
        _native.launch_device(_native.METHOD_DOPR54, s0.data_ptr(), tt.data_ptr(), out.data_ptr(),
                              w.n, w.nt, -23.0, -23.0, ordering=3)
    s = _native.session(0)
    opts = _native.CLaunchOptions(0, 0, 1, 0, 0, 7)
    rc = s.lib.drift_galpy_dopr54_device(
        s.handle, ctypes.c_void_p(s0.data_ptr()), ctypes.c_void_p(tt.data_ptr()),
        ctypes.c_void_p(out.data_ptr()), w.n, w.nt, -23.0, -23.0, _native.INITIAL_STEP_SENTINEL,
        None, None, ctypes.byref(opts), None)
    assert rc == _native.DRIFT_ERR_INVALID_ARGUMENT
    assert "staging" in _native.last_error()


def test_autotuner_choice_is_valid_and_tuned_launches_match(gpu):
    """drift_tune_device picks one of its candidates; SCHEDULE_TUNED launches
    (with and without a stored choice) give the same bits as the plain kernel."""
    import torch

    w = workloads.mw_disk(n=20000, t_end=10.0, steps=201)
    dev = torch.device("cuda:0")
    s0 = torch.from_numpy(w.state0).to(dev)
    tt = torch.from_numpy(w.times).to(dev)
    ref_out, ref_st = _run(w, _native.METHOD_DOP853, _native.SCHEDULE_PLAIN, _native.ORDER_INDEX)
    # no stored choice yet for this size: TUNED falls back to persistent + AUTO
    out, st = _run(w, _native.METHOD_DOP853, _native.SCHEDULE_TUNED, _native.ORDER_AUTO)
    assert _same(out, ref_out) and np.array_equal(st, ref_st)
    lt = math.log(w.rtol)
    best = _native.tune_device(_native.METHOD_DOP853, s0.data_ptr(), tt.data_ptr(), w.n, w.nt,
                               lt, lt, w.potential)
    print("tuned:", best)
    assert best["schedule"] in (_native.SCHEDULE_PLAIN, _native.SCHEDULE_PERSISTENT)
    assert best["block_size"] in (64, 128, 256, 512)
    assert best["ordering"] in (_native.ORDER_INDEX, _native.ORDER_COST)
    assert best["ms"] > 0.0
    out, st = _run(w, _native.METHOD_DOP853, _native.SCHEDULE_TUNED, _native.ORDER_AUTO)
    assert _same(out, ref_out) and np.array_equal(st, ref_st)


def test_autotuner_takes_staged_output_on_a_sample_dense_grid(gpu):
    """At rtol 1e-8 over t = 0..2 with 801 outputs every DOP853 step emits
    many samples, so the tuner must choose staged output whatever its prefix
    timings say; TUNED launches with that choice give the plain kernel's bits."""
    import torch

    w = workloads.mw_disk(n=5000, t_end=2.0, steps=801, rtol=1e-8, atol=1e-8)
    dev = torch.device("cuda:0")
    s0 = torch.from_numpy(w.state0).to(dev)
    tt = torch.from_numpy(w.times).to(dev)
    ref_out, ref_st = _run(w, _native.METHOD_DOP853, _native.SCHEDULE_PLAIN, _native.ORDER_INDEX)
    samples_per_step = w.n * (w.nt - 1) / float(ref_st[:, 0].sum())
    print(f"samples per accepted step over the whole grid: {samples_per_step:.2f}")
    assert samples_per_step > 3.0
    lt = math.log(w.rtol)
    best = _native.tune_device(_native.METHOD_DOP853, s0.data_ptr(), tt.data_ptr(), w.n, w.nt,
                               lt, lt, w.potential)
    print("tuned:", best)
    assert best["staging"] == _native.STAGING_STAGED
    assert best["ms"] > 0.0
    out, st = _run(w, _native.METHOD_DOP853, _native.SCHEDULE_TUNED, _native.ORDER_AUTO)
    assert _same(out, ref_out) and np.array_equal(st, ref_st)


def test_autotuner_rejects_bad_arguments(gpu):
    with pytest.raises(_native.DriftNativeError, match="method"):
        _native.tune_device(7, 0, 0, 10, 10, -23.0, -23.0)
    with pytest.raises(_native.DriftNativeError, match="nt"):
        _native.tune_device(_native.METHOD_DOP853, 0, 0, 10, 1, -23.0, -23.0)
