"""The allocation-free DOP853 entry point against the session entry point.

drift_galpy_dop853_device_noalloc keeps its scratch (work-queue counter, cost
ordering, Milky-Way tables) in a fixed library array per GPU.  Called through
the C ABI twice in a row, the second time with fewer orbits, it must give the
bits of drift_galpy_dop853_device with the same options; it refuses sizes
beyond its scratch with an error and launches nothing."""

from __future__ import annotations

import ctypes
import math

import numpy as np
import pytest

from astrodrift_amd import _native, workloads

pytestmark = pytest.mark.gpu


def _entry():
    lib = _native.load_library()
    f = lib.drift_galpy_dop853_device_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                  ctypes.c_size_t, ctypes.c_double, ctypes.c_double,
                  ctypes.POINTER(_native.CPotential), ctypes.c_void_p,
                  ctypes.POINTER(_native.CLaunchOptions), ctypes.c_void_p]
    return f


def _bits(a):
    return a.view(np.uint64)


def _pair(w, schedule, ordering, precision=_native.PRECISION_FP64):
    """(session entry point, allocation-free entry point) outputs and stats."""
    import torch

    dev = torch.device("cuda:0")
    s0 = torch.from_numpy(w.state0).to(dev)
    tt = torch.from_numpy(w.times).to(dev)
    lt = math.log(w.rtol)
    stream = torch.cuda.current_stream().cuda_stream
    outs = []
    for which in ("session", "noalloc"):
        out = torch.full((w.nt, w.n, 6), float("nan"), dtype=torch.float64, device=dev)
        st = torch.zeros((w.n, 4), dtype=torch.int32, device=dev)
        torch.cuda.synchronize()
        print(f"{which}: n={w.n} nt={w.nt} schedule={schedule} ordering={ordering} "
              f"precision={precision}", flush=True)
        if which == "session":
            _native.launch_device(_native.METHOD_DOP853, s0.data_ptr(), tt.data_ptr(),
                                  out.data_ptr(), w.n, w.nt, lt, lt, w.potential, st.data_ptr(),
                                  precision=precision, schedule=schedule, ordering=ordering,
                                  stream=stream)
        else:
            pot = w.potential.to_c()
            opts = _native.CLaunchOptions(precision, 0, schedule, 0, ordering, 0)
            code = _entry()(s0.data_ptr(), tt.data_ptr(), out.data_ptr(), w.n, w.nt, lt, lt,
                            ctypes.byref(pot), st.data_ptr(), ctypes.byref(opts), stream)
            _native.check(code)
        torch.cuda.synchronize()
        outs.append((out.cpu().numpy(), st.cpu().numpy()))
    return outs


@pytest.mark.parametrize("schedule", [_native.SCHEDULE_PERSISTENT, _native.SCHEDULE_PLAIN])
def test_noalloc_matches_session_entry_on_two_calls(gpu, schedule):
    # first call: cost-ordered (AUTO from 16384 orbits); second: fewer orbits
    for n in (20000, 17000):
        w = workloads.mw_disk(n=n, seed=5, t_end=10.0, steps=101, rtol=1e-10, atol=1e-10)
        (ref, ref_st), (got, got_st) = _pair(w, schedule, _native.ORDER_AUTO)
        assert not np.isnan(got).any()
        assert np.array_equal(_bits(got), _bits(ref)), n
        assert np.array_equal(got_st, ref_st), n


def test_noalloc_index_order_and_mixed_precision(gpu):
    """Index order, and the mixed mode (the float table in the same array)
    after an FP64 call with the same potential."""
    w = workloads.mw_disk(n=5000, seed=9, t_end=10.0, steps=51, rtol=1e-7, atol=1e-7)
    for precision in (_native.PRECISION_FP64, _native.PRECISION_MIXED):
        (ref, ref_st), (got, got_st) = _pair(w, _native.SCHEDULE_PERSISTENT, _native.ORDER_INDEX,
                                             precision)
        assert np.array_equal(_bits(got), _bits(ref)), precision
        assert np.array_equal(got_st, ref_st), precision


def test_noalloc_refuses_sizes_beyond_its_scratch(gpu):
    import torch

    n = (1 << 20) + 1
    dev = torch.device("cuda:0")
    s0 = torch.zeros((n, 6), dtype=torch.float64, device=dev)
    tt = torch.tensor([0.0, 1.0], dtype=torch.float64, device=dev)
    out = torch.empty((2, n, 6), dtype=torch.float64, device=dev)
    pot = _native.PotentialSpec.kepler(1.0).to_c()
    opts = _native.CLaunchOptions(0, 0, _native.SCHEDULE_PERSISTENT, 0, _native.ORDER_COST, 0)
    code = _entry()(s0.data_ptr(), tt.data_ptr(), out.data_ptr(), n, 2, -23.0, -23.0,
                    ctypes.byref(pot), None, ctypes.byref(opts), None)
    assert code == _native.DRIFT_ERR_INVALID_ARGUMENT
    assert "at most 1048576 orbits" in _native.last_error()
    opts.staging = _native.STAGING_STAGED
    opts.ordering = _native.ORDER_INDEX
    code = _entry()(s0.data_ptr(), tt.data_ptr(), out.data_ptr(), n, 2, -23.0, -23.0,
                    ctypes.byref(pot), None, ctypes.byref(opts), None)
    assert code == _native.DRIFT_ERR_UNSUPPORTED
