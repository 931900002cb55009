"""Mixed precision (FP32 stage forces; FP64 state, error norm and control).

The Milky-Way DOP853 workload at rtol = atol = 1e-7, 50 time units: the
mixed-precision run against the pure-FP64 run at the same tolerance and an
FP64 rtol = 1e-13 run (itself checked against the oracle on a subset).
Reports position relative error and |dE/E| at t = 50; the mixed mode's
energy error must stay within 10x of FP64's.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from astrodrift_amd import _native, workloads

import oracle_lib

pytestmark = pytest.mark.gpu


def run(torch, w, tol, precision, schedule=_native.SCHEDULE_PERSISTENT):
    dev = torch.device("cuda:0")
    s0 = torch.from_numpy(w.state0).to(dev)
    tt = torch.from_numpy(w.times).to(dev)
    out = torch.empty((w.nt, w.n, 6), dtype=torch.float64, device=dev)
    st = torch.zeros((w.n, 4), dtype=torch.int32, device=dev)
    lt = math.log(tol)
    _native.launch_device(_native.METHOD_DOP853, s0.data_ptr(), tt.data_ptr(), out.data_ptr(),
                          w.n, w.nt, lt, lt, w.potential, st.data_ptr(), precision=precision,
                          schedule=schedule)
    torch.cuda.synchronize()
    return out.cpu().numpy(), st.cpu().numpy()


def energy(spec, states):
    return np.array([0.5 * (s[3] ** 2 + s[4] ** 2 + s[5] ** 2)
                     + oracle_lib.potential_value(spec, s[0], s[1], s[2]) for s in states])


def test_mixed_precision_error_budget(gpu):
    torch = pytest.importorskip("torch")
    full = workloads.mw_disk(n=1 << 20, t_end=50.0, steps=11, rtol=1e-7, atol=1e-7)
    idx = np.arange(0, full.n, full.n // 4096)
    w = workloads.Workload("mw_subset", "DOP853", np.ascontiguousarray(full.state0[idx]),
                           full.times, 1e-7, 1e-7, full.potential)
    truth, _ = run(torch, w, 1e-13, _native.PRECISION_FP64)
    # the rtol = 1e-13 truth agrees with the oracle on a slice
    ref, _ = oracle_lib.batch(1, w.state0[:64], w.times, math.log(1e-13), math.log(1e-13),
                              w.potential, nthreads=8)
    assert np.max(np.abs(truth[:, :64] - ref) / np.maximum(np.abs(ref), 1e-3)) < 1e-8
    fp64, st64 = run(torch, w, 1e-7, _native.PRECISION_FP64)
    mixed, stmx = run(torch, w, 1e-7, _native.PRECISION_MIXED)
    with pytest.raises(_native.DriftNativeError, match="1e-7"):
        run(torch, w, 1e-9, _native.PRECISION_MIXED)

    def pos_err(a):
        return np.linalg.norm(a[-1, :, :3] - truth[-1, :, :3], axis=1) / \
            np.linalg.norm(truth[-1, :, :3], axis=1)

    e0 = energy(w.potential, w.state0)
    dE64 = np.abs(energy(w.potential, fp64[-1]) - e0) / np.abs(e0)
    dEmx = np.abs(energy(w.potential, mixed[-1]) - e0) / np.abs(e0)
    print(f"t=50, n={w.n}: position rel err median/max FP64 {np.median(pos_err(fp64)):.2e}/"
          f"{pos_err(fp64).max():.2e}, mixed {np.median(pos_err(mixed)):.2e}/"
          f"{pos_err(mixed).max():.2e}; |dE/E| median/max FP64 {np.median(dE64):.2e}/"
          f"{dE64.max():.2e}, mixed {np.median(dEmx):.2e}/{dEmx.max():.2e}; accepted steps "
          f"FP64 {st64[:, 0].sum()} mixed {stmx[:, 0].sum()}")
    assert np.median(dEmx) < 10 * np.median(dE64)
    assert np.median(pos_err(mixed)) < 10 * np.median(pos_err(fp64)) + 1e-6
