"""Python threads sharing the process-wide session of a device.

ctypes releases the GIL during a foreign call, and a libdrift_hip session may
be used by one thread at a time; the Python layer serializes calls on its
session.  Launches from several threads at once must give the bits the same
launches give one after another."""

from __future__ import annotations

import math
import threading

import numpy as np
import pytest

from astrodrift_amd import _native, workloads

pytestmark = pytest.mark.gpu


def _disk(spec, n, seed):
    """Near-circular orbits of `spec` (circular speed from the device force)."""
    rng = np.random.default_rng(seed)
    R = rng.uniform(0.5, 2.0, n)
    phi = rng.uniform(0.0, 2.0 * math.pi, n)
    acc = _native.force_host(spec, np.stack([R, np.zeros(n), np.zeros(n)], 1))
    vc = np.sqrt(-R * acc[:, 0])
    vphi = vc * (1.0 + rng.normal(0.0, 0.05, n))
    vR, vz = rng.normal(0.0, 0.05, n) * vc, rng.normal(0.0, 0.05, n) * vc
    z = rng.normal(0.0, 0.03, n)
    c, s = np.cos(phi), np.sin(phi)
    return np.ascontiguousarray(np.stack([R * c, R * s, z, vR * c - vphi * s, vR * s + vphi * c,
                                          vz], 1))


def _cases():
    cases = []
    for k, (method, spec) in enumerate([
            (_native.METHOD_DOP853, _native.PotentialSpec.milky_way()),
            (_native.METHOD_DOPR54, _native.PotentialSpec.kepler(1.0)),
            (_native.METHOD_DOP853, _native.PotentialSpec.plummer(1.0, 0.1)),
            (_native.METHOD_DOPR54, None)]):
        cases.append((method, spec, _disk(spec, 300 + 37 * k, 100 + k),
                      workloads.grid_times(0.0, 5.0, 51)))
    return cases


def _launch(case):
    method, spec, y0, t = case
    tol = math.log(1e-10)
    out, stats, code = _native.launch_host(method, y0, t, tol, tol, spec,
                                           raise_on_integration_error=False)
    return out, stats, code


def test_threads_sharing_a_session_match_sequential_launches(gpu):
    cases = _cases()
    expected = [_launch(c) for c in cases]
    pos = np.random.default_rng(3).normal(0.0, 1.0, (500, 3))
    force_expected = _native.force_host(_native.PotentialSpec.milky_way(), pos)
    results: dict[tuple[int, int], tuple] = {}
    errors: list[BaseException] = []

    def worker(tid: int) -> None:
        try:
            for rep in range(3):
                k = (tid + rep) % len(cases)
                results[(tid, rep)] = (k, _launch(cases[k]),
                                       _native.force_host(_native.PotentialSpec.milky_way(), pos))
        except BaseException as e:  # reported below, with the thread's case
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(tid,)) for tid in range(6)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=100)
    assert not any(th.is_alive() for th in threads)
    assert not errors, errors
    assert len(results) == 18
    for (tid, rep), (k, (out, stats, code), acc) in results.items():
        e_out, e_stats, e_code = expected[k]
        assert code == e_code == 0, (tid, rep, k)
        assert np.array_equal(out.view(np.uint64), e_out.view(np.uint64)), (tid, rep, k)
        assert np.array_equal(stats, e_stats), (tid, rep, k)
        assert np.array_equal(acc, force_expected), (tid, rep)


def test_released_session_is_replaced_by_a_new_one(gpu):
    """release_session frees the process-wide session; the next call builds a
    new one whose launches give the same bits."""
    w = workloads.mw_disk(n=2000, t_end=2.0, steps=21)
    lt = math.log(w.rtol)
    first = _native.launch_host(_native.METHOD_DOP853, w.state0, w.times, lt, lt, w.potential)
    old = _native.session(0)
    _native.release_session(0)
    assert not old.handle
    again = _native.launch_host(_native.METHOD_DOP853, w.state0, w.times, lt, lt, w.potential)
    assert _native.session(0) is not old
    assert np.array_equal(np.asarray(first[0]).view(np.uint64), np.asarray(again[0]).view(np.uint64))
