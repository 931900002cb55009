"""Packed dense output (DOP853 in 512-thread blocks).

A lane hands an accepted step's dense-output inputs to its wavefront's LDS
stash, and whichever lane later runs the request writes the samples.  Every
output and every step statistic must equal the inline form's (256-thread
blocks) bit for bit: for each static potential, in FP64 and in the mixed mode,
for the plain and the persistent schedule, on grids with one and with several
samples per step, and for orbits that end non-finite.  The statistics arrays
start as garbage: a packed launch zeroes their status words itself before it
ORs statuses in.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from astrodrift_amd import _native, workloads

pytestmark = pytest.mark.gpu


def _run(w, schedule, block, precision=_native.PRECISION_FP64, state0=None, rtol=None):
    import torch

    dev = torch.device("cuda:0")
    y0 = w.state0 if state0 is None else state0
    n = y0.shape[0]
    s0 = torch.from_numpy(np.ascontiguousarray(y0)).to(dev)
    tt = torch.from_numpy(w.times).to(dev)
    out = torch.full((w.nt, n, 6), float("nan"), dtype=torch.float64, device=dev)
    # garbage in the status words: the packed launch clears them itself
    st = torch.full((n, 4), -1, dtype=torch.int32, device=dev)
    lt = math.log(rtol if rtol is not None else w.rtol)
    _native.launch_device(_native.METHOD_DOP853, s0.data_ptr(), tt.data_ptr(), out.data_ptr(), n,
                          w.nt, lt, lt, w.potential, st.data_ptr(), precision=precision,
                          schedule=schedule, block_size=block)
    torch.cuda.synchronize()
    return out.cpu().numpy(), st.cpu().numpy()


def _same(a, b):
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


SCHEDULES = [_native.SCHEDULE_PERSISTENT, _native.SCHEDULE_PLAIN]


@pytest.mark.parametrize("rtol", [1e-10, 1e-7])
@pytest.mark.parametrize("precision", [_native.PRECISION_FP64, _native.PRECISION_MIXED])
def test_packed_equals_inline_mw(gpu, precision, rtol):
    """20,000 Milky-Way orbits (cost ordered), t = 0..10, 201 outputs: at rtol
    1e-10 a step emits under one sample on average, at 1e-7 several."""
    if precision == _native.PRECISION_MIXED and rtol < 1e-7:
        pytest.skip("the mixed mode needs rtol >= 1e-7")
    w = workloads.mw_disk(n=20000, t_end=10.0, steps=201)
    for schedule in SCHEDULES:
        ref_out, ref_st = _run(w, schedule, 256, precision, rtol=rtol)
        out, st = _run(w, schedule, 512, precision, rtol=rtol)
        assert not np.isnan(ref_out).any()
        assert _same(out, ref_out), schedule
        assert np.array_equal(st, ref_st), schedule
        assert (st[:, 3] == 0).all()


@pytest.mark.parametrize("kind", ["galpy", "kepler", "plummer"])
def test_packed_equals_inline_point_masses(gpu, kind):
    """The 1,024-orbit Kepler disk under each point-mass force (no LDS table:
    the stashes take 56 requests a wavefront), over two orbital periods."""
    k = workloads.kepler_disk(n=1024)
    spec = {"galpy": None, "kepler": _native.PotentialSpec.kepler(1.0),
            "plummer": _native.PotentialSpec.plummer(1.0, 0.1)}[kind]
    w = workloads.Workload("kepler_disk_dop853", "DOP853", k.state0,
                           np.linspace(0.0, 4.0 * math.pi, 301), 1e-10, 1e-10, spec)
    for schedule in SCHEDULES:
        ref_out, ref_st = _run(w, schedule, 256)
        out, st = _run(w, schedule, 512)
        assert _same(out, ref_out), schedule
        assert np.array_equal(st, ref_st), schedule


def test_packed_non_finite_orbits_keep_their_flag(gpu):
    """Orbits that start at the galpy Kepler singularity (R = 0: NaN x/y
    accelerations) among regular ones, so that wavefronts hold both finished
    and running orbits while requests are stashed: the packed launch flags the
    same orbits and writes the same NaN samples as the inline one."""
    k = workloads.kepler_disk(n=2048)
    y0 = k.state0.copy()
    y0[::97, 0] = 0.0
    y0[::97, 1] = 0.0
    w = workloads.Workload("kepler_disk_singular", "DOP853", y0, np.linspace(0.0, 6.0, 121),
                           1e-9, 1e-9, None)
    for schedule in SCHEDULES:
        ref_out, ref_st = _run(w, schedule, 256)
        out, st = _run(w, schedule, 512)
        assert (ref_st[::97, 3] & _native.STATUS_NONFINITE).all()
        assert _same(out, ref_out), schedule
        assert np.array_equal(st, ref_st), schedule


def test_packed_is_the_persistent_default(gpu):
    """block_size 0: persistent DOP853 launches with direct output run packed
    (512-thread blocks), plain ones and DOPR54 inline (256)."""
    w = workloads.mw_disk(n=4096)
    _, block = _native.launch_shape(_native.METHOD_DOP853, w.n, w.potential,
                                    schedule=_native.SCHEDULE_PERSISTENT)
    assert block == 512
    _, block = _native.launch_shape(_native.METHOD_DOP853, w.n, w.potential,
                                    schedule=_native.SCHEDULE_PLAIN)
    assert block == 256
    _, block = _native.launch_shape(_native.METHOD_DOPR54, w.n, w.potential,
                                    schedule=_native.SCHEDULE_PERSISTENT)
    assert block == 256


def test_default_schedules_of_the_host_buffer_entry_points(gpu):
    """No options: DOP853 launches persistent (a grid of resident blocks),
    DOPR54 plain (one lane per orbit, ceil(n / 256) blocks)."""
    n = 1 << 20
    w = workloads.mw_disk(n=256, t_end=1.0, steps=11)
    grid, block = _native.launch_shape(_native.METHOD_DOPR54, n, w.potential)
    assert (grid, block) == ((n + 255) // 256, 256)
    grid, block = _native.launch_shape(_native.METHOD_DOP853, n, w.potential)
    assert block == 512 and grid < n // 512


def test_packed_is_refused_where_it_does_not_exist(gpu):
    """block_size 512 needs DOP853 and a static potential."""
    w = workloads.mw_disk(n=256, t_end=1.0, steps=11)
    with pytest.raises(_native.DriftNativeError, match="block_size 512"):
        _native.launch_shape(_native.METHOD_DOPR54, w.n, w.potential,
                             schedule=_native.SCHEDULE_PERSISTENT, block_size=512)
