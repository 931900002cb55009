"""Dispatch log records through the Python API: the GPU cases of the
reference's logging tests (tests/test_python_logging.py:64-127), with the
reference's logger names and messages.  The backend logger of this build is
drift.dispatch.gpu.hip (the reference's are .cuda_oxide and .rust_cuda).  The
reference's `record.filename` asserts name its Rust source files and are not
carried over."""

from __future__ import annotations

import logging
import re
from collections.abc import Iterator

import numpy as np
import pytest

import astrodrift_amd as dft

pytestmark = pytest.mark.gpu

GPU_LOGGER_NAME = "drift.dispatch.gpu"
GPU_BACKEND_LOGGERS = {f"{GPU_LOGGER_NAME}.hip"}
N_PARTICLES = 50
INITIAL_STATE = np.tile(np.array([1.0, 0.0, 0.0, 0.0, 1.0, 0.0], dtype=np.float64),
                        (N_PARTICLES, 1))


def _config(method=None) -> dft.Config:
    kepler = dft.Potential.kepler(1.0)
    kw = {} if method is None else {"method": method}
    sim = dft.Config(engine=dft.Engine.GPU, ts=(0.0, 2.0 * np.pi, 16), **kw)
    sim.add(dft.test_particles(INITIAL_STATE), dft.background(kepler))
    return sim


@pytest.fixture(autouse=True)
def restore_drift_logger() -> Iterator[None]:
    logger = logging.getLogger("drift")
    saved = (logger.level, logger.propagate, list(logger.handlers))
    try:
        yield
    finally:
        logger.setLevel(saved[0])
        logger.propagate = saved[1]
        logger.handlers[:] = saved[2]


def test_gpu_run_emits_info_records_on_the_drift_logger(gpu, caplog) -> None:
    with caplog.at_level(logging.INFO, logger="drift"):
        _config().run()
    records = [r for r in caplog.records if r.name == GPU_LOGGER_NAME]
    messages = [r.getMessage() for r in records]
    assert any(m.startswith("GPU integration starting") for m in messages)
    assert any(m.startswith("GPU integration finished") for m in messages)
    assert all(r.levelno == logging.INFO for r in records)
    # no debug record leaks at INFO
    assert not any(r.name in GPU_BACKEND_LOGGERS for r in caplog.records)


def test_gpu_debug_records_include_dispatch_and_backend_phases(gpu, caplog) -> None:
    with caplog.at_level(logging.DEBUG, logger="drift"):
        _config().run()
    dispatch = [r.getMessage() for r in caplog.records if r.name == GPU_LOGGER_NAME]
    assert any(m.startswith("launching GALPY DOPR54 kernel") for m in dispatch)
    assert any(m.startswith("stage 1: dispatching") for m in dispatch)
    assert any(m.startswith("stage 1: 50 particles finished") for m in dispatch)

    backend = [r for r in caplog.records if r.name in GPU_BACKEND_LOGGERS]
    assert len({r.name for r in backend}) == 1
    backend_messages = [r.getMessage() for r in backend]
    assert any("context initialized and kernel module loaded" in m for m in backend_messages)
    assert any("bytes of host data to the device" in m for m in backend_messages)
    assert any("result bytes to the host" in m for m in backend_messages)


def test_launch_record_carries_the_launch_shape_and_byte_counts(gpu, caplog) -> None:
    """The launch record names the grid and block the integrator ran with (the
    persistent schedule: resident blocks, capped by n), and the backend's byte
    counts are the inputs and the (nt, n, 6) trajectories."""
    with caplog.at_level(logging.DEBUG, logger="drift"):
        _config(dft.Method.DOP853).run()
    msgs = [r.getMessage() for r in caplog.records]
    launch = [m for m in msgs if m.startswith("launching GALPY DOP853 kernel")]
    assert len(launch) == 1, msgs
    m = re.fullmatch(r"launching GALPY DOP853 kernel: 50 particles, grid=(\d+), block=(\d+), "
                     r"16 output times", launch[0])
    assert m, launch[0]
    grid, block = int(m.group(1)), int(m.group(2))
    # 50 orbits: one block; persistent DOP853 runs its packed dense output in
    # 512-thread blocks
    assert block == 512 and grid == 1
    ran = [x for x in msgs if x.startswith("integrated on the device")]
    assert ran and f"grid={grid}, block={block}" in ran[0]
    assert any(x.startswith(f"copied {(50 * 6 + 16) * 8} bytes of host data") for x in msgs)
    assert any(f"copied {16 * 50 * 6 * 8} result bytes to the host" in x for x in msgs)
    order = [i for i, x in enumerate(msgs) if x.startswith(("launching GALPY", "stage 1: 50"))]
    assert len(order) == 2 and msgs[order[0]].startswith("launching")


def test_standard_python_levels_silence_the_records(gpu, caplog) -> None:
    with caplog.at_level(logging.WARNING, logger="drift"):
        _config().run()
    assert [r for r in caplog.records if r.name.startswith(GPU_LOGGER_NAME)] == []
