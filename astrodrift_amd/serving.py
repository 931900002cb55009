"""Continuous-batching server over libdrift_hip's serving loop.

A :class:`Server` wraps drift_server_create / submit / wait (include/drift_hip.h):
requests (host arrays) queue first-come first-served, are admitted while their
device footprint fits the HBM budget, and advance together one output chunk
per ragged launch with their orbit state resident on the device.  Requests
that arrive while others run join at the next chunk boundary.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

from . import _native


class CRequest(ctypes.Structure):
    _fields_ = [
        ("method", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("potential", _native.CPotential),
        ("n", ctypes.c_size_t),
        ("nt", ctypes.c_size_t),
        ("rtol", ctypes.c_double),
        ("atol", ctypes.c_double),
        ("dt_one_init", ctypes.c_double),
        ("state0", ctypes.c_void_p),
        ("times", ctypes.c_void_p),
        ("state_out", ctypes.c_void_p),
        ("stats", ctypes.c_void_p),
    ]


class CServerOptions(ctypes.Structure):
    _fields_ = [
        ("hbm_budget_bytes", ctypes.c_uint64),
        ("chunk_samples", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("block_size", ctypes.c_int32),
        ("max_attempts", ctypes.c_uint32),
        ("chunk_target_ms", ctypes.c_double),
    ]


class CRequestInfo(ctypes.Structure):
    _fields_ = [
        ("state", ctypes.c_int32),
        ("status", ctypes.c_int32),
        ("submit_s", ctypes.c_double),
        ("admit_s", ctypes.c_double),
        ("complete_s", ctypes.c_double),
        ("accepted_steps", ctypes.c_uint64),
        ("chunks", ctypes.c_uint64),
    ]


QUEUED, RUNNING, DONE, FAILED = 0, 1, 2, 3


def _declare(lib) -> None:
    c = ctypes
    if getattr(lib, "_serving_declared", False):
        return
    lib.drift_server_create.restype = c.c_int
    lib.drift_server_create.argtypes = [c.c_int, c.POINTER(CServerOptions), c.POINTER(c.c_void_p)]
    lib.drift_server_submit.restype = c.c_int
    lib.drift_server_submit.argtypes = [c.c_void_p, c.POINTER(CRequest), c.POINTER(c.c_uint64)]
    lib.drift_server_query.restype = c.c_int
    lib.drift_server_query.argtypes = [c.c_void_p, c.c_uint64, c.POINTER(CRequestInfo)]
    lib.drift_server_wait.restype = c.c_int
    lib.drift_server_wait.argtypes = [c.c_void_p, c.c_uint64, c.c_double, c.POINTER(CRequestInfo)]
    lib.drift_server_stats.restype = c.c_int
    lib.drift_server_stats.argtypes = [c.c_void_p, c.POINTER(c.c_uint64), c.POINTER(c.c_uint64),
                                       c.POINTER(c.c_uint64)]
    lib.drift_server_busy.restype = c.c_int
    lib.drift_server_busy.argtypes = [c.c_void_p, c.POINTER(c.c_double), c.POINTER(c.c_double),
                                      c.POINTER(c.c_uint64)]
    lib.drift_server_destroy.restype = c.c_int
    lib.drift_server_destroy.argtypes = [c.c_void_p]
    lib._serving_declared = True


@dataclass
class Request:
    """A submitted request and the host buffers the server writes into."""

    ticket: int
    method: int
    state0: np.ndarray
    times: np.ndarray
    out: np.ndarray
    stats: np.ndarray


class Server:
    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0, chunk_samples: int = 0,
                 precision: int = _native.PRECISION_FP64, block_size: int = 0,
                 max_attempts: int = 0, chunk_target_ms: float = 0.0):
        """chunk_samples = 0 adapts the chunk length to chunk_target_ms (0 = 10 ms)."""
        self.lib = _native.load_library()
        _declare(self.lib)
        opts = CServerOptions(hbm_budget_bytes, chunk_samples, precision, block_size, max_attempts,
                              chunk_target_ms)
        handle = ctypes.c_void_p()
        _native.check(self.lib.drift_server_create(device, ctypes.byref(opts),
                                                   ctypes.byref(handle)))
        self.handle = handle
        self._live: dict[int, Request] = {}

    def submit(self, method: int, state0, times, rtol: float, atol: float,
               potential: _native.PotentialSpec | None = None,
               dt_one: float = _native.INITIAL_STEP_SENTINEL, out: np.ndarray | None = None
               ) -> Request:
        """Queues one request (linear tolerances); returns its Request handle."""
        y0 = np.ascontiguousarray(state0, dtype=np.float64).reshape(-1, 6)
        t = np.ascontiguousarray(times, dtype=np.float64)
        n, nt = y0.shape[0], t.shape[0]
        if out is None:
            out = np.empty((nt, n, 6), dtype=np.float64)
        stats = np.zeros(n, dtype=_native.STATS_DTYPE)
        pot = (potential or _native.PotentialSpec.galpy_kepler()).to_c()
        req = CRequest(method, 0, pot, n, nt, math.log(rtol), math.log(atol), dt_one,
                       y0.ctypes.data, t.ctypes.data, out.ctypes.data, stats.ctypes.data)
        ticket = ctypes.c_uint64()
        _native.check(self.lib.drift_server_submit(self.handle, ctypes.byref(req),
                                                   ctypes.byref(ticket)))
        r = Request(ticket.value, method, y0, t, out, stats)
        self._live[r.ticket] = r
        return r

    def wait(self, request: Request, timeout_s: float = -1.0) -> CRequestInfo:
        info = CRequestInfo()
        _native.check(self.lib.drift_server_wait(self.handle, request.ticket, timeout_s,
                                                 ctypes.byref(info)))
        return info

    def query(self, request: Request) -> CRequestInfo:
        info = CRequestInfo()
        _native.check(self.lib.drift_server_query(self.handle, request.ticket, ctypes.byref(info)))
        return info

    def stats(self) -> dict:
        """Chunk launches so far: replays of captured hipGraphs (graph_launches),
        captures (one per chunk shape) and chunks launched without a graph."""
        v = [ctypes.c_uint64() for _ in range(3)]
        _native.check(self.lib.drift_server_stats(self.handle, *[ctypes.byref(x) for x in v]))
        return {"graph_launches": v[0].value, "graph_captures": v[1].value,
                "eager_chunks": v[2].value}

    def busy(self) -> dict:
        """drift_server_busy: seconds in chunk launches (the GPU's share of the
        wall time), seconds and bytes of result copies to host memory."""
        busy, copy = ctypes.c_double(), ctypes.c_double()
        nbytes = ctypes.c_uint64()
        _native.check(self.lib.drift_server_busy(self.handle, ctypes.byref(busy),
                                                 ctypes.byref(copy), ctypes.byref(nbytes)))
        return {"busy_s": busy.value, "copy_s": copy.value, "copy_bytes": nbytes.value}

    def close(self) -> None:
        if self.handle:
            self.lib.drift_server_destroy(self.handle)
            self.handle = ctypes.c_void_p()
            self._live.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def latency_percentiles(infos, qs=(50, 99)) -> dict:
    lat = np.array([i.complete_s - i.submit_s for i in infos])
    return {f"p{q}": float(np.percentile(lat, q)) for q in qs}
