"""ctypes binding of the CPU oracle (oracle/liboracle.so) for the tests.

Test infrastructure: the oracle is what the HIP path is compared
against; it is never called by the product.
"""

from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent.parent
ORACLE_DIR = REPO / "oracle"
ORACLE_LIB = ORACLE_DIR / "liboracle.so"

REGIONS = ["init", "attempt", "accept", "dense", "sample", "substep", "rhs"]
OP_KINDS = ["add", "mul", "div", "sqrt", "trans"]


class OPotential(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("reserved", ctypes.c_int32)] + [
        (name, ctypes.c_double)
        for name in (
            "amp", "b", "bulge_amp", "bulge_alpha", "bulge_rc",
            "disk_amp", "disk_a", "disk_b", "halo_amp", "halo_a",
        )
    ] + [
        ("origin_tracks", ctypes.c_void_p),
        ("origin_params", ctypes.c_void_p),
        ("n_origins", ctypes.c_int32),
        ("n_segments", ctypes.c_int32),
        ("t_start", ctypes.c_double),
        ("t_end", ctypes.c_double),
    ]


class OStats(ctypes.Structure):
    _fields_ = [
        (name, ctypes.c_int64)
        for name in ("attempts", "accepted", "rejected", "rhs_evals", "samples", "dense_steps")
    ] + [("status", ctypes.c_int32), ("reserved", ctypes.c_int32)]


OSTATS_DTYPE = np.dtype(
    [(name, "<i8") for name in ("attempts", "accepted", "rejected", "rhs_evals", "samples",
                                "dense_steps")] + [("status", "<i4"), ("reserved", "<i4")]
)

_lib = None


def load() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not ORACLE_LIB.exists():
            subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], check=True)
        lib = ctypes.CDLL(str(ORACLE_LIB))
        c = ctypes
        P = c.POINTER(OPotential)
        lib.oracle_orbit.restype = c.c_int
        lib.oracle_orbit.argtypes = [c.c_int, P, c.c_void_p, c.c_void_p, c.c_int, c.c_double,
                                     c.c_double, c.c_double, c.c_void_p, c.c_void_p]
        lib.oracle_batch.restype = c.c_int64
        lib.oracle_batch.argtypes = [c.c_int, P, c.c_int64, c.c_void_p, c.c_void_p, c.c_int,
                                     c.c_double, c.c_double, c.c_double, c.c_void_p, c.c_void_p,
                                     c.c_int, c.c_int64]
        lib.oracle_count_ops.restype = None
        lib.oracle_count_ops.argtypes = [c.c_int, P, c.c_void_p, c.c_void_p, c.c_int, c.c_double,
                                         c.c_double, c.c_double, c.c_void_p, c.c_void_p]
        lib.oracle_force.restype = None
        lib.oracle_force.argtypes = [P, c.c_double, c.c_double, c.c_double, c.c_void_p]
        lib.oracle_force_at.restype = None
        lib.oracle_force_at.argtypes = [P, c.c_double, c.c_double, c.c_double, c.c_double,
                                        c.c_void_p]
        lib.oracle_potential_value.restype = c.c_double
        lib.oracle_potential_value.argtypes = [P, c.c_double, c.c_double, c.c_double]
        lib.oracle_lower_gamma.restype = c.c_double
        lib.oracle_lower_gamma.argtypes = [c.c_double, c.c_double]
        lib.oracle_upper_gamma.restype = c.c_double
        lib.oracle_upper_gamma.argtypes = [c.c_double, c.c_double]
        lib.oracle_ln.restype = c.c_double
        lib.oracle_ln.argtypes = [c.c_double]
        lib.oracle_sample_times.restype = None
        lib.oracle_sample_times.argtypes = [c.c_double, c.c_double, c.c_int64, c.c_void_p]
        _lib = lib
    return _lib


def to_opot(spec) -> OPotential:
    """From astrodrift_amd.PotentialSpec (or None = galpy unit Kepler),
    moving origins included."""
    if spec is None:
        return OPotential(kind=0)
    pot = OPotential(spec.kind, 0, spec.amp, spec.b, spec.bulge_amp, spec.bulge_alpha,
                     spec.bulge_rc, spec.disk_amp, spec.disk_a, spec.disk_b, spec.halo_amp,
                     spec.halo_a)
    o = getattr(spec, "origins", None)
    if o is not None:
        tracks = np.ascontiguousarray(o.tracks, dtype=np.float64)
        params = np.ascontiguousarray(o.params, dtype=np.float64)
        pot.origin_tracks = tracks.ctypes.data
        pot.origin_params = params.ctypes.data
        pot.n_origins = tracks.shape[0]
        pot.n_segments = tracks.shape[1]
        pot.t_start = o.t_start
        pot.t_end = o.t_end
        pot._keep = (tracks, params)  # the arrays the pointers refer to
    return pot


def _p(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


def orbit(method: int, y0, times, rtol_ln, atol_ln, spec=None, dt_one=-9999.99):
    """One orbit: returns (result (nt, 6), stats record, err)."""
    lib = load()
    y = np.ascontiguousarray(y0, dtype=np.float64)
    t = np.ascontiguousarray(times, dtype=np.float64)
    res = np.zeros((t.size, 6))
    st = OStats()
    pot = to_opot(spec)
    err = lib.oracle_orbit(method, ctypes.byref(pot), _p(y), _p(t), t.size, dt_one, rtol_ln,
                           atol_ln, _p(res), ctypes.byref(st))
    return res, st, err


def batch(method: int, state0, times, rtol_ln, atol_ln, spec=None, dt_one=-9999.99,
          nthreads: int = 1, stride: int = 1, want_out: bool = True):
    """Batch: returns (out (nt, n, 6) or None, stats array)."""
    lib = load()
    y = np.ascontiguousarray(state0, dtype=np.float64).reshape(-1, 6)
    t = np.ascontiguousarray(times, dtype=np.float64)
    n = y.shape[0]
    out = np.full((t.size, n, 6), np.nan) if want_out else None
    stats = np.zeros(n, dtype=OSTATS_DTYPE)
    pot = to_opot(spec)
    lib.oracle_batch(method, ctypes.byref(pot), n, _p(y), _p(t), t.size, dt_one, rtol_ln, atol_ln,
                     _p(out) if out is not None else None, _p(stats), nthreads, stride)
    return out, stats


def count_ops(method: int, y0, times, rtol_ln, atol_ln, spec=None, dt_one=-9999.99):
    """Counting run: returns ({region: {kind: count}}, stats)."""
    lib = load()
    y = np.ascontiguousarray(y0, dtype=np.float64)
    t = np.ascontiguousarray(times, dtype=np.float64)
    ops = np.zeros((len(REGIONS), len(OP_KINDS)), dtype=np.int64)
    st = OStats()
    pot = to_opot(spec)
    lib.oracle_count_ops(method, ctypes.byref(pot), _p(y), _p(t), t.size, dt_one, rtol_ln,
                         atol_ln, _p(ops), ctypes.byref(st))
    table = {r: dict(zip(OP_KINDS, map(int, ops[i]))) for i, r in enumerate(REGIONS)}
    return table, st


def force(spec, x, y, z):
    lib = load()
    out = np.zeros(3)
    pot = to_opot(spec)
    lib.oracle_force(ctypes.byref(pot), x, y, z, _p(out))
    return out


def force_at(spec, t, x, y, z):
    """Static force plus the moving origins' at time t."""
    lib = load()
    out = np.zeros(3)
    pot = to_opot(spec)
    lib.oracle_force_at(ctypes.byref(pot), t, x, y, z, _p(out))
    return out


def potential_value(spec, x, y, z) -> float:
    pot = to_opot(spec)
    return load().oracle_potential_value(ctypes.byref(pot), x, y, z)


def sample_times(start, end, steps) -> np.ndarray:
    out = np.zeros(steps)
    load().oracle_sample_times(start, end, steps, _p(out))
    return out
