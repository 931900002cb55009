"""ctypes binding of libdrift_hip.so (the C ABI in include/drift_hip.h).

The library is the only compute path: if it is missing or cannot be loaded
every integration raises :class:`NativeLibraryError`; there is no fallback.
"""

from __future__ import annotations

import contextlib
import ctypes
import math
import threading
import time
from dataclasses import dataclass, field, replace
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libdrift_hip.so"
HEADER_PATH = Path(__file__).resolve().parent.parent / "include" / "drift_hip.h"

# drift_status
DRIFT_OK = 0
DRIFT_ERR_INVALID_ARGUMENT = 1
DRIFT_ERR_UNSUPPORTED = 2
DRIFT_ERR_DEVICE = 3
DRIFT_ERR_INTEGRATION = 4
DRIFT_ERR_OUT_OF_MEMORY = 5

METHOD_DOPR54 = 0
METHOD_DOP853 = 1

POTENTIAL_GALPY_KEPLER = 0
POTENTIAL_KEPLER = 1
POTENTIAL_PLUMMER = 2
POTENTIAL_MW = 3
POTENTIAL_NONE = 4  # no static force (moving origins only, or free motion)

PRECISION_FP64 = 0
PRECISION_MIXED = 1

SCHEDULE_PLAIN = 0
SCHEDULE_PERSISTENT = 1
SCHEDULE_TUNED = 2  # what tune_device stored for this method / potential / precision / log2 n

ORDER_AUTO = 0   # cost order for launches of >= 16384 orbits (either schedule)
ORDER_INDEX = 1
ORDER_COST = 2
STAGING_AUTO = 0
STAGING_DIRECT = 1
STAGING_STAGED = 2

ABI_VERSION = 6

STATUS_CODE_MASK = 0xFF
STATUS_NONFINITE = 0x100
STATUS_STEP_LIMIT = 0x200

# src/integrators/galpy/mod.rs:8
INITIAL_STEP_SENTINEL = -9999.99


class NativeLibraryError(RuntimeError):
    """libdrift_hip.so is missing or failed to load."""


class DriftNativeError(RuntimeError):
    """A C-ABI call returned a non-zero drift_status."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


class CPotential(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("origins", ctypes.c_int32),
        ("amp", ctypes.c_double),
        ("b", ctypes.c_double),
        ("bulge_amp", ctypes.c_double),
        ("bulge_alpha", ctypes.c_double),
        ("bulge_rc", ctypes.c_double),
        ("disk_amp", ctypes.c_double),
        ("disk_a", ctypes.c_double),
        ("disk_b", ctypes.c_double),
        ("halo_amp", ctypes.c_double),
        ("halo_a", ctypes.c_double),
    ]


class COrbitStats(ctypes.Structure):
    _fields_ = [
        ("accepted", ctypes.c_uint32),
        ("rejected", ctypes.c_uint32),
        ("dense_steps", ctypes.c_uint32),
        ("status", ctypes.c_int32),
    ]


STATS_DTYPE = np.dtype(
    [("accepted", "<u4"), ("rejected", "<u4"), ("dense_steps", "<u4"), ("status", "<i4")]
)


class CLaunchOptions(ctypes.Structure):
    _fields_ = [
        ("precision", ctypes.c_int32),
        ("block_size", ctypes.c_int32),
        ("schedule", ctypes.c_int32),
        ("max_attempts", ctypes.c_uint32),
        ("ordering", ctypes.c_int32),
        ("staging", ctypes.c_int32),
    ]


# ---------------------------------------------------------------------------
# Potential descriptors
# ---------------------------------------------------------------------------


def _lower_gamma(a: float, x: float) -> float:
    """gamma(a, x), series (A&S 6.5.29) / Lentz continued fraction (6.5.31)."""
    if x <= 0.0:
        return 0.0
    if x < a + 1.0:
        ap, term, total = a, 1.0 / a, 1.0 / a
        for _ in range(1000):
            ap += 1.0
            term *= x / ap
            total += term
            if abs(term) < abs(total) * 1e-17:
                break
        return total * math.exp(-x + a * math.log(x))
    tiny = 1e-300
    b = x + 1.0 - a
    c = 1.0 / tiny
    d = 1.0 / b
    h = d
    for i in range(1, 1000):
        an = -i * (i - a)
        b += 2.0
        d = an * d + b
        d = tiny if abs(d) < tiny else d
        c = b + an / c
        c = tiny if abs(c) < tiny else c
        d = 1.0 / d
        delta = d * c
        h *= delta
        if abs(delta - 1.0) < 1e-17:
            break
    return math.gamma(a) - math.exp(-x + a * math.log(x)) * h


ORIGIN_COEFS = 18


@dataclass(frozen=True, eq=False)
class Origins:
    """Moving origins (the reference's CustomOrigin, shared/src/potential.rs:
    264-356): origin i is a point mass (b = 0) or Plummer sphere of amplitude
    params[i, 0] and radius params[i, 1] whose centre follows the quintic
    tracks[i, s] (x, y, z coefficients, tau^5 .. tau^0) on segment s of
    n_segments equal segments of [t_start, t_end], at local time tau in [0, 1).
    """

    tracks: np.ndarray  # (n_origins, n_segments, 18)
    params: np.ndarray  # (n_origins, 2): amp, b
    t_start: float
    t_end: float

    def __post_init__(self):
        tracks = np.ascontiguousarray(self.tracks, dtype=np.float64)
        params = np.ascontiguousarray(self.params, dtype=np.float64).reshape(-1, 2)
        if tracks.ndim != 3 or tracks.shape[2] != ORIGIN_COEFS:
            raise ValueError("tracks must have shape (n_origins, n_segments, 18)")
        if tracks.shape[0] != params.shape[0]:
            raise ValueError("tracks and params must describe the same origins")
        object.__setattr__(self, "tracks", tracks)
        object.__setattr__(self, "params", params)

    @property
    def n_origins(self) -> int:
        return self.tracks.shape[0]

    @property
    def n_segments(self) -> int:
        return self.tracks.shape[1]

    @staticmethod
    def hermite(states, accels, times, params) -> "Origins":
        """Tracks of sampled orbits (drift_origin_tracks_hermite, host only):
        states (nt, n, 6) time-major as the integrators write them, accels
        (nt, n, 3) the accelerations at those samples, times (nt,) uniformly
        spaced; params (n, 2) or one (amp, b) pair for every orbit."""
        lib = load_library()
        y = np.ascontiguousarray(states, dtype=np.float64)
        a = np.ascontiguousarray(accels, dtype=np.float64)
        t = np.ascontiguousarray(times, dtype=np.float64)
        if y.ndim != 3 or y.shape[2] != 6 or a.shape != y.shape[:2] + (3,) or t.shape != y.shape[:1]:
            raise ValueError("states (nt, n, 6), accels (nt, n, 3) and times (nt,) must agree")
        nt, n = y.shape[0], y.shape[1]
        tracks = np.empty((n, nt - 1, ORIGIN_COEFS))
        t0, t1 = ctypes.c_double(), ctypes.c_double()
        check(lib.drift_origin_tracks_hermite(_ptr(y), _ptr(a), _ptr(t), nt, n, _ptr(tracks),
                                              ctypes.byref(t0), ctypes.byref(t1)))
        pb = np.broadcast_to(np.asarray(params, dtype=np.float64).reshape(-1, 2), (n, 2))
        return Origins(tracks, pb.copy(), t0.value, t1.value)

    @staticmethod
    def fixed(params, n_segments: int, t_start: float, t_end: float) -> "Origins":
        """Origins resting at (0, 0, 0): a static point mass or Plummer
        sphere expressed as a moving origin."""
        pb = np.asarray(params, dtype=np.float64).reshape(-1, 2)
        return Origins(np.zeros((pb.shape[0], n_segments, ORIGIN_COEFS)), pb, t_start, t_end)

    @staticmethod
    def concat(sets: list) -> "Origins":
        """One set holding the origins of several sets on the same segments."""
        first = sets[0]
        for o in sets[1:]:
            if (o.n_segments, o.t_start, o.t_end) != (first.n_segments, first.t_start, first.t_end):
                raise ValueError("origin sets to combine must share their segments")
        return Origins(np.concatenate([o.tracks for o in sets]),
                       np.concatenate([o.params for o in sets]), first.t_start, first.t_end)


@dataclass(frozen=True)
class PotentialSpec:
    """A force model the kernels evaluate (G = 1 natural units): a static
    kind plus, optionally, moving origins."""

    kind: int
    amp: float = 0.0
    b: float = 0.0
    bulge_amp: float = 0.0
    bulge_alpha: float = 0.0
    bulge_rc: float = 0.0
    disk_amp: float = 0.0
    disk_a: float = 0.0
    disk_b: float = 0.0
    halo_amp: float = 0.0
    halo_a: float = 0.0
    origins: Origins | None = field(default=None, compare=False)

    def to_c(self, origins_id: int = 0) -> CPotential:
        return CPotential(
            self.kind,
            origins_id,
            self.amp,
            self.b,
            self.bulge_amp,
            self.bulge_alpha,
            self.bulge_rc,
            self.disk_amp,
            self.disk_a,
            self.disk_b,
            self.halo_amp,
            self.halo_a,
        )

    @staticmethod
    def galpy_kepler() -> "PotentialSpec":
        """The reference's hard-coded unit Kepler force (dopr54.rs:75-88)."""
        return PotentialSpec(POTENTIAL_GALPY_KEPLER)

    @staticmethod
    def none() -> "PotentialSpec":
        """No static force (with moving origins, or free motion)."""
        return PotentialSpec(POTENTIAL_NONE)

    def with_origins(self, origins: Origins | None) -> "PotentialSpec":
        return replace(self, origins=origins)

    @staticmethod
    def kepler(amp: float) -> "PotentialSpec":
        return PotentialSpec(POTENTIAL_KEPLER, amp=float(amp))

    @staticmethod
    def plummer(amp: float, b: float) -> "PotentialSpec":
        return PotentialSpec(POTENTIAL_PLUMMER, amp=float(amp), b=float(b))

    @staticmethod
    def milky_way(
        bulge_alpha: float = 1.8,
        bulge_rc: float = 1.9 / 8.0,
        bulge_frac: float = 0.05,
        disk_a: float = 3.0 / 8.0,
        disk_b: float = 0.28 / 8.0,
        disk_frac: float = 0.6,
        halo_a: float = 16.0 / 8.0,
        halo_frac: float = 0.35,
    ) -> "PotentialSpec":
        """Bulge + Miyamoto-Nagai disk + NFW halo normalised so that each
        component contributes `*_frac` of vc^2 at R = 1, z = 0 (vc(1) = 1 when
        the fractions sum to 1).  The defaults are the model behind the
        reference's BovyPotential (shared/src/potential.rs:94-114)."""
        disk_amp = disk_frac * (1.0 + (disk_a + disk_b) ** 2) ** 1.5
        halo_amp = halo_frac / (math.log(1.0 + 1.0 / halo_a) - 1.0 / (halo_a + 1.0))
        a = 1.5 - 0.5 * bulge_alpha
        bulge_amp = bulge_frac / (
            2.0 * math.pi * bulge_rc ** (3.0 - bulge_alpha) * _lower_gamma(a, 1.0 / bulge_rc**2)
        )
        return PotentialSpec(
            POTENTIAL_MW,
            bulge_amp=bulge_amp,
            bulge_alpha=bulge_alpha,
            bulge_rc=bulge_rc,
            disk_amp=disk_amp,
            disk_a=disk_a,
            disk_b=disk_b,
            halo_amp=halo_amp,
            halo_a=halo_a,
        )


class CJob(ctypes.Structure):
    _fields_ = [
        ("method", ctypes.c_int32),
        ("potential", ctypes.c_int32),
        ("n", ctypes.c_size_t),
        ("nt", ctypes.c_size_t),
        ("rtol", ctypes.c_double),
        ("atol", ctypes.c_double),
        ("dt_one_init", ctypes.c_double),
        ("span_hint", ctypes.c_double),
        ("d_state0", ctypes.c_void_p),
        ("d_times", ctypes.c_void_p),
        ("d_state_out", ctypes.c_void_p),
        ("d_stats", ctypes.c_void_p),
    ]


class CLaunchReport(ctypes.Structure):
    _fields_ = [
        ("setup_s", ctypes.c_double),
        ("h2d_s", ctypes.c_double),
        ("kernel_s", ctypes.c_double),
        ("d2h_s", ctypes.c_double),
        ("h2d_bytes", ctypes.c_uint64),
        ("d2h_bytes", ctypes.c_uint64),
        ("grid", ctypes.c_int32),
        ("block", ctypes.c_int32),
    ]


# ---------------------------------------------------------------------------
# Library loading
# ---------------------------------------------------------------------------

_lib = None
_lib_lock = threading.Lock()

_D = ctypes.POINTER(ctypes.c_double)


def _declare(lib: ctypes.CDLL) -> None:
    c = ctypes
    lib.drift_get_abi_version.restype = c.c_int
    lib.drift_get_abi_version.argtypes = []
    lib.drift_last_error.restype = c.c_char_p
    lib.drift_last_error.argtypes = []
    lib.drift_session_create.restype = c.c_int
    lib.drift_session_create.argtypes = [c.c_int, c.POINTER(c.c_void_p)]
    lib.drift_session_destroy.restype = c.c_int
    lib.drift_session_destroy.argtypes = [c.c_void_p]
    host = [c.c_void_p, c.c_void_p, c.c_void_p, c.c_void_p, c.c_size_t, c.c_size_t,
            c.c_double, c.c_double]
    lib.drift_galpy_dopr54.restype = c.c_int
    lib.drift_galpy_dopr54.argtypes = host + [c.c_double, c.POINTER(CPotential), c.c_void_p]
    lib.drift_galpy_dop853.restype = c.c_int
    lib.drift_galpy_dop853.argtypes = host + [c.POINTER(CPotential), c.c_void_p]
    lib.drift_galpy_dopr54_device.restype = c.c_int
    lib.drift_galpy_dopr54_device.argtypes = host + [
        c.c_double, c.POINTER(CPotential), c.c_void_p, c.POINTER(CLaunchOptions), c.c_void_p]
    lib.drift_galpy_dop853_device.restype = c.c_int
    lib.drift_galpy_dop853_device.argtypes = host + [
        c.POINTER(CPotential), c.c_void_p, c.POINTER(CLaunchOptions), c.c_void_p]
    lib.drift_tune_device.restype = c.c_int
    lib.drift_tune_device.argtypes = [c.c_void_p, c.c_int32, c.c_void_p, c.c_void_p, c.c_size_t,
                                      c.c_size_t, c.c_double, c.c_double, c.c_double,
                                      c.POINTER(CPotential), c.c_size_t,
                                      c.POINTER(CLaunchOptions), c.POINTER(c.c_double), c.c_void_p]
    lib.drift_ragged_device.restype = c.c_int
    lib.drift_ragged_device.argtypes = [c.c_void_p, c.POINTER(CJob), c.c_size_t,
                                        c.POINTER(CPotential), c.c_size_t,
                                        c.POINTER(CLaunchOptions), c.c_void_p]
    lib.drift_tune_ragged.restype = c.c_int
    lib.drift_tune_ragged.argtypes = [c.c_void_p, c.POINTER(CJob), c.c_size_t,
                                      c.POINTER(CPotential), c.c_size_t,
                                      c.POINTER(CLaunchOptions), c.c_size_t,
                                      c.POINTER(c.c_int32), c.c_void_p]
    lib.drift_force_device.restype = c.c_int
    lib.drift_force_device.argtypes = [c.c_void_p, c.POINTER(CPotential), c.c_void_p, c.c_void_p,
                                       c.c_void_p, c.c_size_t, c.c_int32, c.c_void_p]
    lib.drift_force.restype = c.c_int
    lib.drift_force.argtypes = [c.c_void_p, c.POINTER(CPotential), c.c_void_p, c.c_void_p,
                                c.c_void_p, c.c_size_t]
    lib.drift_origins_create.restype = c.c_int
    lib.drift_origins_create.argtypes = [c.c_void_p, c.c_void_p, c.c_void_p, c.c_int32, c.c_int32,
                                         c.c_double, c.c_double, c.POINTER(c.c_int32)]
    lib.drift_origins_destroy.restype = c.c_int
    lib.drift_origins_destroy.argtypes = [c.c_void_p, c.c_int32]
    lib.drift_origin_tracks_hermite.restype = c.c_int
    lib.drift_origin_tracks_hermite.argtypes = [c.c_void_p, c.c_void_p, c.c_void_p, c.c_size_t,
                                                c.c_size_t, c.c_void_p, _D, _D]
    lib.drift_mw_table_check.restype = c.c_int
    lib.drift_mw_table_check.argtypes = [c.POINTER(CPotential), c.c_int32, _D]
    lib.drift_launch_shape.restype = c.c_int
    lib.drift_launch_shape.argtypes = [c.c_void_p, c.c_int32, c.POINTER(CPotential), c.c_size_t,
                                       c.POINTER(CLaunchOptions), c.POINTER(c.c_int32),
                                       c.POINTER(c.c_int32)]
    lib.drift_session_last_launch.restype = c.c_int
    lib.drift_session_last_launch.argtypes = [c.c_void_p, c.POINTER(CLaunchReport)]


def load_library() -> ctypes.CDLL:
    """Loads libdrift_hip.so once; raises NativeLibraryError if unavailable."""
    global _lib
    with _lib_lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise NativeLibraryError(
                    f"{LIB_PATH} is not built; run `python -m astrodrift_amd.build` "
                    "(hipcc, gfx950)"
                )
            try:
                lib = ctypes.CDLL(str(LIB_PATH))
            except OSError as exc:  # pragma: no cover - environment specific
                raise NativeLibraryError(f"cannot load {LIB_PATH}: {exc}") from exc
            _declare(lib)
            if lib.drift_get_abi_version() != ABI_VERSION:
                raise NativeLibraryError("libdrift_hip.so ABI version mismatch")
            _lib = lib
        return _lib


def last_error() -> str:
    msg = load_library().drift_last_error()
    return msg.decode() if msg else ""


def check(code: int) -> None:
    if code != DRIFT_OK:
        raise DriftNativeError(code, last_error() or f"drift status {code}")


class Session:
    """Owns a libdrift_hip session (device, stream and cached device buffers).

    A session may be used by one thread at a time, and ctypes releases the
    GIL during a foreign call, so every call on the handle holds `lock`."""

    def __init__(self, device: int = 0):
        started = time.perf_counter()
        self.lib = load_library()
        self.lock = threading.RLock()
        handle = ctypes.c_void_p()
        check(self.lib.drift_session_create(int(device), ctypes.byref(handle)))
        self.handle = handle
        self.device = int(device)
        # seconds taken to load the library and create the session (device
        # context and stream); reported once with the first launch's setup
        self.created_s = time.perf_counter() - started
        self.creation_reported = False

    def close(self) -> None:
        with self.lock:
            if self.handle:
                self.lib.drift_session_destroy(self.handle)
                self.handle = ctypes.c_void_p()

    def __del__(self):  # pragma: no cover - interpreter shutdown order varies
        try:
            self.close()
        except Exception:
            pass


_sessions: dict[int, Session] = {}
_sessions_lock = threading.Lock()


def session(device: int = 0) -> Session:
    """The process-wide session for `device` (created on first use)."""
    with _sessions_lock:
        s = _sessions.get(device)
        if s is None:
            s = Session(device)
            _sessions[device] = s
        return s


def release_session(device: int = 0) -> None:
    """Destroys the process-wide session for `device`, freeing the device
    memory it caches (launch outputs of the host-buffer entry points, the
    staged-output buffer, tuner scratch, force tables); the next call creates
    a new one.  Close BoundPotential objects of the session first."""
    with _sessions_lock:
        s = _sessions.pop(device, None)
    if s is not None:
        s.close()


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


class BoundPotential:
    """A PotentialSpec whose moving origins stay uploaded to the session of
    `device` until close() (or the end of a `with` block), for repeated
    launches without a per-launch upload."""

    def __init__(self, spec: PotentialSpec, device: int = 0):
        self.session = session(device)
        self.spec = spec
        self.id = 0
        o = spec.origins
        if o is not None:
            ident = ctypes.c_int32(0)
            with self.session.lock:
                code = self.session.lib.drift_origins_create(
                    self.session.handle, _ptr(o.tracks), _ptr(o.params), o.n_origins,
                    o.n_segments, o.t_start, o.t_end, ctypes.byref(ident))
            check(code)
            self.id = ident.value
        self.c = spec.to_c(self.id)

    def close(self) -> None:
        if self.id:
            with self.session.lock:
                self.session.lib.drift_origins_destroy(self.session.handle, self.id)
            self.id = 0

    def __enter__(self) -> "BoundPotential":
        return self

    def __exit__(self, *exc) -> None:
        self.close()


@contextlib.contextmanager
def _c_potential(s: Session, potential):
    """The C descriptor of `potential` (None = NULL), with its moving origins
    uploaded to the session for the duration of the block (a BoundPotential
    keeps its own upload).  Callers hold s.lock around the block."""
    if potential is None:
        yield None
        return
    if isinstance(potential, BoundPotential):
        yield potential.c
        return
    o = potential.origins
    if o is None:
        yield potential.to_c()
        return
    ident = ctypes.c_int32(0)
    check(s.lib.drift_origins_create(s.handle, _ptr(o.tracks), _ptr(o.params), o.n_origins,
                                     o.n_segments, o.t_start, o.t_end, ctypes.byref(ident)))
    try:
        yield potential.to_c(ident.value)
    finally:
        s.lib.drift_origins_destroy(s.handle, ident.value)


def _as_state(state0) -> np.ndarray:
    a = np.ascontiguousarray(state0, dtype=np.float64)
    if a.ndim == 1:
        if a.size % 6:
            raise ValueError("state0 must have 6N values")
        a = a.reshape(-1, 6)
    if a.ndim != 2 or a.shape[1] != 6:
        raise ValueError("state0 must have shape (N, 6)")
    return a


def launch_host(
    method: int,
    state0,
    times,
    rtol_ln: float,
    atol_ln: float,
    potential: PotentialSpec | None = None,
    dt_one: float = INITIAL_STEP_SENTINEL,
    device: int = 0,
    raise_on_integration_error: bool = True,
):
    """Host-buffer launch (drift_galpy_dopr54 / drift_galpy_dop853).

    Returns (out, stats, status): out has shape (nt, n, 6) time-major, stats a
    structured array of drift_orbit_stats, status the drift_status code.
    """
    s = session(device)
    y0 = _as_state(state0)
    t = np.ascontiguousarray(times, dtype=np.float64)
    n, nt = y0.shape[0], t.shape[0]
    out = np.empty((nt, n, 6), dtype=np.float64)
    stats = np.zeros(n, dtype=STATS_DTYPE)
    with s.lock, _c_potential(s, potential) as pot:
        pot_ref = ctypes.byref(pot) if pot is not None else None
        if method == METHOD_DOPR54:
            code = s.lib.drift_galpy_dopr54(s.handle, _ptr(y0), _ptr(t), _ptr(out), n, nt,
                                            rtol_ln, atol_ln, dt_one, pot_ref, _ptr(stats))
        else:
            code = s.lib.drift_galpy_dop853(s.handle, _ptr(y0), _ptr(t), _ptr(out), n, nt,
                                            rtol_ln, atol_ln, pot_ref, _ptr(stats))
    if code != DRIFT_OK and (code != DRIFT_ERR_INTEGRATION or raise_on_integration_error):
        check(code)
    return out, stats, code


def launch_shape(method: int, n: int, potential: PotentialSpec | None = None,
                 precision: int = PRECISION_FP64, schedule: int | None = None,
                 block_size: int = 0, device: int = 0) -> tuple[int, int]:
    """drift_launch_shape: (grid, block) of the integrator kernel for n orbits;
    schedule None is the host-buffer entry points' shape (persistent for
    DOP853, plain for DOPR54)."""
    s = session(device)
    if potential is None:
        pot = None
    elif isinstance(potential, BoundPotential):
        pot = potential.c
    else:
        # only the presence of moving origins matters to the shape
        pot = potential.to_c(1 if potential.origins is not None else 0)
    opts = None if schedule is None else CLaunchOptions(precision, block_size, schedule, 0, 0, 0)
    grid, block = ctypes.c_int32(0), ctypes.c_int32(0)
    with s.lock:
        check(s.lib.drift_launch_shape(s.handle, method,
                                       ctypes.byref(pot) if pot is not None else None, n,
                                       ctypes.byref(opts) if opts is not None else None,
                                       ctypes.byref(grid), ctypes.byref(block)))
    return grid.value, block.value


def last_launch(device: int = 0) -> dict:
    """drift_session_last_launch: the phases of the session's last host-buffer
    launch (seconds; device phases from HIP events), its byte counts and the
    integrator's launch shape."""
    s = session(device)
    r = CLaunchReport()
    with s.lock:
        check(s.lib.drift_session_last_launch(s.handle, ctypes.byref(r)))
    return {name: getattr(r, name) for name, _ in CLaunchReport._fields_}


def launch_device(
    method: int,
    d_state0: int,
    d_times: int,
    d_out: int,
    n: int,
    nt: int,
    rtol_ln: float,
    atol_ln: float,
    potential: PotentialSpec | None = None,
    d_stats: int = 0,
    precision: int = PRECISION_FP64,
    schedule: int = SCHEDULE_PLAIN,
    block_size: int = 0,
    max_attempts: int = 0,
    stream: int = 0,
    dt_one: float = INITIAL_STEP_SENTINEL,
    device: int = 0,
    ordering: int = ORDER_AUTO,
    staging: int = STAGING_AUTO,
) -> None:
    """Device-resident launch on raw device pointers (e.g. torch data_ptr()).

    With moving origins the set is uploaded before the launch and freed after
    it (freeing waits for the device)."""
    s = session(device)
    opts = CLaunchOptions(precision, block_size, schedule, max_attempts, ordering, staging)
    args = (s.handle, ctypes.c_void_p(d_state0), ctypes.c_void_p(d_times), ctypes.c_void_p(d_out),
            n, nt, rtol_ln, atol_ln)
    with s.lock, _c_potential(s, potential) as pot:
        pot_ref = ctypes.byref(pot) if pot is not None else None
        if method == METHOD_DOPR54:
            code = s.lib.drift_galpy_dopr54_device(*args, dt_one, pot_ref,
                                                   ctypes.c_void_p(d_stats), ctypes.byref(opts),
                                                   ctypes.c_void_p(stream))
        else:
            code = s.lib.drift_galpy_dop853_device(*args, pot_ref, ctypes.c_void_p(d_stats),
                                                   ctypes.byref(opts), ctypes.c_void_p(stream))
    check(code)


def tune_device(method: int, d_state0: int, d_times: int, n: int, nt: int, rtol_ln: float,
                atol_ln: float, potential: PotentialSpec | None = None,
                precision: int = PRECISION_FP64, prefix_samples: int = 0, stream: int = 0,
                dt_one: float = INITIAL_STEP_SENTINEL, device: int = 0) -> dict:
    """drift_tune_device: times the launch candidates on a prefix of the time
    grid and stores the fastest in the session for SCHEDULE_TUNED launches."""
    s = session(device)
    best = CLaunchOptions(precision, 0, 0, 0, 0, 0)
    ms = ctypes.c_double(0.0)
    with s.lock, _c_potential(s, potential) as pot:
        code = s.lib.drift_tune_device(s.handle, method, ctypes.c_void_p(d_state0),
                                       ctypes.c_void_p(d_times), n, nt, rtol_ln, atol_ln, dt_one,
                                       ctypes.byref(pot) if pot is not None else None,
                                       prefix_samples, ctypes.byref(best), ctypes.byref(ms),
                                       ctypes.c_void_p(stream))
    check(code)
    return {"schedule": best.schedule, "block_size": best.block_size,
            "ordering": best.ordering, "staging": best.staging, "ms": ms.value}


@dataclass
class DeviceJob:
    """One job of a ragged batch, on device pointers."""

    method: int
    potential: int  # index into the potentials list
    n: int
    nt: int
    rtol_ln: float
    atol_ln: float
    d_state0: int
    d_times: int
    d_out: int
    d_stats: int = 0
    dt_one: float = INITIAL_STEP_SENTINEL
    span_hint: float = 0.0

    def to_c(self) -> CJob:
        return CJob(self.method, self.potential, self.n, self.nt, self.rtol_ln, self.atol_ln,
                    self.dt_one, self.span_hint, self.d_state0, self.d_times, self.d_out,
                    self.d_stats)


def ragged_device(jobs: list[DeviceJob], potentials: list, precision: int = PRECISION_FP64,
                  block_size: int = 0, max_attempts: int = 0, stream: int = 0,
                  device: int = 0) -> None:
    """drift_ragged_device: every job to completion, one persistent launch per
    method (returns when done)."""
    s = session(device)
    arr = (CJob * len(jobs))(*[j.to_c() for j in jobs])
    pots = (CPotential * len(potentials))(
        *[(p.to_c() if p is not None else PotentialSpec.galpy_kepler().to_c()) for p in potentials])
    opts = CLaunchOptions(precision, block_size, 0, max_attempts, ORDER_AUTO, 0)
    with s.lock:
        code = s.lib.drift_ragged_device(s.handle, arr, len(jobs), pots, len(potentials),
                                         ctypes.byref(opts), ctypes.c_void_p(stream))
    check(code)


def tune_ragged(jobs: list[DeviceJob], potentials: list, precision: int = PRECISION_FP64,
                prefix_samples: int = 0, stream: int = 0, device: int = 0) -> dict:
    """drift_tune_ragged: times block sizes 64/128/256 per method on a prefix
    of every job's grid and stores the fastest in the session for ragged
    launches with block_size 0."""
    s = session(device)
    arr = (CJob * len(jobs))(*[j.to_c() for j in jobs])
    pots = (CPotential * len(potentials))(
        *[(p.to_c() if p is not None else PotentialSpec.galpy_kepler().to_c()) for p in potentials])
    opts = CLaunchOptions(precision, 0, 0, 0, ORDER_AUTO, 0)
    best = (ctypes.c_int32 * 2)()
    with s.lock:
        code = s.lib.drift_tune_ragged(s.handle, arr, len(jobs), pots, len(potentials),
                                       ctypes.byref(opts), prefix_samples, best,
                                       ctypes.c_void_p(stream))
    check(code)
    return {"dopr54_block": best[METHOD_DOPR54], "dop853_block": best[METHOD_DOP853]}


def force_device(potential: PotentialSpec | None, d_pos: int, d_acc: int, n: int,
                 precision: int = PRECISION_FP64, stream: int = 0, device: int = 0,
                 d_t: int = 0) -> None:
    """drift_force_device: accelerations at n points (d_t: n times, or 0 for
    t = 0)."""
    s = session(device)
    with s.lock, _c_potential(s, potential) as pot:
        code = s.lib.drift_force_device(s.handle, ctypes.byref(pot) if pot is not None else None,
                                        ctypes.c_void_p(d_pos), ctypes.c_void_p(d_t),
                                        ctypes.c_void_p(d_acc), n, precision,
                                        ctypes.c_void_p(stream))
    check(code)


def force_host(potential: PotentialSpec | None, pos, t=None, device: int = 0) -> np.ndarray:
    """drift_force: FP64 accelerations (n, 3) at host positions pos (n, 3) and
    times t (n,) or None for t = 0."""
    s = session(device)
    p = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    n = p.shape[0]
    tt = None if t is None else np.ascontiguousarray(np.broadcast_to(t, (n,)), dtype=np.float64)
    acc = np.empty((n, 3))
    with s.lock, _c_potential(s, potential) as pot:
        code = s.lib.drift_force(s.handle, ctypes.byref(pot) if pot is not None else None,
                                 _ptr(p), _ptr(tt) if tt is not None else None, _ptr(acc), n)
    check(code)
    return acc


def mw_table_check(potential: PotentialSpec, samples_per_segment: int = 200) -> float:
    """Worst relative error of the MW spherical-force table (host only)."""
    lib = load_library()
    out = ctypes.c_double()
    pot = potential.to_c()
    check(lib.drift_mw_table_check(ctypes.byref(pot), samples_per_segment, ctypes.byref(out)))
    return out.value
