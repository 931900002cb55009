"""The reference's Python API (``dft.Config`` / ``sim.add`` / ``sim.run``) on
the gfx950 engine.

Mirrors src/interface/{mod,container,recipe,error}.rs and
python/drift/{__init__,selectors}.py: the same constructors, validation
messages, result order and shape ``(nt, n, 6)``, and exception types.  The GPU
engine runs every state-bearing container through libdrift_hip.so
(drift_galpy_dopr54 / drift_galpy_dop853).  Differences from the reference,
all deliberate:

* the per-container particle limit (MAX_PARTICLES = 1000,
  shared/src/lib.rs:25) is lifted, and the output-time limit is raised from
  1024 to MAX_OUTPUT_TIMES;
* ``Config(..., arbitrary_ts=True)`` takes output times that are not
  uniformly spaced and passes them to the integrators as given (the
  reference's FIXME, src/interface/mod.rs:89-90); DOPR54, which steps the
  fixed interval t[1] - t[0], refuses such grids;
* ``Engine.CPU`` configurations validate as before but ``run()`` raises
  NotImplementedError: this package replaces the GPU engine only;
* ``Config(..., forces="model")`` integrates each container in the potential
  of its force sources instead of the reference's hard-coded unit galpy Kepler
  force (the default, ``forces="reference"``, keeps the reference behaviour,
  src/dispatch/gpu.rs:155-156).  A particle group with an attached potential
  acts as a force source through moving origins (the reference's
  CustomKepler / CustomPlummer recipes, shared/src/model.rs:180-235, which it
  never constructs): the group is integrated first and each of its particles
  becomes a point mass or Plummer sphere following its trajectory.
"""

from __future__ import annotations

import itertools
import logging
import math
import time
import warnings
from dataclasses import dataclass
from typing import Any, Sequence

import numpy as np

from . import _native
from .selectors import Engine, Implementation, Method

MAX_MODEL_COMPONENTS = 11  # shared/src/lib.rs:19
MAX_OUTPUT_TIMES = 1 << 20  # reference: 1024 (shared/src/lib.rs:27)
MIN_RTOL = 1e-12  # shared/src/lib.rs:29
MIN_ATOL = 1e-12  # shared/src/lib.rs:30
INPUT_STATE_DIM = 6
GRID_ULP_TOLERANCE = 8.0  # src/interface/mod.rs:68
_EPS = np.finfo(np.float64).eps
_MIN_POSITIVE = np.finfo(np.float64).tiny

_log_gpu = logging.getLogger("drift.dispatch.gpu")
_log_hip = logging.getLogger("drift.dispatch.gpu.hip")


class DriftError(RuntimeError):
    """Base class for operational errors reported by drift."""


class IntegrationError(DriftError):
    """An integration could not be completed."""


DriftError.__module__ = "astrodrift_amd"
IntegrationError.__module__ = "astrodrift_amd"


# ---------------------------------------------------------------------------
# Output grid and tolerance (src/interface/mod.rs:29-171, shared/src/config.rs)
# ---------------------------------------------------------------------------


@dataclass(frozen=True)
class OutputGrid:
    start: float = 0.0
    end: float = 2.0 * math.pi
    steps: int = 100
    # arbitrary requested times (Config(..., arbitrary_ts=True)), passed to the
    # integrators as given; None for the affine grid
    explicit: tuple | None = None

    def sample(self, index: int) -> float:
        """OutputGrid::sample, shared/src/config.rs:57-66."""
        if self.explicit is not None:
            return self.explicit[index]
        if index == self.steps - 1:
            return self.end
        return self.start + (self.end - self.start) * float(index) / float(self.steps - 1)

    def times(self) -> np.ndarray:
        """sample_times, src/dispatch/mod.rs:14-18."""
        return np.array([self.sample(i) for i in range(self.steps)], dtype=np.float64)

    def is_affine(self) -> bool:
        """Whether the times are the affine grid's within GRID_ULP_TOLERANCE
        (src/interface/mod.rs:110-131)."""
        if self.explicit is None:
            return True
        affine = OutputGrid(self.start, self.end, self.steps)
        step = (self.end - self.start) / float(self.steps - 1)
        for index, actual in enumerate(self.explicit):
            expected = affine.sample(index)
            scale = max(abs(actual), abs(expected), abs(step), _MIN_POSITIVE)
            if abs(actual - expected) > GRID_ULP_TOLERANCE * _EPS * scale:
                return False
        return True


def _grid_new(start: float, end: float, steps: int) -> OutputGrid:
    if not (2 <= steps <= MAX_OUTPUT_TIMES):
        raise ValueError(f"ts must contain between 2 and {MAX_OUTPUT_TIMES} output times")
    if not (math.isfinite(start) and math.isfinite(end)):
        raise ValueError("ts endpoints must be finite")
    span = end - start
    if span == 0.0 or not math.isfinite(span):
        raise ValueError("ts must span a non-zero finite interval")
    return OutputGrid(float(start), float(end), int(steps))


def _grid_from_times(times: Sequence[float], arbitrary: bool = False) -> OutputGrid:
    times = [float(t) for t in times]
    if not (2 <= len(times) <= MAX_OUTPUT_TIMES):
        raise ValueError(f"ts must contain between 2 and {MAX_OUTPUT_TIMES} output times")
    if any(not math.isfinite(t) for t in times):
        raise ValueError("all ts values must be finite")
    grid = _grid_new(times[0], times[-1], len(times))
    increasing = grid.end > grid.start
    for a, b in zip(times, times[1:]):
        if (b <= a) if increasing else (b >= a):
            raise ValueError("ts values must be strictly increasing or strictly decreasing")
    if arbitrary:
        # the reference's FIXME (src/interface/mod.rs:89-90): keep the times as given
        return OutputGrid(grid.start, grid.end, grid.steps, tuple(times))
    if not OutputGrid(grid.start, grid.end, grid.steps, tuple(times)).is_affine():
        raise ValueError(
            "ts values must be uniformly spaced; arbitrary output times are not yet supported"
        )
    return grid


def _parse_grid(obj: Any, arbitrary: bool = False) -> OutputGrid:
    """PyOutputGrid::extract, src/interface/mod.rs:31-65.  With `arbitrary`,
    a sequence of times need not be uniformly spaced."""
    if isinstance(obj, tuple):
        if len(obj) != 3:
            raise ValueError("time-grid tuple must have 3 elements: (start, end, steps)")
        start, end, steps = obj
        if isinstance(steps, bool) or not isinstance(steps, (int, np.integer)):
            raise TypeError("time-grid steps must be an integer")
        if steps < 0:
            raise OverflowError("time-grid steps must be non-negative")
        return _grid_new(float(start), float(end), int(steps))
    if isinstance(obj, np.ndarray) and obj.dtype == np.float64 and obj.ndim == 1:
        if not obj.flags.c_contiguous:
            raise ValueError("NumPy time arrays must be contiguous")
        return _grid_from_times(obj.tolist(), arbitrary)
    try:
        seq = [float(v) for v in obj]
    except TypeError:
        raise ValueError(
            "Expected (start, end, steps) tuple or 1D numpy.linspace array"
        ) from None
    return _grid_from_times(seq, arbitrary)


def _parse_tolerance(obj: Any) -> tuple[float, float]:
    """PyTolerance::extract, src/interface/mod.rs:139-171."""
    if isinstance(obj, tuple) and len(obj) == 2:
        rtol, atol = float(obj[0]), float(obj[1])
    else:
        try:
            rtol = atol = float(obj)
        except (TypeError, ValueError):
            raise ValueError(
                "Expected (rtol, atol) tuple or single float tolerance value"
            ) from None
    if not (math.isfinite(rtol) and math.isfinite(atol)) or rtol <= 0.0 or atol <= 0.0:
        raise ValueError("rtol and atol must be finite and greater than zero")
    return rtol, atol


def log_tolerance(rtol: float, atol: float) -> tuple[float, float]:
    """LogTolerance::from_linear, src/integrators/galpy/mod.rs:19-24."""
    return math.log(rtol), math.log(atol)


def _selector(value: Any, cls: type, name: str):
    """extract_value, src/interface/selector.rs:4-22."""
    if value is None:
        return list(cls)[0]
    if type(value) is not cls:
        raise TypeError(f"expected a drift.{name} member (astrodrift_amd.{name})")
    return value


def _rust_display(x: float) -> str:
    if math.isfinite(x) and x == int(x) and abs(x) < 1e16:
        return str(int(x))
    return repr(float(x))


def _rust_duration(seconds: float) -> str:
    """A duration as the reference's records print one (Rust `{:.3?}` of a
    Duration: the largest of s, ms, µs, ns below the value, 3 decimals)."""
    for scale, unit in ((1.0, "s"), (1e-3, "ms"), (1e-6, "\u00b5s")):
        if seconds >= scale:
            return f"{seconds / scale:.3f}{unit}"
    return f"{seconds * 1e9:.3f}ns"


# ---------------------------------------------------------------------------
# Potentials and containers (src/interface/recipe.rs, container.rs)
# ---------------------------------------------------------------------------


def _python_float(value: float) -> str:
    if math.isnan(value):
        return "float('nan')"
    if value == math.inf:
        return "float('inf')"
    if value == -math.inf:
        return "float('-inf')"
    return repr(float(value))


class Potential:
    """A gravitational potential definition.

    Construct potentials with :meth:`Potential.kepler`,
    :meth:`Potential.plummer`, or :meth:`Potential.bovy`.
    """

    __slots__ = ("_name", "_amp", "_radius")

    def __init__(self, name: str = "kepler", amp: float = 1.0, radius: float = 0.0):
        self._name = name
        self._amp = float(amp)
        self._radius = float(radius)

    def __repr__(self) -> str:
        if self._name == "kepler":
            return f"Potential.kepler(amp={_python_float(self._amp)})"
        if self._name == "plummer":
            return (
                f"Potential.plummer(amp={_python_float(self._amp)}, "
                f"radius={_python_float(self._radius)})"
            )
        return "Potential.bovy()"

    @staticmethod
    def kepler(amp: float) -> "Potential":
        """Point-mass potential (G = 1); ``amp`` is the total mass."""
        return Potential("kepler", float(amp))

    @staticmethod
    def plummer(amp: float, radius: float) -> "Potential":
        """Plummer sphere with total mass ``amp`` and scale radius ``radius``."""
        return Potential("plummer", float(amp), float(radius))

    @staticmethod
    def bovy() -> "Potential":
        """The built-in composite background potential: cut-off power-law
        bulge + Miyamoto-Nagai disk + NFW halo (shared/src/potential.rs:87-124)."""
        return Potential("bovy")

    def spec(self) -> _native.PotentialSpec:
        """The kernel force model for this potential."""
        if self._name == "kepler":
            return _native.PotentialSpec.kepler(self._amp)
        if self._name == "plummer":
            return _native.PotentialSpec.plummer(self._amp, self._radius)
        return _native.PotentialSpec.milky_way()


_next_identity = itertools.count()


class Container:
    """A state-bearing particle group or stationary background potential."""

    __slots__ = ("_num_particles", "_potential", "_state", "_identity")

    def __init__(self, num_particles, potential, state):
        self._num_particles = num_particles
        self._potential = potential
        self._state = state
        self._identity = next(_next_identity)

    @property
    def num_particles(self) -> int | None:
        """Number of particles in this group (``None`` for background containers)."""
        return self._num_particles

    def __repr__(self) -> str:
        if self._num_particles is not None and self._potential is not None:
            return (
                f"Container(kind='particles', num_particles={self._num_particles}, "
                f"potential={self._potential!r})"
            )
        if self._num_particles is not None:
            return f"Container(kind='test_particles', num_particles={self._num_particles})"
        if self._potential is not None:
            return f"Container(kind='background', potential={self._potential!r})"
        return "Container(kind='background', potential=None)"


def _validate_istate(istate: Any) -> np.ndarray:
    """validate_istate, src/interface/container.rs:59-93 (no particle limit)."""
    if not isinstance(istate, np.ndarray):
        raise TypeError("istate must be a numpy.ndarray of float64")
    if istate.dtype != np.float64:
        raise TypeError("istate must be a float64 array")
    shape = istate.shape
    if istate.ndim == 1:
        particles, ok = shape[0] // INPUT_STATE_DIM, shape[0] % INPUT_STATE_DIM == 0
    elif istate.ndim == 2:
        particles, ok = shape[0], shape[1] == INPUT_STATE_DIM
    else:
        particles, ok = 0, False
    if not ok:
        raise ValueError(
            "istate must have shape (N, 6) with columns [x, y, z, vx, vy, vz] "
            f"(or a flat array of length 6N); got shape {list(shape)}"
        )
    if particles == 0:
        raise ValueError("istate must contain at least one particle")
    if not np.all(np.isfinite(istate)):
        raise ValueError("istate positions and velocities must be finite")
    state = np.array(istate, dtype=np.float64, order="C").reshape(particles, INPUT_STATE_DIM)
    state.flags.writeable = False
    return state


def test_particles(istate: np.ndarray) -> Container:
    """A state-bearing particle group without an attached potential."""
    state = _validate_istate(istate)
    return Container(state.shape[0], None, state)


test_particles.__test__ = False  # not a pytest test despite its name


def particles(potential: Potential, istate: np.ndarray) -> Container:
    """A state-bearing particle group with an attached potential."""
    if not isinstance(potential, Potential):
        raise TypeError("potential must be a Potential")
    if potential._name not in ("kepler", "plummer"):
        raise NotImplementedError(
            "only Kepler and Plummer potentials can currently be attached to particles"
        )
    state = _validate_istate(istate)
    return Container(state.shape[0], potential, state)


def background(potential: Potential) -> Container:
    """A stationary background potential for use as a force source."""
    if not isinstance(potential, Potential):
        raise TypeError("potential must be a Potential")
    return Container(None, potential, None)


# ---------------------------------------------------------------------------
# Config (src/interface/mod.rs:178-469)
# ---------------------------------------------------------------------------


def validate_configuration(engine: Engine, method: Method, implementation: Implementation) -> None:
    """validate_configuration, src/integrators/mod.rs:23-36."""
    if implementation is not Implementation.GALPY:
        raise NotImplementedError(
            f"{engine.value} + {implementation.value} {method.value} is not implemented"
        )


class Config:
    """Configuration and registered container graph for an integration.

    The default selectors are ``Engine.CPU``, ``Method.DOPR54``, and
    ``Implementation.GALPY``.  Register force relationships with :meth:`add`,
    then execute the complete model with :meth:`run`.
    """

    def __init__(
        self,
        engine: Engine | None = None,
        method: Method | None = None,
        implementation: Implementation | None = None,
        ts: Any = None,
        tolerance: Any = None,
        *,
        forces: str = "reference",
        arbitrary_ts: bool = False,
        device: int = 0,
    ):
        self.engine = _selector(engine, Engine, "Engine")
        self.method = _selector(method, Method, "Method")
        self.implementation = _selector(implementation, Implementation, "Implementation")
        self.grid = OutputGrid() if ts is None else _parse_grid(ts, bool(arbitrary_ts))
        self.rtol, self.atol = (MIN_RTOL, MIN_ATOL) if tolerance is None else _parse_tolerance(tolerance)
        if forces not in ("reference", "model"):
            raise ValueError("forces must be 'reference' or 'model'")
        self.forces = forces
        self.device = int(device)
        validate_configuration(self.engine, self.method, self.implementation)
        self._dependencies: list[tuple[int, int]] = []  # (source, dependent)
        self._containers: list[Container] = []

    # -- graph registration -------------------------------------------------

    def _path_exists(self, start: int, target: int) -> bool:
        pending, visited = [start], set()
        while pending:
            current = pending.pop()
            if current == target:
                return True
            if current in visited:
                continue
            visited.add(current)
            pending.extend(dep for src, dep in self._dependencies if src == current)
        return False

    def add(self, node: Container, *requires: Container) -> None:
        """Register that ``node`` is integrated with ``requires`` as inputs."""
        if not isinstance(node, Container) or not all(isinstance(r, Container) for r in requires):
            raise TypeError("add() takes Container arguments")
        unique: list[Container] = []
        for r in requires:
            if all(r._identity != u._identity for u in unique):
                unique.append(r)
        containers = list(self._containers)
        for c in [node, *unique]:
            if all(c._identity != k._identity for k in containers):
                containers.append(c)
        if len(containers) > MAX_MODEL_COMPONENTS:
            raise ValueError(
                f"add() would register {len(containers)} containers but a model supports "
                f"at most {MAX_MODEL_COMPONENTS}"
            )
        if not unique:
            raise ValueError("add() requires at least one dependency")
        if any(self._path_exists(node._identity, r._identity) for r in unique):
            raise ValueError("adding these dependencies would create a cycle")
        for r in unique:
            edge = (r._identity, node._identity)
            if edge not in self._dependencies:
                self._dependencies.append(edge)
        self._containers = containers

    def dependency(self, node: Container, *requires: Container) -> None:
        """Deprecated alias for :meth:`add`."""
        warnings.warn(
            "Config.dependency() is deprecated; use Config.add(node, *requires)",
            DeprecationWarning,
            stacklevel=2,
        )
        self.add(node, *requires)

    def info(self) -> str:
        """Return a human-readable summary of this configuration."""
        return (
            f"Config(engine={self.engine.value}, method={self.method.value}, "
            f"implementation={self.implementation.value}, "
            f"times=({_rust_display(self.grid.start)}, {_rust_display(self.grid.end)}, "
            f"{self.grid.steps}), tolerance=(rtol={self.rtol:.6e}, atol={self.atol:.6e}), "
            f"containers={len(self._containers)}, dependencies={len(self._dependencies)})"
        )

    def __repr__(self) -> str:
        return self.info()

    # -- execution ----------------------------------------------------------

    def _sources(self, node: Container) -> list[Container]:
        return [c for c in self._containers if (c._identity, node._identity) in self._dependencies]

    def _integration_order(self) -> list[Container]:
        """State-bearing containers, each after the particle groups it feels
        (forces="model"); first-registration order otherwise."""
        pending = [c for c in self._containers if c._state is not None]
        if self.forces == "reference":
            return pending
        order: list[Container] = []
        done: set[int] = set()
        while pending:
            ready = next(c for c in pending
                         if all(s._identity in done for s in self._sources(c)
                                if s._state is not None))  # add() keeps the graph acyclic
            order.append(ready)
            done.add(ready._identity)
            pending.remove(ready)
        return order

    def _force_for(self, node: Container, specs: dict | None = None,
                   outputs: dict | None = None,
                   times: np.ndarray | None = None) -> _native.PotentialSpec | None:
        """The force a container feels: None (the reference's unit galpy
        Kepler) for forces="reference"; otherwise the sum over its force
        sources, a background's static potential or a particle group's moving
        origins (one per particle, on quintic Hermite tracks of the group's
        integrated trajectory)."""
        if self.forces == "reference":
            return None  # unit galpy Kepler, as the reference kernels hard-code
        statics = [c._potential for c in self._sources(node)
                   if c._state is None and c._potential is not None]
        groups = [c for c in self._sources(node) if c._state is not None and c._potential is not None]
        bovy = [p for p in statics if p._name == "bovy"]
        points = [p for p in statics if p._name != "bovy"]
        if len(bovy) > 1:
            raise NotImplementedError(
                "forces='model' supports at most one bovy background per integrated container"
            )
        if not groups and len(points) + len(bovy) == 1:
            return statics[0].spec()
        base = bovy[0].spec() if bovy else _native.PotentialSpec.none()
        if not groups and not points:
            return base
        if times is None:
            times = self.grid.times()
        sets = []
        for g in groups:
            states = outputs[g._identity]
            nt, n = states.shape[0], states.shape[1]
            pos = states[:, :, :3].reshape(-1, 3)
            t_of = np.repeat(times, n)
            acc = _native.force_host(specs[g._identity], pos, t_of,
                                     device=self.device).reshape(nt, n, 3)
            amp = g._potential._amp
            b = g._potential._radius if g._potential._name == "plummer" else 0.0
            try:
                sets.append(_native.Origins.hermite(states, acc, times, (amp, b)))
            except _native.DriftNativeError as exc:
                raise NotImplementedError(
                    f"particle groups as force sources need a uniform time grid: {exc}"
                ) from None
        span = (float(times.min()), float(times.max()))
        for p in points:
            b = p._radius if p._name == "plummer" else 0.0
            sets.append(_native.Origins.fixed((p._amp, b), max(len(times) - 1, 1), *span))
        return base.with_origins(_native.Origins.concat(sets))

    def run(self) -> list:
        """Integrate the registered model.

        Results follow first-registration order.  Each state-bearing container
        produces a float64 array shaped ``(time, particle, 6)``; stationary
        background containers produce ``None``.

        Raises ``IntegrationError`` when the backend cannot complete the
        integration.
        """
        for c in self._containers:
            if c._state is not None and not any(dep == c._identity for _, dep in self._dependencies):
                raise ValueError(
                    "each particle container must be registered with add(node, *requires) "
                    "before integration"
                )
        if not any(c._state is not None for c in self._containers):
            return [None for _ in self._containers]
        if self.engine is Engine.CPU:
            raise NotImplementedError(
                "astrodrift_amd provides the GPU engine only; use Engine.GPU"
            )
        if self.method is Method.DOPR54 and not self.grid.is_affine():
            raise NotImplementedError(
                "arbitrary output times need Method.DOP853: DOPR54 steps the fixed interval "
                "t[1] - t[0] (src/integrators/galpy/dopr54.rs:713)"
            )
        times = self.grid.times()
        rtol_ln, atol_ln = log_tolerance(self.rtol, self.atol)
        method = _native.METHOD_DOPR54 if self.method is Method.DOPR54 else _native.METHOD_DOP853
        _log_gpu.info(
            "GPU integration starting: implementation=%s, method=%s, %d output times over "
            "[%s, %s], rtol=%.2e, atol=%.2e",
            self.implementation.value,
            self.method.value,
            self.grid.steps,
            _rust_display(self.grid.start),
            _rust_display(self.grid.end),
            self.rtol,
            self.atol,
        )
        started = time.perf_counter()
        outputs: dict[int, np.ndarray] = {}
        specs: dict[int, Any] = {}
        stage = 0
        total = 0
        for c in self._integration_order():
            stage += 1
            n = c._state.shape[0]
            total += n
            _log_gpu.debug("stage %d: dispatching %d particles to the GPU", stage, n)
            stage_started = time.perf_counter()
            try:
                potential = self._force_for(c, specs, outputs, times)
                # src/dispatch/gpu.rs:165-172, with this backend's launch shape
                grid, block = _native.launch_shape(method, n, potential, device=self.device)
                _log_gpu.debug(
                    "launching GALPY %s kernel: %d particles, grid=%d, block=%d, %d output times",
                    self.method.value, n, grid, block, len(times),
                )
                out, _stats, _code = _native.launch_host(
                    method, c._state, times, rtol_ln, atol_ln, potential, device=self.device
                )
                report = _native.last_launch(self.device)
            except _native.DriftNativeError as exc:
                raise IntegrationError(str(exc)) from exc
            self._log_backend_phases(report)
            outputs[c._identity] = out
            specs[c._identity] = potential
            _log_gpu.debug(
                "stage %d: %d particles finished in %s", stage, n,
                _rust_duration(time.perf_counter() - stage_started),
            )
        results = [outputs.get(c._identity) if c._state is not None else None
                   for c in self._containers]
        _log_gpu.info(
            "GPU integration finished: %d stage(s), %d particle trajectories in %s",
            stage,
            total,
            _rust_duration(time.perf_counter() - started),
        )
        return results

    def _log_backend_phases(self, report: dict) -> None:
        """The backend's per-phase records (src/dispatch/gpu/cuda_oxide.rs:15-34,
        77-84) on drift.dispatch.gpu.hip.  The session, which holds the device
        context and stream, is created once per process: its creation time is
        added to the first launch's setup."""
        if not _log_hip.isEnabledFor(logging.DEBUG):
            return
        sess = _native.session(self.device)
        setup = report["setup_s"]
        if not sess.creation_reported:
            setup += sess.created_s
            sess.creation_reported = True
        _log_hip.debug("context initialized and kernel module loaded in %s",
                       _rust_duration(setup))
        _log_hip.debug("copied %d bytes of host data to the device in %s", report["h2d_bytes"],
                       _rust_duration(report["h2d_s"]))
        _log_hip.debug("integrated on the device (grid=%d, block=%d) in %s", report["grid"],
                       report["block"], _rust_duration(report["kernel_s"]))
        _log_hip.debug("synchronized device work and copied %d result bytes to the host in %s",
                       report["d2h_bytes"], _rust_duration(report["d2h_s"]))


__all__ = [
    "Config",
    "Container",
    "DriftError",
    "IntegrationError",
    "OutputGrid",
    "Potential",
    "background",
    "log_tolerance",
    "particles",
    "test_particles",
    "validate_configuration",
]
