"""Seeded synthetic initial conditions for the benchmark workloads.

* ``kepler_disk``: 1,024 orbits in the unit Kepler potential (G M = 1),
  R ~ U[0.5, 2], phi ~ U[0, 2 pi), z = 0, tangential speed R^-1/2 (1 + N(0, 0.05)),
  vertical speed N(0, 0.02); the velocity is then tilted out of the plane by an
  inclination ~ U[0, 0.3] rad about the radial direction.  DOPR54,
  rtol = atol = 1e-10, ts = (0, 20 pi, 101).
* ``tidal_stream``: one serving request: N stars placed around a progenitor
  on an eccentric Milky-Way orbit (R ~ U[0.5, 2], speed 0.7-1.0 of circular,
  tangential) as a Plummer sphere of scale 0.002 with isotropic velocity
  dispersion 0.005; 101 outputs over a span ~ U[10, 100].
* ``ragged_mix``: a ragged batch of jobs: n_j log-uniform on [16, 65,536],
  potential uniform over {Kepler (amp 1), Plummer (amp 1, b 0.1), Milky Way},
  method uniform over {DOPR54, DOP853}, span ~ U[5, 200], nt ~ U{11..501},
  rtol = atol log-uniform on [1e-12, 1e-8]; near-circular disk orbits
  (R ~ U[0.3, 2], z ~ N(0, 0.03), v_phi = vc(R)(1 + N(0, 0.05)),
  v_R, v_z ~ N(0, 0.05)).  4,096 jobs (~32M orbits, ~400 GB of output) is the
  full mix; the benchmark runs a prefix that fits one GPU.
* ``mw_disk``: 1,048,576 orbits in the Milky-Way model (bulge alpha = 1.8,
  rc = 0.2375, vc^2 fraction 0.05; Miyamoto-Nagai a = 0.375, b = 0.035, 0.6;
  NFW a = 2.0, 0.35), exponential-disk initial conditions: R ~ R exp(-R/0.3125)
  truncated to [0.1, 3], z ~ sech^2(z/0.0375), v_phi = vc(R) + N(0, 0.05),
  v_R, v_z ~ N(0, 0.05).  DOP853, rtol = atol = 1e-12, ts = (0, 50, 1001).

numpy's PCG64 generator with the fixed seeds below makes every draw
reproducible; the same seed gives the same orbits on any host.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from ._native import PotentialSpec

KEPLER_SEED = 0x1024
MW_SEED = 0x100000
RAGGED_SEED = 0x4096


@dataclass
class Workload:
    name: str
    method: str          # "DOPR54" or "DOP853"
    state0: np.ndarray   # (n, 6)
    times: np.ndarray    # (nt,)
    rtol: float
    atol: float
    potential: PotentialSpec | None  # None = the reference's unit galpy Kepler force

    @property
    def n(self) -> int:
        return self.state0.shape[0]

    @property
    def nt(self) -> int:
        return self.times.shape[0]


def grid_times(start: float, end: float, steps: int) -> np.ndarray:
    """OutputGrid::sample for every index (shared/src/config.rs:57-66)."""
    t = np.array(
        [start + (end - start) * float(i) / float(steps - 1) for i in range(steps)],
        dtype=np.float64,
    )
    t[-1] = end
    return t


def kepler_disk(n: int = 1024, seed: int = KEPLER_SEED) -> Workload:
    rng = np.random.default_rng(seed)
    R = rng.uniform(0.5, 2.0, n)
    phi = rng.uniform(0.0, 2.0 * math.pi, n)
    vt = R**-0.5 * (1.0 + rng.normal(0.0, 0.05, n))
    vz = rng.normal(0.0, 0.02, n)
    inc = rng.uniform(0.0, 0.3, n)
    c, s = np.cos(phi), np.sin(phi)
    # unit vectors: radial (c, s, 0), azimuthal (-s, c, 0), vertical (0, 0, 1)
    # tilt the (azimuthal, vertical) velocity by `inc` about the radial axis
    v_az = vt * np.cos(inc) - vz * np.sin(inc)
    v_vert = vt * np.sin(inc) + vz * np.cos(inc)
    state = np.stack([R * c, R * s, np.zeros(n), -s * v_az, c * v_az, v_vert], axis=1)
    return Workload(
        "kepler_disk_dopr54", "DOPR54", np.ascontiguousarray(state),
        grid_times(0.0, 20.0 * math.pi, 101), 1e-10, 1e-10, None,
    )


def mw_vcirc(spec: PotentialSpec, R: np.ndarray) -> np.ndarray:
    """Circular speed in the z = 0 plane of a Milky-Way PotentialSpec."""
    from scipy.special import gamma as gamma_fn
    from scipy.special import gammainc

    R = np.asarray(R, dtype=np.float64)
    a_b = 1.5 - 0.5 * spec.bulge_alpha
    m_b = (
        2.0 * math.pi * spec.bulge_amp * spec.bulge_rc ** (3.0 - spec.bulge_alpha)
        * gammainc(a_b, R**2 / spec.bulge_rc**2) * gamma_fn(a_b)
    )
    vc2_b = m_b / R
    vc2_d = spec.disk_amp * R**2 / (R**2 + (spec.disk_a + spec.disk_b) ** 2) ** 1.5
    u = R / spec.halo_a
    vc2_h = spec.halo_amp * (np.log1p(u) - u / (1.0 + u)) / R
    return np.sqrt(vc2_b + vc2_d + vc2_h)


def mw_disk(n: int = 1 << 20, seed: int = MW_SEED, t_end: float = 50.0, steps: int = 1001,
            rtol: float = 1e-12, atol: float = 1e-12) -> Workload:
    rng = np.random.default_rng(seed)
    spec = PotentialSpec.milky_way()
    R = np.empty(0)
    while R.size < n:  # Gamma(2, 0.3125) truncated to [0.1, 3] by rejection
        draw = rng.gamma(2.0, 0.3125, 2 * n)
        R = np.concatenate([R, draw[(draw >= 0.1) & (draw <= 3.0)]])
    R = R[:n]
    phi = rng.uniform(0.0, 2.0 * math.pi, n)
    z = 0.0375 * np.arctanh(2.0 * rng.uniform(1e-12, 1.0 - 1e-12, n) - 1.0)
    vphi = mw_vcirc(spec, R) + rng.normal(0.0, 0.05, n)
    vR = rng.normal(0.0, 0.05, n)
    vz = rng.normal(0.0, 0.05, n)
    c, s = np.cos(phi), np.sin(phi)
    state = np.stack([R * c, R * s, z, vR * c - vphi * s, vR * s + vphi * c, vz], axis=1)
    return Workload(
        "mw_disk_dop853", "DOP853", np.ascontiguousarray(state),
        grid_times(0.0, t_end, steps), rtol, atol, spec,
    )


def mw_potential(spec: PotentialSpec, x, y, z):
    """Phi (G = 1) of the Milky-Way model in closed form, for energy
    diagnostics: the power-law bulge with its Gaussian cutoff (scipy's
    regularised incomplete gamma functions), the Miyamoto-Nagai disk and the
    NFW halo.  Vectorised over numpy arrays."""
    from scipy.special import gamma as gamma_fn
    from scipy.special import gammainc, gammaincc

    r2 = x * x + y * y + z * z
    r = np.sqrt(r2)
    a_lo = 1.5 - 0.5 * spec.bulge_alpha
    a_hi = 1.0 - 0.5 * spec.bulge_alpha
    xx = r2 / spec.bulge_rc ** 2
    mass = (2.0 * math.pi * spec.bulge_amp * spec.bulge_rc ** (3.0 - spec.bulge_alpha)
            * gammainc(a_lo, xx) * gamma_fn(a_lo))
    outer = (2.0 * math.pi * spec.bulge_amp * spec.bulge_rc ** (2.0 - spec.bulge_alpha)
             * gammaincc(a_hi, xx) * gamma_fn(a_hi))
    phi_bulge = -mass / r - outer
    zb = np.sqrt(z * z + spec.disk_b ** 2)
    phi_disk = -spec.disk_amp / np.sqrt(x * x + y * y + (spec.disk_a + zb) ** 2)
    phi_halo = -spec.halo_amp * np.log1p(r / spec.halo_a) / r
    return phi_bulge + phi_disk + phi_halo


def mw_energy(spec: PotentialSpec, states: np.ndarray) -> np.ndarray:
    """Specific energy of (n, 6) phase-space states in the Milky-Way model."""
    return 0.5 * np.sum(states[:, 3:] ** 2, axis=1) + mw_potential(
        spec, states[:, 0], states[:, 1], states[:, 2])


def tidal_stream(n: int, rng: np.random.Generator, span: float | None = None,
                 steps: int = 101, rtol: float = 1e-10, atol: float = 1e-10) -> Workload:
    spec = PotentialSpec.milky_way()
    R = rng.uniform(0.5, 2.0)
    phi = rng.uniform(0.0, 2.0 * math.pi)
    v = rng.uniform(0.7, 1.0) * mw_vcirc(spec, np.array([R]))[0]
    c, s = math.cos(phi), math.sin(phi)
    prog = np.array([R * c, R * s, 0.0, -v * s, v * c, 0.0])
    # Plummer sphere: r = b (u^(-2/3) - 1)^(-1/2), isotropic directions,
    # truncated at 25 scale radii
    u = rng.uniform(1e-9, 1.0, n)
    r = np.minimum(0.002 / np.sqrt(u ** (-2.0 / 3.0) - 1.0), 0.05)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    state = np.tile(prog, (n, 1))
    state[:, :3] += r[:, None] * d
    state[:, 3:] += rng.normal(0.0, 0.005, (n, 3))
    t_end = span if span is not None else rng.uniform(10.0, 100.0)
    return Workload("tidal_stream", "DOP853", np.ascontiguousarray(state),
                    grid_times(0.0, t_end, steps), rtol, atol, spec)


RAGGED_POTENTIALS = (PotentialSpec.kepler(1.0), PotentialSpec.plummer(1.0, 0.1),
                     PotentialSpec.milky_way())


def vcirc(spec: PotentialSpec, R: np.ndarray) -> np.ndarray:
    """Circular speed in the z = 0 plane of a Kepler, Plummer or Milky-Way spec."""
    R = np.asarray(R, dtype=np.float64)
    if spec.kind == 1:  # Kepler
        return np.sqrt(spec.amp / R)
    if spec.kind == 2:  # Plummer
        return np.sqrt(spec.amp * R * R / (R * R + spec.b * spec.b) ** 1.5)
    return mw_vcirc(spec, R)


@dataclass
class RaggedJob:
    method: str          # "DOPR54" or "DOP853"
    potential: int       # index into RAGGED_POTENTIALS
    state0: np.ndarray   # (n, 6)
    times: np.ndarray    # (nt,)
    rtol: float

    @property
    def n(self) -> int:
        return self.state0.shape[0]

    @property
    def nt(self) -> int:
        return self.times.shape[0]


def disk_orbits(spec: PotentialSpec, n: int, seed: int) -> np.ndarray:
    """n disk orbits of the ragged mix's distribution in one potential: R
    uniform in [0.3, 2], z ~ N(0, 0.03), circular speed with 5% scatter,
    radial and vertical speeds ~ N(0, 0.05); (n, 6) AoS."""
    rng = np.random.default_rng(seed)
    R = rng.uniform(0.3, 2.0, n)
    phi = rng.uniform(0.0, 2.0 * math.pi, n)
    z = rng.normal(0.0, 0.03, n)
    vphi = vcirc(spec, R) * (1.0 + rng.normal(0.0, 0.05, n))
    vR, vz = rng.normal(0.0, 0.05, n), rng.normal(0.0, 0.05, n)
    c, s = np.cos(phi), np.sin(phi)
    return np.ascontiguousarray(
        np.stack([R * c, R * s, z, vR * c - vphi * s, vR * s + vphi * c, vz], axis=1))


def ragged_mix(njobs: int = 4096, seed: int = RAGGED_SEED) -> list[RaggedJob]:
    rng = np.random.default_rng(seed)
    jobs = []
    for _ in range(njobs):
        n = int(np.exp(rng.uniform(math.log(16), math.log(65536))))
        pot = int(rng.integers(0, 3))
        method = "DOPR54" if rng.integers(0, 2) == 0 else "DOP853"
        span = rng.uniform(5.0, 200.0)
        nt = int(rng.integers(11, 502))
        tol = math.exp(rng.uniform(math.log(1e-12), math.log(1e-8)))
        R = rng.uniform(0.3, 2.0, n)
        phi = rng.uniform(0.0, 2.0 * math.pi, n)
        z = rng.normal(0.0, 0.03, n)
        vphi = vcirc(RAGGED_POTENTIALS[pot], R) * (1.0 + rng.normal(0.0, 0.05, n))
        vR, vz = rng.normal(0.0, 0.05, n), rng.normal(0.0, 0.05, n)
        c, s = np.cos(phi), np.sin(phi)
        state = np.stack([R * c, R * s, z, vR * c - vphi * s, vR * s + vphi * c, vz], axis=1)
        jobs.append(RaggedJob(method, pot, np.ascontiguousarray(state),
                              grid_times(0.0, span, nt), tol))
    return jobs
