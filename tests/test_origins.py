"""Moving origins on the CPU: the track builder (libdrift_hip.so's host-only
drift_origin_tracks_hermite) and the oracle's CustomOrigin restatement.

The reference defines CustomOrigin (shared/src/potential.rs:264-356) but never
constructs one, so there is no reference output to compare with.  The oracle's
moving-origin force is pinned instead by what the physics and the algebra fix:
it reduces exactly to the static force for an origin at rest, it is linear in
the origins' amplitudes, it follows the origin's track across segments, an
orbit about a uniformly moving origin is the static orbit carried along
(Galilean invariance, for both integrators, which pins the stage times), and
where the reference's own reading of the table is unambiguous (t_start = 0,
unit segments, as many origins as segments) both readings give the same
positions.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

import oracle_lib
from astrodrift_amd import _native
from astrodrift_amd._native import Origins, PotentialSpec


def eval_tracks(o: Origins, t: float) -> np.ndarray:
    """Origin positions (n_origins, 3) at time t, from the layout the header
    documents: u = (t - t_start) n_segments / (t_end - t_start), s = floor(u)
    clamped, tau = u - s, coefficients tau^5 .. tau^0 per coordinate."""
    dt = (o.t_end - o.t_start) / o.n_segments
    u = (t - o.t_start) / dt
    s = min(max(math.floor(u), 0), o.n_segments - 1)
    tau = u - s
    powers = np.array([tau**5, tau**4, tau**3, tau**2, tau, 1.0])
    return o.tracks[:, s, :].reshape(-1, 3, 6) @ powers


def circular_samples(times, radius=1.0, omega=0.7, phase=0.3):
    """(nt, 1, 6) states and (nt, 1, 3) accelerations of uniform circular motion."""
    th = omega * times + phase
    x = np.stack([radius * np.cos(th), radius * np.sin(th), 0.05 * np.sin(2 * th)], axis=-1)
    v = np.stack([-radius * omega * np.sin(th), radius * omega * np.cos(th),
                  0.1 * omega * np.cos(2 * th)], axis=-1)
    a = np.stack([-radius * omega**2 * np.cos(th), -radius * omega**2 * np.sin(th),
                  -0.2 * omega**2 * np.sin(2 * th)], axis=-1)
    return np.concatenate([x, v], axis=-1)[:, None, :], a[:, None, :]


# ---------------------------------------------------------------------------
# Track builder
# ---------------------------------------------------------------------------


def test_hermite_tracks_reproduce_quintic_motion():
    rng = np.random.default_rng(5)
    coef = rng.normal(size=(2, 3, 6))  # 2 orbits, x/y/z, t^0..t^5
    times = np.linspace(-1.0, 3.0, 9)

    def motion(t):
        p = np.array([t**k for k in range(6)])
        dp = np.array([k * t ** (k - 1) if k else 0.0 for k in range(6)])
        ddp = np.array([k * (k - 1) * t ** (k - 2) if k > 1 else 0.0 for k in range(6)])
        return coef @ p, coef @ dp, coef @ ddp

    states = np.zeros((times.size, 2, 6))
    accels = np.zeros((times.size, 2, 3))
    for k, t in enumerate(times):
        x, v, a = motion(t)
        states[k, :, :3], states[k, :, 3:], accels[k] = x, v, a
    o = Origins.hermite(states, accels, times, (1.0, 0.0))
    assert (o.t_start, o.t_end, o.n_segments) == (-1.0, 3.0, 8)
    for t in np.linspace(-1.0, 3.0, 57):
        x, _, _ = motion(t)
        np.testing.assert_allclose(eval_tracks(o, t), x, rtol=0, atol=1e-11 * (1 + abs(x).max()))


def test_hermite_tracks_match_samples_and_derivatives():
    times = np.linspace(0.0, 10.0, 21)
    states, accels = circular_samples(times)
    o = Origins.hermite(states, accels, times, (0.5, 0.1))
    dt = 0.5
    for s in range(o.n_segments):
        c = o.tracks[0, s].reshape(3, 6)  # tau^5 .. tau^0
        p0, v0, a0 = c[:, 5], c[:, 4], 2 * c[:, 3]
        p1 = c.sum(axis=1)
        v1 = c @ np.array([5, 4, 3, 2, 1, 0])
        a1 = c @ np.array([20, 12, 6, 2, 0, 0])
        np.testing.assert_allclose(p0, states[s, 0, :3], atol=1e-15)
        np.testing.assert_allclose(v0, states[s, 0, 3:] * dt, atol=1e-15)
        np.testing.assert_allclose(a0, accels[s, 0] * dt * dt, atol=1e-15)
        np.testing.assert_allclose(p1, states[s + 1, 0, :3], atol=1e-14)
        np.testing.assert_allclose(v1, states[s + 1, 0, 3:] * dt, atol=1e-14)
        np.testing.assert_allclose(a1, accels[s + 1, 0] * dt * dt, atol=1e-14)
    np.testing.assert_array_equal(o.params, [[0.5, 0.1]])


def test_hermite_tracks_of_a_decreasing_grid_run_forward_in_time():
    times = np.linspace(0.0, 6.0, 13)
    states, accels = circular_samples(times)
    fwd = Origins.hermite(states, accels, times, (1.0, 0.0))
    bwd = Origins.hermite(states[::-1].copy(), accels[::-1].copy(), times[::-1].copy(), (1.0, 0.0))
    assert (bwd.t_start, bwd.t_end) == (0.0, 6.0)
    np.testing.assert_array_equal(fwd.tracks, bwd.tracks)


def test_hermite_tracks_validate_their_input():
    times = np.array([0.0, 1.0, 2.5])
    states, accels = circular_samples(times)
    with pytest.raises(_native.DriftNativeError, match="uniformly"):
        Origins.hermite(states, accels, times, (1.0, 0.0))
    times = np.array([0.0, 1.0, 1.0])
    with pytest.raises(_native.DriftNativeError, match="monotone"):
        Origins.hermite(states, accels, times, (1.0, 0.0))
    times = np.array([0.0, 1.0, 2.0])
    accels[1, 0, 0] = np.nan
    with pytest.raises(_native.DriftNativeError, match="finite"):
        Origins.hermite(states, accels, times, (1.0, 0.0))


# ---------------------------------------------------------------------------
# Oracle CustomOrigin
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("amp,b", [(1.3, 0.0), (0.7, 0.25)])
def test_origin_at_rest_is_the_static_force(amp, b):
    moving = PotentialSpec.none().with_origins(Origins.fixed((amp, b), 4, 0.0, 8.0))
    static = PotentialSpec.kepler(amp) if b == 0.0 else PotentialSpec.plummer(amp, b)
    rng = np.random.default_rng(1)
    for _ in range(50):
        x, y, z = rng.uniform(-2, 2, 3)
        t = rng.uniform(-3.0, 11.0)  # inside and outside the tabulated span
        np.testing.assert_array_equal(oracle_lib.force_at(moving, t, x, y, z),
                                      oracle_lib.force(static, x, y, z))


def test_origin_forces_add_to_the_static_force_and_are_linear():
    times = np.linspace(0.0, 4.0, 9)
    states, accels = circular_samples(times)
    one = Origins.hermite(states, accels, times, (0.3, 0.05))
    split = Origins.concat([Origins.hermite(states, accels, times, (0.1, 0.05)),
                            Origins.hermite(states, accels, times, (0.2, 0.05))])
    mw = PotentialSpec.milky_way()
    rng = np.random.default_rng(2)
    for _ in range(30):
        x, y, z = rng.uniform(-2, 2, 3)
        t = rng.uniform(0.0, 4.0)
        f_one = oracle_lib.force_at(mw.with_origins(one), t, x, y, z)
        f_split = oracle_lib.force_at(mw.with_origins(split), t, x, y, z)
        np.testing.assert_allclose(f_one, f_split, rtol=1e-13, atol=1e-15)
        alone = oracle_lib.force_at(PotentialSpec.none().with_origins(one), t, x, y, z)
        np.testing.assert_allclose(f_one, oracle_lib.force(mw, x, y, z) + alone, rtol=1e-14,
                                   atol=1e-15)


def test_origin_follows_its_track_across_segments():
    times = np.linspace(0.0, 20.0, 201)
    states, accels = circular_samples(times)
    o = Origins.hermite(states, accels, times, (0.8, 0.0))
    spec = PotentialSpec.none().with_origins(o)
    rng = np.random.default_rng(3)
    for t in rng.uniform(0.0, 20.0, 40):
        x, y, z = rng.uniform(-2, 2, 3)
        c, _ = circular_samples(np.array([t]))
        d = np.array([x, y, z]) - c[0, 0, :3]
        r = np.linalg.norm(d)
        expect = -0.8 * d / r**3
        np.testing.assert_allclose(oracle_lib.force_at(spec, t, x, y, z), expect, rtol=1e-9)


def test_reference_reading_of_the_table_agrees_where_it_is_unambiguous():
    """CustomOrigin::origins (shared/src/potential.rs:290-331) with t_start = 0,
    final_time = division (unit segments) and as many origins as segments."""
    rng = np.random.default_rng(4)
    n = 3
    tracks = rng.normal(size=(n, n, 18))
    o = Origins(tracks, np.tile([1.0, 0.0], (n, 1)), 0.0, float(n))
    table = tracks.reshape(-1)
    for t in rng.uniform(0.0, n, 20):
        dt = o.t_end / n
        t0 = int(math.floor(t / dt))
        pos = []
        for i in range(n):
            p0 = 18 * (i * n + t0)
            tl = t - t0
            pos.append([sum(table[p0 + 6 * c + k] * tl ** (5 - k) for k in range(6))
                        for c in range(3)])
        pos = np.array(pos)
        np.testing.assert_allclose(eval_tracks(o, t), pos, rtol=1e-13, atol=1e-13)
        x = np.array([5.0, -4.0, 3.0])
        expect = sum(-(x - p) / np.linalg.norm(x - p) ** 3 for p in pos)
        np.testing.assert_allclose(
            oracle_lib.force_at(PotentialSpec.none().with_origins(o), t, *x), expect, rtol=1e-12)


def uniform_track(o0, u, t0, t1, nseg, params):
    times = np.linspace(t0, t1, nseg + 1)
    x = o0[None, :] + u[None, :] * times[:, None]
    states = np.concatenate([x, np.broadcast_to(u, x.shape)], axis=-1)[:, None, :]
    return Origins.hermite(states, np.zeros((times.size, 1, 3)), times, params)


@pytest.mark.parametrize("method", [_native.METHOD_DOP853, _native.METHOD_DOPR54])
def test_orbit_about_a_uniformly_moving_origin_is_the_static_orbit_carried_along(method):
    o0 = np.array([0.4, -0.3, 0.1])
    u = np.array([0.05, 0.02, -0.03])
    track = uniform_track(o0, u, 0.0, 12.0, 12, (1.0, 0.0))
    moving = PotentialSpec.none().with_origins(track)
    static = PotentialSpec.kepler(1.0)
    times = np.linspace(0.0, 10.0, 41)
    tol = math.log(1e-12)
    rel = np.array([1.0, 0.1, 0.05, 0.05, 1.05, 0.1])
    shift = np.concatenate([o0, u])
    ref, _, err_s = oracle_lib.orbit(method, rel, times, tol, tol, static)
    got, _, err_m = oracle_lib.orbit(method, rel + shift, times, tol, tol, moving)
    assert err_s == 0 and err_m == 0
    carried = ref + np.concatenate([o0[None, :] + u[None, :] * times[:, None],
                                    np.broadcast_to(u, (times.size, 3))], axis=-1)
    np.testing.assert_allclose(got, carried, rtol=0, atol=2e-8)
    # and the origin's motion matters: the static orbit from the same start differs
    unmoved, _, _ = oracle_lib.orbit(method, rel + shift, times, tol, tol,
                                     PotentialSpec.none().with_origins(
                                         uniform_track(o0, 0 * u, 0.0, 12.0, 12, (1.0, 0.0))))
    assert np.abs(unmoved - got).max() > 1e-2
