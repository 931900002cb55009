"""GPU parity of moving origins (CustomOrigin, shared/src/potential.rs:264-356)
through the C ABI against the CPU oracle, and the model-driven API that
builds them from an integrated particle group.

Bar: 1e-8 relative on positions and velocities at rtol = atol = 1e-10, as
for the static potentials; the API chain (satellite integrated, tracks built,
stream integrated about it) against the same chain run with the oracle, 1e-7
(the tracks are built from two integrations that each agree to 1e-8).
"""

from __future__ import annotations

import math

import numpy as np
import pytest

import astrodrift_amd as dft
from astrodrift_amd import _native
from astrodrift_amd._native import Origins, PotentialSpec

import oracle_lib
from test_gpu_parity import REL_TOL, rel_err

pytestmark = pytest.mark.gpu


def satellite_track(t1=10.0, nseg=100, amp=0.05, b=0.05) -> Origins:
    """A Plummer sphere on a slightly inclined circle of radius 1 (period 9)."""
    times = np.linspace(0.0, t1, nseg + 1)
    om = 2 * math.pi / 9.0
    th = om * times
    x = np.stack([np.cos(th), np.sin(th), 0.1 * np.sin(th)], axis=-1)
    v = np.stack([-om * np.sin(th), om * np.cos(th), 0.1 * om * np.cos(th)], axis=-1)
    a = -om * om * x
    states = np.concatenate([x, v], axis=-1)[:, None, :]
    return Origins.hermite(states, a[:, None, :], times, (amp, b))


def orbits_near_track(n, seed=11):
    rng = np.random.default_rng(seed)
    R = rng.uniform(0.7, 1.3, n)
    phi = rng.uniform(-0.6, 0.6, n)
    vc = 2 * math.pi / 9.0 * R
    y0 = np.zeros((n, 6))
    y0[:, 0], y0[:, 1] = R * np.cos(phi), R * np.sin(phi)
    y0[:, 2] = rng.normal(0, 0.05, n)
    y0[:, 3] = -vc * np.sin(phi) + rng.normal(0, 0.05, n)
    y0[:, 4] = vc * np.cos(phi) + rng.normal(0, 0.05, n)
    y0[:, 5] = rng.normal(0, 0.05, n)
    return y0


STATICS = {
    "mw": PotentialSpec.milky_way(),
    "none": PotentialSpec.none(),
    "kepler": PotentialSpec.kepler(0.4),
}


@pytest.mark.parametrize("method", [_native.METHOD_DOP853, _native.METHOD_DOPR54])
@pytest.mark.parametrize("static", sorted(STATICS))
def test_moving_origin_parity(gpu, method, static):
    spec = STATICS[static].with_origins(satellite_track())
    y0 = orbits_near_track(256)
    times = np.linspace(0.0, 10.0, 51)
    tol = math.log(1e-10)
    print(f"moving origins: static={static} method={method} n=256", flush=True)
    out, stats, code = _native.launch_host(method, y0, times, tol, tol, spec)
    cpu, cst = oracle_lib.batch(method, y0, times, tol, tol, spec, nthreads=8)
    ep, ev = rel_err(out, cpu)
    acc_g, acc_c = int(stats["accepted"].sum()), int(cst["accepted"].sum())
    print(f"  pos_rel={ep:.3e} vel_rel={ev:.3e} accepted gpu/cpu={acc_g}/{acc_c}")
    assert code == 0
    assert ep < REL_TOL and ev < REL_TOL
    assert abs(acc_g - acc_c) <= 1e-3 * acc_c


def satellites(k, t1=10.0, nseg=100, amp=0.02, b=0.05) -> Origins:
    """k Plummer spheres on inclined circles of radii 0.8..1.2 at spread
    phases (k origins on the same segments)."""
    times = np.linspace(0.0, t1, nseg + 1)
    r = np.linspace(0.8, 1.2, k)
    om = 2 * math.pi / 9.0 / r ** 1.5
    th = om[None, :] * times[:, None] + np.linspace(0.0, 2 * math.pi, k, endpoint=False)[None, :]
    x = np.stack([r * np.cos(th), r * np.sin(th), 0.1 * r * np.sin(th)], axis=-1)
    v = np.stack([-r * om * np.sin(th), r * om * np.cos(th), 0.1 * r * om * np.cos(th)], axis=-1)
    a = -(om ** 2)[None, :, None] * x
    return Origins.hermite(np.concatenate([x, v], axis=-1), a, times, (amp, b))


@pytest.mark.parametrize("static", ["mw", "none"])
@pytest.mark.parametrize("k,nseg,nt", [(3, 100, 401), (16, 200, 101)])
def test_windowed_launch_equals_one_launch(gpu, static, k, nseg, nt):
    """Two or more origins: a persistent DOP853 launch runs in time windows
    with the tracks' rows in LDS, pausing every orbit after each window's last
    sample and resuming it in the next (3 origins: windows of about 190
    samples; 16 origins: 50 windows of 2 samples, where many steps overshoot
    the staged rows and read the tracks).  Bit for bit the plain schedule's
    single launch, step counts included, and the oracle on a subset."""
    import torch

    spec = STATICS[static].with_origins(satellites(k, nseg=nseg))
    y0 = orbits_near_track(8192, seed=5)
    times = np.linspace(0.0, 10.0, nt)
    tol = math.log(1e-10)
    d_y0 = torch.from_numpy(y0).to(gpu)
    d_t = torch.from_numpy(times).to(gpu)
    res = {}
    for schedule in (_native.SCHEDULE_PLAIN, _native.SCHEDULE_PERSISTENT):
        d_out = torch.full((nt, y0.shape[0], 6), float("nan"), dtype=torch.float64, device=gpu)
        d_st = torch.zeros((y0.shape[0], 4), dtype=torch.int32, device=gpu)
        print(f"windowed moving origins: k={k} static={static} schedule={schedule}", flush=True)
        _native.launch_device(_native.METHOD_DOP853, d_y0.data_ptr(), d_t.data_ptr(),
                              d_out.data_ptr(), y0.shape[0], nt, tol, tol, spec, d_st.data_ptr(),
                              schedule=schedule, ordering=_native.ORDER_COST)
        torch.cuda.synchronize()
        res[schedule] = (d_out.cpu().numpy(), d_st.cpu().numpy())
    (po, ps), (qo, qs) = res[_native.SCHEDULE_PLAIN], res[_native.SCHEDULE_PERSISTENT]
    assert not np.isnan(po).any() and (ps[:, 3] == 0).all()
    assert np.array_equal(qo.view(np.uint64), po.view(np.uint64))
    assert np.array_equal(qs, ps)
    idx = np.arange(0, y0.shape[0], 256)
    cpu, _ = oracle_lib.batch(_native.METHOD_DOP853, y0[idx], times, tol, tol, spec, nthreads=8)
    ep, ev = rel_err(qo[:, idx], cpu)
    print(f"  pos_rel={ep:.3e} vel_rel={ev:.3e}")
    assert ep < REL_TOL and ev < REL_TOL


@pytest.mark.parametrize("schedule", [_native.SCHEDULE_PLAIN, _native.SCHEDULE_PERSISTENT])
def test_moving_origin_device_schedules(gpu, schedule):
    import torch

    spec = STATICS["mw"].with_origins(satellite_track())
    y0 = orbits_near_track(4096, seed=3)
    times = np.linspace(0.0, 10.0, 41)
    tol = math.log(1e-10)
    d_y0 = torch.from_numpy(y0).to(gpu)
    d_t = torch.from_numpy(times).to(gpu)
    d_out = torch.empty((times.size, y0.shape[0], 6), dtype=torch.float64, device=gpu)
    print(f"moving origins on device: schedule={schedule} n=4096", flush=True)
    _native.launch_device(_native.METHOD_DOP853, d_y0.data_ptr(), d_t.data_ptr(),
                          d_out.data_ptr(), y0.shape[0], times.size, tol, tol, spec,
                          schedule=schedule, ordering=_native.ORDER_COST)
    torch.cuda.synchronize()
    out = d_out.cpu().numpy()
    idx = np.arange(0, 4096, 16)
    cpu, _ = oracle_lib.batch(_native.METHOD_DOP853, y0[idx], times, tol, tol, spec, nthreads=8)
    ep, ev = rel_err(out[:, idx], cpu)
    print(f"  pos_rel={ep:.3e} vel_rel={ev:.3e}")
    assert ep < REL_TOL and ev < REL_TOL


def test_force_with_times_matches_oracle(gpu):
    spec = STATICS["mw"].with_origins(
        Origins.concat([satellite_track(), satellite_track(amp=0.02, b=0.0)]))
    rng = np.random.default_rng(8)
    pos = rng.uniform(-2, 2, (512, 3))
    t = rng.uniform(-1.0, 12.0, 512)  # including times outside the tracks
    got = _native.force_host(spec, pos, t)
    want = np.array([oracle_lib.force_at(spec, t[k], *pos[k]) for k in range(512)])
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14)
    static = _native.force_host(STATICS["mw"], pos)
    assert np.abs(got - static).max() > 1e-3


def test_orbits_about_a_uniformly_moving_origin_at_scale(gpu):
    """Galilean invariance on 65,536 orbits: about a point mass moving at
    constant velocity, each orbit is its static counterpart carried along."""
    o0 = np.array([0.3, -0.2, 0.1])
    u = np.array([0.04, -0.03, 0.02])
    times = np.linspace(0.0, 20.0, 81)
    grid = np.linspace(0.0, 21.0, 22)
    xs = o0[None, :] + u[None, :] * grid[:, None]
    track = Origins.hermite(
        np.concatenate([xs, np.broadcast_to(u, xs.shape)], axis=-1)[:, None, :],
        np.zeros((grid.size, 1, 3)), grid, (1.0, 0.0))
    rng = np.random.default_rng(21)
    n = 65536
    R = rng.uniform(0.5, 2.0, n)
    phi = rng.uniform(0, 2 * math.pi, n)
    rel = np.zeros((n, 6))
    rel[:, 0], rel[:, 1] = R * np.cos(phi), R * np.sin(phi)
    rel[:, 3], rel[:, 4] = -np.sin(phi) / np.sqrt(R), np.cos(phi) / np.sqrt(R)
    rel[:, 5] = rng.normal(0, 0.05, n)
    tol = math.log(1e-12)
    print("galilean: n=65536 DOP853", flush=True)
    static, _, _ = _native.launch_host(_native.METHOD_DOP853, rel, times, tol, tol,
                                       PotentialSpec.kepler(1.0))
    moving, stats, _ = _native.launch_host(_native.METHOD_DOP853, rel + np.concatenate([o0, u]),
                                           times, tol, tol, PotentialSpec.none().with_origins(track))
    carried = static.copy()
    carried[..., :3] += o0 + u * times[:, None, None]
    carried[..., 3:] += u
    err = np.abs(moving - carried).max()
    print(f"  max |moving - carried| = {err:.3e}, accepted {int(stats['accepted'].sum())}")
    assert err < 1e-7


def test_moving_origin_refusals(gpu):
    import torch

    spec = STATICS["mw"].with_origins(satellite_track())
    y0 = torch.zeros((64, 6), dtype=torch.float64, device=gpu)
    y0[:, 0] = 1.0
    y0[:, 4] = 1.0
    t = torch.linspace(0, 1, 5, dtype=torch.float64, device=gpu)
    out = torch.empty((5, 64, 6), dtype=torch.float64, device=gpu)
    with pytest.raises(_native.DriftNativeError, match="FP64"):
        _native.launch_device(_native.METHOD_DOP853, y0.data_ptr(), t.data_ptr(), out.data_ptr(),
                              64, 5, math.log(1e-7), math.log(1e-7), spec,
                              precision=_native.PRECISION_MIXED)
    job = _native.DeviceJob(_native.METHOD_DOP853, 0, 64, 5, math.log(1e-8), math.log(1e-8),
                            y0.data_ptr(), t.data_ptr(), out.data_ptr())
    s = _native.session(0)
    with _native._c_potential(s, spec) as pot:
        pots = (_native.CPotential * 1)(pot)
        arr = (_native.CJob * 1)(job.to_c())
        opts = _native.CLaunchOptions(0, 0, 0, 0, 0, 0)
        rc = s.lib.drift_ragged_device(s.handle, arr, 1, pots, 1, opts, None)
    assert rc == _native.DRIFT_ERR_UNSUPPORTED and "single-batch" in _native.last_error()
    bogus = STATICS["mw"].to_c(origins_id=12345)
    rc = s.lib.drift_galpy_dop853_device(s.handle, y0.data_ptr(), t.data_ptr(), out.data_ptr(),
                                         64, 5, -18.0, -18.0, bogus, None, None, None)
    assert rc == _native.DRIFT_ERR_INVALID_ARGUMENT and "not a moving-origin set" in _native.last_error()
    galpy = PotentialSpec.galpy_kepler().with_origins(satellite_track())
    with pytest.raises(_native.DriftNativeError, match="galpy"):
        _native.launch_host(_native.METHOD_DOP853, y0.cpu().numpy(), t.cpu().numpy(), -18.0,
                            -18.0, galpy)


def test_model_api_stream_about_a_satellite(gpu):
    """A satellite (a Plummer sphere attached to a particle group) integrated
    in the Milky-Way model, then a stream integrated in the Milky Way plus the
    satellite moving along its computed orbit."""
    mw = dft.Potential.bovy()
    sat0 = np.array([[1.0, 0.0, 0.05, 0.0, 1.0, 0.05]])
    stream0 = sat0 + np.random.default_rng(5).normal(0, 0.01, (200, 6))
    ts = (0.0, 8.0, 81)
    tol = 1e-10
    cfg = dft.Config(engine=dft.Engine.GPU, method=dft.Method.DOP853, ts=ts, tolerance=tol,
                     forces="model")
    bg = dft.background(mw)
    stream = dft.test_particles(stream0)
    sat = dft.particles(dft.Potential.plummer(2e-3, 0.02), sat0)
    cfg.add(stream, bg, sat)
    cfg.add(sat, bg)
    print("model API: satellite + 200-particle stream", flush=True)
    got_stream, none, got_sat = cfg.run()  # first-registration order
    assert none is None
    # the same chain on the oracle
    times = cfg.grid.times()
    lt = math.log(tol)
    mw_spec = mw.spec()
    sat_cpu, _ = oracle_lib.batch(_native.METHOD_DOP853, sat0, times, lt, lt, mw_spec)
    acc = np.array([[oracle_lib.force_at(mw_spec, times[k], *sat_cpu[k, 0, :3])]
                    for k in range(times.size)])
    track = Origins.hermite(sat_cpu, acc, times, (2e-3, 0.02))
    stream_cpu, _ = oracle_lib.batch(_native.METHOD_DOP853, stream0, times, lt, lt,
                                     mw_spec.with_origins(track), nthreads=8)
    ep, ev = rel_err(got_sat, sat_cpu)
    print(f"  satellite pos_rel={ep:.3e} vel_rel={ev:.3e}")
    assert ep < REL_TOL and ev < REL_TOL
    ep, ev = rel_err(got_stream, stream_cpu)
    print(f"  stream pos_rel={ep:.3e} vel_rel={ev:.3e}")
    assert ep < 1e-7 and ev < 1e-7
    # the satellite's pull is visible against the stream integrated without it
    alone, _ = oracle_lib.batch(_native.METHOD_DOP853, stream0, times, lt, lt, mw_spec, nthreads=8)
    assert np.abs(alone - stream_cpu).max() > 1e-5


def test_model_api_chained_groups(gpu):
    """Force sources chained through two particle groups: satellite A in the
    Milky Way; group B in the Milky Way plus A; a stream in the Milky Way plus
    B.  Registration order differs from integration order.  Against the same
    chain on the oracle (tracks built with the same Hermite builder)."""
    mw = dft.Potential.bovy()
    a0 = np.array([[1.0, 0.0, 0.05, 0.0, 1.0, 0.05]])
    b0 = np.array([[1.2, 0.1, 0.0, -0.08, 0.95, 0.0], [1.15, -0.1, 0.02, 0.05, 0.97, 0.0]])
    # the stream starts 0.05 from B's first particle: an orbit that passes
    # through a point mass would make the oracle, which like the reference has
    # no step limit, crawl at the step floor
    stream0 = b0[0] + np.array([0.05, 0.0, 0.0, 0.0, 0.0, 0.0]) + \
        np.random.default_rng(9).normal(0, 0.01, (96, 6))
    ts = (0.0, 6.0, 61)
    tol = 1e-10
    cfg = dft.Config(engine=dft.Engine.GPU, method=dft.Method.DOP853, ts=ts, tolerance=tol,
                     forces="model")
    bg = dft.background(mw)
    sat_a = dft.particles(dft.Potential.plummer(3e-3, 0.02), a0)
    grp_b = dft.particles(dft.Potential.plummer(1e-3, 0.01), b0)
    stream = dft.test_particles(stream0)
    cfg.add(stream, bg, grp_b)
    cfg.add(grp_b, bg, sat_a)
    cfg.add(sat_a, bg)
    print("model API: satellite -> group -> stream", flush=True)
    got_stream, none, got_b, got_a = cfg.run()
    assert none is None
    times = cfg.grid.times()
    lt = math.log(tol)
    mws = mw.spec()

    def tracks(states, spec, params):
        acc = np.array([[oracle_lib.force_at(spec, times[k], *states[k, i, :3])
                         for i in range(states.shape[1])] for k in range(times.size)])
        return Origins.hermite(states, acc, times, params)

    a_cpu, _ = oracle_lib.batch(_native.METHOD_DOP853, a0, times, lt, lt, mws)
    b_spec = mws.with_origins(tracks(a_cpu, mws, (3e-3, 0.02)))
    b_cpu, _ = oracle_lib.batch(_native.METHOD_DOP853, b0, times, lt, lt, b_spec)
    s_spec = mws.with_origins(tracks(b_cpu, b_spec, (1e-3, 0.01)))
    s_cpu, _ = oracle_lib.batch(_native.METHOD_DOP853, stream0, times, lt, lt, s_spec, nthreads=8)
    for label, got, want, bar in (("A", got_a, a_cpu, 1e-8), ("B", got_b, b_cpu, 1e-7),
                                  ("stream", got_stream, s_cpu, 1e-6)):
        ep, ev = rel_err(got, want)
        print(f"  {label}: pos_rel={ep:.3e} vel_rel={ev:.3e}")
        assert ep < bar and ev < bar
