"""Ragged batches: jobs of different particle counts, potentials, methods,
time grids and tolerances in one persistent launch per method.

Each job's result is compared with the oracle on the same inputs (1e-8
relative, the parity bar) and with the same job run alone through the plain
kernel.  The ragged kernel dispatches the force at run time and is compiled
apart from the plain one, so the two round a few operations differently; an
orbit can then take an accept/reject decision the other way (a step count one
apart and a difference at the tolerance's level), which the comparisons with
the plain kernel allow for 0.2% of the orbits.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from astrodrift_amd import _native, workloads

import oracle_lib

pytestmark = pytest.mark.gpu

POTS = [_native.PotentialSpec.kepler(1.0), _native.PotentialSpec.plummer(1.0, 0.1),
        _native.PotentialSpec.milky_way()]


def disk_ics(spec, n, rng):
    """Near-circular disk orbits in `spec` (circular speed from the oracle force)."""
    R = rng.uniform(0.3, 2.0, n)
    phi = rng.uniform(0, 2 * math.pi, n)
    z = rng.normal(0, 0.03, n)
    vc = np.array([math.sqrt(-r * oracle_lib.force(spec, r, 0.0, 0.0)[0]) for r in R])
    vphi = vc * (1 + rng.normal(0, 0.05, n))
    vR, vz = rng.normal(0, 0.05, n), rng.normal(0, 0.05, n)
    c, s = np.cos(phi), np.sin(phi)
    return np.ascontiguousarray(np.stack([R * c, R * s, z, vR * c - vphi * s, vR * s + vphi * c,
                                          vz], 1))


def make_jobs(count, seed, span=(5.0, 20.0)):
    rng = np.random.default_rng(seed)
    jobs = []
    for _ in range(count):
        n = int(np.exp(rng.uniform(math.log(16), math.log(512))))
        pot = int(rng.integers(0, 3))
        method = int(rng.integers(0, 2))
        t_end = rng.uniform(*span)
        nt = int(rng.integers(11, 102))
        tol = math.exp(rng.uniform(math.log(1e-12), math.log(1e-8)))
        jobs.append(dict(n=n, pot=pot, method=method, times=workloads.grid_times(0.0, t_end, nt),
                         tol=tol, y0=disk_ics(POTS[pot], n, rng)))
    return jobs


def run_ragged(jobs, torch):
    dev = torch.device("cuda:0")
    keep, descs = [], []
    for j in jobs:
        s0 = torch.from_numpy(j["y0"]).to(dev)
        tt = torch.from_numpy(j["times"]).to(dev)
        out = torch.full((len(j["times"]), j["n"], 6), float("nan"), dtype=torch.float64,
                         device=dev)
        st = torch.zeros((j["n"], 4), dtype=torch.int32, device=dev)
        keep.append((s0, tt, out, st))
        lt = math.log(j["tol"])
        descs.append(_native.DeviceJob(j["method"], j["pot"], j["n"], len(j["times"]), lt, lt,
                                       s0.data_ptr(), tt.data_ptr(), out.data_ptr(),
                                       st.data_ptr(), span_hint=float(j["times"][-1])))
    _native.ragged_device(descs, POTS)
    torch.cuda.synchronize()
    return [(o.cpu().numpy(), s.cpu().numpy()) for (_, _, o, s) in keep]


def rel(a, b):
    dp = np.linalg.norm(a[..., :3] - b[..., :3], axis=-1) / np.linalg.norm(b[..., :3], axis=-1)
    dv = np.linalg.norm(a[..., 3:] - b[..., 3:], axis=-1) / np.linalg.norm(b[..., 3:], axis=-1)
    return max(dp.max(), dv.max())


def test_ragged_batch_matches_oracle_and_plain_kernel(gpu):
    torch = pytest.importorskip("torch")
    jobs = make_jobs(24, seed=21)
    results = run_ragged(jobs, torch)
    worst_oracle = worst_plain = 0.0
    orbits = differing = 0
    for j, (out, st) in zip(jobs, results):
        lt = math.log(j["tol"])
        ref, rst = oracle_lib.batch(j["method"], j["y0"], j["times"], lt, lt, POTS[j["pot"]],
                                    nthreads=8)
        plain, pst, _ = _native.launch_host(j["method"], j["y0"], j["times"], lt, lt, POTS[j["pot"]])
        assert not np.isnan(out).any()
        np.testing.assert_array_equal(out[0], j["y0"])
        worst_oracle = max(worst_oracle, rel(out, ref))
        worst_plain = max(worst_plain, rel(out, plain))
        assert abs(int(st[:, 0].sum()) - int(rst["accepted"].sum())) <= 1e-3 * rst["accepted"].sum()
        same = st[:, 0] == pst["accepted"].astype(np.int32)
        orbits += same.size
        differing += int((~same).sum())
        if not same.all():
            k = np.flatnonzero(~same)
            print(f"  job pot={j['pot']} method={j['method']} tol={j['tol']:.1e}: accepted differs "
                  f"for {k.size} orbits, e.g. {int(st[k[0], 0])} vs {int(pst['accepted'][k[0]])}; "
                  f"their rel diff {rel(out[:, k], plain[:, k]):.3e}")
    print(f"ragged: worst rel vs oracle {worst_oracle:.3e}, vs plain kernel {worst_plain:.3e}; "
          f"{differing} of {orbits} orbits with different accepted counts")
    assert worst_oracle < 1e-8
    assert worst_plain < 1e-8
    assert differing <= 0.002 * orbits


def test_ragged_empty_and_single_jobs(gpu):
    torch = pytest.importorskip("torch")
    jobs = make_jobs(1, seed=5)
    (out, st), = run_ragged(jobs, torch)
    lt = math.log(jobs[0]["tol"])
    ref, _ = oracle_lib.batch(jobs[0]["method"], jobs[0]["y0"], jobs[0]["times"], lt, lt,
                              POTS[jobs[0]["pot"]])
    assert rel(out, ref) < 1e-8
    _native.ragged_device([], POTS)  # no jobs: a no-op


def test_ragged_lds_and_global_mw_tables_fp64_and_mixed(gpu):
    """Two different Milky-Way potentials in one launch: the one with more
    orbits has its table staged in LDS, the other is read from global memory.
    Each job against the plain kernel at the same precision (the same force
    code): FP64 and the mixed mode (FP32 stage forces, float tables), and
    against the FP64 oracle within each mode's error budget."""
    torch = pytest.importorskip("torch")
    pots = [_native.PotentialSpec.milky_way(),
            _native.PotentialSpec.milky_way(bulge_alpha=1.0, disk_frac=0.55, halo_frac=0.4)]
    rng = np.random.default_rng(33)
    dev = torch.device("cuda:0")
    for precision, tol, bar in ((_native.PRECISION_FP64, 1e-10, 1e-8),
                                (_native.PRECISION_MIXED, 1e-7, 1e-3)):
        jobs = []
        for k in range(8):
            pot = 0 if k < 5 else 1  # potential 0 has more orbits: its table goes to LDS
            n = int(rng.integers(64, 256))
            jobs.append(dict(n=n, pot=pot, method=k % 2, times=workloads.grid_times(0.0, 10.0, 41),
                             y0=disk_ics(pots[pot], n, rng)))
        keep, descs = [], []
        lt = math.log(tol)
        for j in jobs:
            s0 = torch.from_numpy(j["y0"]).to(dev)
            tt = torch.from_numpy(j["times"]).to(dev)
            out = torch.full((41, j["n"], 6), float("nan"), dtype=torch.float64, device=dev)
            st = torch.zeros((j["n"], 4), dtype=torch.int32, device=dev)
            keep.append((s0, tt, out, st))
            descs.append(_native.DeviceJob(j["method"], j["pot"], j["n"], 41, lt, lt, s0.data_ptr(),
                                           tt.data_ptr(), out.data_ptr(), st.data_ptr()))
        print(f"ragged two MW potentials, precision={precision}", flush=True)
        _native.ragged_device(descs, pots, precision=precision)
        torch.cuda.synchronize()
        worst = worst_plain = 0.0
        for j, (s0, tt, out, st) in zip(jobs, keep):
            o = out.cpu().numpy()
            assert not np.isnan(o).any()
            ref, _ = oracle_lib.batch(j["method"], j["y0"], j["times"], lt, lt, pots[j["pot"]],
                                      nthreads=8)
            worst = max(worst, rel(o, ref))
            plain = torch.empty_like(out)
            _native.launch_device(j["method"], s0.data_ptr(), tt.data_ptr(), plain.data_ptr(),
                                  j["n"], 41, lt, lt, pots[j["pot"]], precision=precision,
                                  schedule=_native.SCHEDULE_PLAIN)
            torch.cuda.synchronize()
            worst_plain = max(worst_plain, rel(o, plain.cpu().numpy()))
        print(f"  worst rel vs oracle {worst:.3e}, vs plain kernel {worst_plain:.3e}")
        assert worst < bar
        assert worst_plain < 1e-8


def test_large_ragged_batch_matches_jobs_alone_and_plain_kernel(gpu):
    """A batch of four jobs of 9-10 thousand orbits: each job's result is
    bit-identical to the same job run alone (an orbit's arithmetic does not
    depend on what else the launch carries) and agrees with the plain kernel
    to the parity bar, with the same accepted counts on all but 0.2% of the
    orbits."""
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(77)
    jobs = []
    for k, (pot, method) in enumerate([(2, 1), (0, 0), (1, 1), (2, 0)]):
        n = int(rng.integers(9000, 10000))
        jobs.append(dict(n=n, pot=pot, method=method, times=workloads.grid_times(0.0, 8.0, 21),
                         tol=1e-9, y0=disk_ics(POTS[pot], n, rng)))
    print("large ragged batch:", [(j["n"], j["pot"], j["method"]) for j in jobs], flush=True)
    results = run_ragged(jobs, torch)
    worst = 0.0
    orbits = differing = 0
    for j, (out, st) in zip(jobs, results):
        (alone, ast), = run_ragged([j], torch)
        np.testing.assert_array_equal(out, alone)
        np.testing.assert_array_equal(st, ast)
        lt = math.log(j["tol"])
        plain, pst, _ = _native.launch_host(j["method"], j["y0"], j["times"], lt, lt, POTS[j["pot"]])
        worst = max(worst, rel(out, plain))
        orbits += j["n"]
        differing += int(np.count_nonzero(st[:, 0] != pst["accepted"].astype(np.int32)))
    print(f"  batch == jobs alone; worst rel vs plain kernel {worst:.3e}; {differing} of "
          f"{orbits} orbits with different accepted counts")
    assert worst < 1e-8
    assert differing <= 0.002 * orbits


def test_ragged_autotuner_choice_keeps_every_bit(gpu):
    """drift_tune_ragged picks a block size per method from {64, 128, 256};
    launches with block_size 0 then use it, and the block size changes no
    result: every output and statistic equals the 256-thread launch's."""
    torch = pytest.importorskip("torch")
    jobs = make_jobs(16, seed=33)
    assert {j["method"] for j in jobs} == {0, 1}
    dev = torch.device("cuda:0")
    keep, descs = [], []
    for j in jobs:
        s0 = torch.from_numpy(j["y0"]).to(dev)
        tt = torch.from_numpy(j["times"]).to(dev)
        out = torch.full((len(j["times"]), j["n"], 6), float("nan"), dtype=torch.float64,
                         device=dev)
        st = torch.zeros((j["n"], 4), dtype=torch.int32, device=dev)
        keep.append((s0, tt, out, st))
        lt = math.log(j["tol"])
        descs.append(_native.DeviceJob(j["method"], j["pot"], j["n"], len(j["times"]), lt, lt,
                                       s0.data_ptr(), tt.data_ptr(), out.data_ptr(),
                                       st.data_ptr(), span_hint=float(j["times"][-1])))
    _native.ragged_device(descs, POTS, block_size=256)
    torch.cuda.synchronize()
    ref = [(o.cpu().numpy(), s.cpu().numpy()) for (_, _, o, s) in keep]
    best = _native.tune_ragged(descs, POTS)
    print("tuned:", best)
    assert best["dopr54_block"] in (64, 128, 256) and best["dop853_block"] in (64, 128, 256)
    for (_, _, o, s) in keep:
        o.fill_(float("nan"))
        s.zero_()
    _native.ragged_device(descs, POTS)
    torch.cuda.synchronize()
    for (o, s), (ro, rs) in zip([(o.cpu().numpy(), s.cpu().numpy()) for (_, _, o, s) in keep],
                                ref):
        assert np.array_equal(o.view(np.uint64), ro.view(np.uint64))
        assert np.array_equal(s, rs)
    with pytest.raises(_native.DriftNativeError, match="no jobs"):
        _native.tune_ragged([], POTS)


def test_ragged_mix_draw_matches_oracle(gpu):
    """The first 64 jobs of the seeded ragged mix (workloads.ragged_mix, its
    real distribution): 16 to 61,185 orbits per job, spans up to 198 time
    units, up to 496 outputs, rtol 1e-12 to 1e-8, all three potentials and
    both methods, 468,039 orbits in one call.  Five orbits of every job,
    strided over the job, against the oracle on the same inputs, the
    reference's error-summary style of check (tests/galpy_integrator_tests.rs:
    174-236): the parity bar of 1e-8 relative per sample, except for an orbit
    whose trajectory the problem itself does not fix to that accuracy.  The
    oracle run from the same start with x0 moved by one ulp measures that:
    over 143 time units a chaotic Milky-Way orbit moves 2e-7 that way, and
    any change of rounding (the kernels contract multiply-adds and use the
    second-order form) moves it as much.  Such an orbit must agree within 10x
    its one-ulp sensitivity, and at most 3% of the checked orbits may need
    that allowance.  Step counts are equal on all but 3% of the checked
    orbits (a step decision within a few ulps of its threshold, or on a
    chaotic orbit, may fall the other way)."""
    torch = pytest.importorskip("torch")
    jobs = workloads.ragged_mix(64)
    assert max(j.n for j in jobs) >= 32768
    assert max(float(j.times[-1]) for j in jobs) >= 150.0
    assert max(j.nt for j in jobs) >= 400
    dev = torch.device("cuda:0")
    keep, descs = [], []
    for j in jobs:
        s0 = torch.from_numpy(j.state0).to(dev)
        tt = torch.from_numpy(j.times).to(dev)
        out = torch.full((j.nt, j.n, 6), float("nan"), dtype=torch.float64, device=dev)
        st = torch.zeros((j.n, 4), dtype=torch.int32, device=dev)
        keep.append((s0, tt, out, st))  # the inputs must outlive the launch
        lt = math.log(j.rtol)
        method = _native.METHOD_DOP853 if j.method == "DOP853" else _native.METHOD_DOPR54
        descs.append(_native.DeviceJob(method, j.potential, j.n, j.nt, lt, lt, s0.data_ptr(),
                                       tt.data_ptr(), out.data_ptr(), st.data_ptr(),
                                       span_hint=float(j.times[-1])))
    print(f"ragged mix draw: {len(jobs)} jobs, {sum(j.n for j in jobs)} orbits", flush=True)
    _native.ragged_device(descs, POTS)
    torch.cuda.synchronize()
    worst, checked, flips, chaotic = 0.0, 0, 0, 0
    for k, (j, (_, _, out, st)) in enumerate(zip(jobs, keep)):
        idx = np.unique(np.linspace(0, j.n - 1, 5).astype(np.int64))
        got = out[:, torch.from_numpy(idx).to(dev)].cpu().numpy()
        gst = st[torch.from_numpy(idx).to(dev)].cpu().numpy()
        lt = math.log(j.rtol)
        method = _native.METHOD_DOP853 if j.method == "DOP853" else _native.METHOD_DOPR54
        ref, rst = oracle_lib.batch(method, np.ascontiguousarray(j.state0[idx]), j.times, lt, lt,
                                    POTS[j.potential], nthreads=8)
        assert not np.isnan(got).any(), k
        np.testing.assert_array_equal(got[0], j.state0[idx])
        checked += idx.size
        flips += int(np.count_nonzero((gst[:, 0] != rst["accepted"].astype(np.int32)) |
                                      (gst[:, 1] != rst["rejected"].astype(np.int32))))
        for q in range(idx.size):
            err = rel(got[:, q], ref[:, q])
            if err < 1e-8:
                worst = max(worst, err)
                continue
            # the problem's own sensitivity: the oracle from x0 moved one ulp
            y1 = np.ascontiguousarray(j.state0[idx[q]:idx[q] + 1]).copy()
            y1[0, 0] = np.nextafter(y1[0, 0], np.inf)
            per, _ = oracle_lib.batch(method, y1, j.times, lt, lt, POTS[j.potential])
            ulp = rel(per[:, 0], ref[:, q])
            print(f"  job {k} orbit {int(idx[q])}: {j.method} pot {j.potential} "
                  f"span={j.times[-1]:.0f} rtol={j.rtol:.1e}: rel {err:.3e}, the oracle's "
                  f"one-ulp sensitivity {ulp:.3e}", flush=True)
            assert err < 10.0 * ulp, (k, int(idx[q]), err, ulp)
            chaotic += 1
    print(f"ragged mix draw: worst rel vs oracle {worst:.3e} over {checked - chaotic} of "
          f"{checked} orbits of {len(jobs)} jobs; {chaotic} chaotic orbits within 10x their "
          f"one-ulp sensitivity; {flips} with different step counts", flush=True)
    assert worst < 1e-8
    assert chaotic <= 0.03 * checked
    assert flips <= 0.03 * checked


def test_job_filling_the_device_is_cost_ordered_without_changing_bits(gpu):
    """A job of 160,000 orbits fills the device's lanes (2 waves of 64 per
    SIMD on 256 CUs: 131,072), so its work items run in cost order; the same
    orbits as two jobs of 80,000, which stay in index order, give the same
    trajectories and statistics bit for bit, for both methods."""
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(91)
    n = 160000
    y0 = disk_ics(POTS[2], n, rng)
    times = workloads.grid_times(0.0, 2.0, 11)
    for method in (_native.METHOD_DOPR54, _native.METHOD_DOP853):
        whole = dict(n=n, pot=2, method=method, times=times, tol=1e-9, y0=y0)
        halves = [dict(n=n // 2, pot=2, method=method, times=times, tol=1e-9,
                       y0=np.ascontiguousarray(y0[k * (n // 2):(k + 1) * (n // 2)]))
                  for k in range(2)]
        (out, st), = run_ragged([whole], torch)
        parts = run_ragged(halves, torch)
        assert not np.isnan(out).any()
        np.testing.assert_array_equal(out, np.concatenate([p[0] for p in parts], axis=1))
        np.testing.assert_array_equal(st, np.concatenate([p[1] for p in parts], axis=0))
