"""Golden fixtures (tests/golden/, written by make_golden.py).

CPU: the fixture format round-trips bit-exactly, the oracle reproduces every
committed trajectory bit for bit (regression pin) and the canonical DOPR54
fixture yields the reference's known-answer tail.  GPU: the kernels, run on
each fixture's inputs through the C ABI, agree with it within 1e-8 relative
(the reference's GPU error-summary mode, tests/galpy_integrator_tests.rs:174-236).
"""

from __future__ import annotations

import math
from pathlib import Path

import numpy as np
import pytest

from astrodrift_amd import _native
from astrodrift_amd.fixtures import GalpyFixture, error_summary, run_fixture_gpu, ulp_distance

import oracle_lib

GOLDEN = Path(__file__).resolve().parent / "golden"
FULL = sorted(p for p in GOLDEN.glob("*.fixture") if p.name != "dopr54_reference.fixture")


def method_of(fx: GalpyFixture, path: Path) -> int:
    return 0 if "dopr54" in path.name else 1


def potential_of(fx: GalpyFixture):
    for c in fx.comments:
        parts = c.split()
        if len(parts) >= 3 and parts[1] == "potential":
            if parts[2] == "mw":
                return _native.PotentialSpec.milky_way()
            if parts[2] == "plummer":
                return _native.PotentialSpec.plummer(float(parts[3]), float(parts[4]))
    return None


def test_reference_fixture_round_trip_and_kat():
    fx = GalpyFixture.read(GOLDEN / "dopr54_reference.fixture")
    assert fx.expected_state_bits.size == 0  # input-only, like the reference's
    assert GalpyFixture.parse(fx.format()).format() == fx.format()
    tail = [int(h, 16) for h in (GOLDEN / "dopr54_reference_tail.txt").read_text().split()]
    res, _, err = oracle_lib.orbit(0, fx.initial_state, fx.times, fx.rtol, fx.atol,
                                   dt_one=fx.dt_one)
    assert err == 0
    assert list(res[-1].view(np.uint64)) == tail


@pytest.mark.parametrize("path", FULL, ids=[p.stem for p in FULL])
def test_oracle_reproduces_golden_bits(path):
    fx = GalpyFixture.read(path)
    assert GalpyFixture.parse(fx.format()).format() == fx.format()
    res, _, err = oracle_lib.orbit(method_of(fx, path), fx.initial_state, fx.times, fx.rtol,
                                   fx.atol, potential_of(fx), dt_one=fx.dt_one)
    assert err == 0
    assert np.array_equal(res.view(np.uint64).reshape(-1), fx.expected_state_bits)


def test_fixture_validation_messages():
    fx = GalpyFixture.read(GOLDEN / "dopr54_reference.fixture")
    with pytest.raises(ValueError, match="output times"):
        GalpyFixture.parse(fx.format().replace("nt 1001", "nt 1000"))
    with pytest.raises(ValueError, match="missing"):
        GalpyFixture.parse("dim 6\n")


def test_ulp_metric():
    a = np.array([1.0, -1.0, 0.0, 1.0])
    b = np.array([np.nextafter(1.0, 2.0), np.nextafter(-1.0, -2.0), -0.0, 1.0])
    assert list(ulp_distance(a, b)) == [1, 1, 1, 0]
    s = error_summary(a, b)
    assert s["mismatched"] == 3 and s["max_ulp"] == 1


@pytest.mark.gpu
@pytest.mark.parametrize("path", FULL, ids=[p.stem for p in FULL])
def test_gpu_matches_golden(gpu, path):
    fx = GalpyFixture.read(path)
    got, stats, code = run_fixture_gpu(method_of(fx, path), fx, potential_of(fx))
    assert code == 0
    want = fx.expected_states()
    summary = error_summary(got, want)
    print(path.stem, summary, "accepted", stats["accepted"])
    dp = np.linalg.norm(got[:, :3] - want[:, :3], axis=1) / np.linalg.norm(want[:, :3], axis=1)
    dv = np.linalg.norm(got[:, 3:] - want[:, 3:], axis=1) / np.linalg.norm(want[:, 3:], axis=1)
    assert dp.max() < 1e-8 and dv.max() < 1e-8
    np.testing.assert_array_equal(got[0], want[0])
    if path.stem == "dop853_reference":
        # the canonical DOP853 dump: the reference's own GPU bar
        # (tests/galpy_integrator_tests.rs:327-336)
        assert summary["max_abs"] < 1e-12, summary


@pytest.mark.gpu
def test_gpu_reference_kat_fixture(gpu):
    fx = GalpyFixture.read(GOLDEN / "dopr54_reference.fixture")
    got, _, code = run_fixture_gpu(0, fx)
    tail = np.array([int(h, 16) for h in (GOLDEN / "dopr54_reference_tail.txt").read_text().split()],
                    dtype=np.uint64).view(np.float64)
    assert code == 0
    print("tail ULPs", ulp_distance(got[-1], tail))
    # measured 1.7e-14: the second-order restatement and FMA contraction move
    # the last bits (bitwise equality is the oracle's bar, test above)
    assert np.max(np.abs(got[-1] - tail)) < 1e-13
