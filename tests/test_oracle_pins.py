"""Pins the CPU oracle before the GPU results are compared with it.

Each check is chosen so a plausible restatement mistake (dropped term, wrong
sign or index, transposed coefficient) fails it:

* DOPR54: the reference's known-answer tail bits
  (tests/galpy_integrator_tests.rs:35-42, inputs per
  scripts/generate_galpy_fixtures.py:347-373 and tests/galpy_support/mod.rs).
* DOP853 (no committed vector in the reference): the analytic two-body
  solution via Kepler's equation, dense output against step-end states of a
  second run, order conditions of the tableau.
* The time grid: OutputGrid::sample endpoints (shared/src/config.rs:69-85).
"""

from __future__ import annotations

import math
import re
import struct
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest

import oracle_lib

REPO = Path(__file__).resolve().parent.parent

KAT_TOL = struct.unpack("<d", struct.pack("<Q", 0xC03BA18A998FFFA1))[0]
KAT_TAIL = [0xBFEAD9AC890CBF36, 0xBFE1689EF5F2F535, 0, 0x3FE1689EF5F2DA37, 0xBFEAD9AC890CA90A, 0]


def bits(x: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", float(x)))[0]


def test_kat_tolerance_is_the_fixture_round_trip_of_ln_1e12():
    # generate_galpy_fixtures.py writes f"{np.log(1e-12):.16g}"
    assert float(f"{math.log(1e-12):.16g}") == KAT_TOL
    assert bits(math.log(1e-12)) == 0xC03BA18A998FFFA0


def test_dopr54_known_answer_tail_bits():
    t = np.linspace(0.0, 10.0, 1001)
    assert bits(t[1]) == 0x3F847AE147AE147B
    res, st, err = oracle_lib.orbit(0, [1, 0, 0, 0, 1, 0], t, KAT_TOL, KAT_TOL)
    assert err == 0
    assert [bits(v) for v in res[-1]] == KAT_TAIL
    # 3.3e-12 from the analytic circular orbit x = cos t, y = sin t
    exact = np.array([math.cos(10.0), math.sin(10.0), 0.0, -math.sin(10.0), math.cos(10.0), 0.0])
    assert np.max(np.abs(res[-1] - exact)) < 5e-12


def kepler_solution(y0: np.ndarray, t: float, mu: float = 1.0) -> np.ndarray:
    """Two-body state at time t via universal Kepler's equation (elliptic)."""
    r0v, v0v = y0[:3], y0[3:]
    r0 = np.linalg.norm(r0v)
    v0 = np.linalg.norm(v0v)
    a = 1.0 / (2.0 / r0 - v0 * v0 / mu)
    n = math.sqrt(mu / a**3)
    sigma0 = float(np.dot(r0v, v0v)) / math.sqrt(mu)
    # solve for the eccentric-anomaly difference dE: n t = dE + sigma0/sqrt(a) (1 - cos dE)
    #                                                    - (1 - r0/a) sin dE
    M = n * t
    dE = M
    for _ in range(100):
        f = dE + sigma0 / math.sqrt(a) * (1 - math.cos(dE)) - (1 - r0 / a) * math.sin(dE) - M
        fp = 1 + sigma0 / math.sqrt(a) * math.sin(dE) - (1 - r0 / a) * math.cos(dE)
        step = f / fp
        dE -= step
        if abs(step) < 1e-16:
            break
    F = 1 - a / r0 * (1 - math.cos(dE))
    G = t + math.sqrt(a**3 / mu) * (math.sin(dE) - dE)
    rv = F * r0v + G * v0v
    r = np.linalg.norm(rv)
    Fd = -math.sqrt(mu * a) / (r * r0) * math.sin(dE)
    Gd = 1 - a / r * (1 - math.cos(dE))
    return np.concatenate([rv, Fd * r0v + Gd * v0v])


@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("ecc_case", [0, 1, 2])
def test_eccentric_kepler_orbits_match_the_closed_form(method, ecc_case):
    y0 = [np.array([1.0, 0.0, 0.0, 0.0, 1.2, 0.1]),
          np.array([0.7, 0.3, 0.1, -0.4, 1.1, 0.0]),
          np.array([1.5, -0.2, 0.3, 0.1, 0.5, 0.2])][ecc_case]
    t = np.linspace(0.0, 12.0, 61)
    tol = math.log(1e-12)
    res, st, err = oracle_lib.orbit(method, y0, t, tol, tol)
    assert err == 0
    exact = np.array([kepler_solution(y0, ti) for ti in t])
    # DOPR54 samples at the accumulated time, DOP853 by dense output
    assert np.max(np.abs(res - exact)) < 2e-9, np.max(np.abs(res - exact))
    assert st.accepted > 0


def test_dop853_dense_output_matches_a_run_ending_at_the_sample():
    """Dense output at t_k equals the state of a run whose own grid ends at
    t_k, to the integrator's accuracy (a dropped dense term would not)."""
    y0 = np.array([0.8, 0.1, 0.05, -0.2, 1.1, 0.05])
    t = np.linspace(0.0, 7.3, 74)
    tol = math.log(1e-13)
    res, st, _ = oracle_lib.orbit(1, y0, t, tol, tol)
    assert st.dense_steps == st.accepted
    for k in (5, 23, 50, 73):
        exact = kepler_solution(y0, t[k])
        assert np.max(np.abs(res[k] - exact)) < 1e-10


def _constants(path: Path, namespace: str) -> dict[str, float]:
    text = path.read_text()
    start = text.index(f"namespace {namespace} {{")
    text = text[start:text.index("}  // namespace", start)]
    out = {}
    for name, value in re.findall(r"\b([A-Za-z][A-Za-z0-9_]*)\s*=\s*(-?[0-9.]+e?-?[0-9]*)\s*[;,]",
                                  text):
        out[name.lower()] = float(value)
    return out


def test_dop853_tableau_order_conditions_and_product_copy_agree():
    ora = _constants(REPO / "oracle" / "drift_oracle.hpp", "dop853c")
    prod = _constants(REPO / "astrodrift_amd" / "csrc" / "integrators.hpp", "hw853")
    names = [k for k in ora if re.fullmatch(r"(c|a|b|bhh|d|er)\d+", k)]
    assert len(names) > 120
    for k in names:
        assert prod[k] == ora[k], k  # the product restates the same published table
    c = {i: ora.get(f"c{i}", 0.0) for i in range(1, 17)}
    c[12] = 1.0  # stage 12 is evaluated at t + h (dop853.rs:575-583)
    rows = {}
    for k, v in ora.items():
        m = re.fullmatch(r"a(\d+)", k)
        if m:
            s = m.group(1)
            # a<i><j> with i in 2..16, j < i; two- or three-digit splits
            for split in range(1, len(s)):
                i, j = int(s[:split]), int(s[split:])
                if 2 <= i <= 16 and 1 <= j < i and (s[:split] != "1" or len(s) == 2):
                    rows.setdefault(i, {})[j] = v
                    break
    for i, row in rows.items():
        assert abs(sum(row.values()) - c[i]) < 1e-14, (i, sum(row.values()), c[i])
    b = {1: ora["b1"], 6: ora["b6"], 7: ora["b7"], 8: ora["b8"], 9: ora["b9"], 10: ora["b10"],
         11: ora["b11"], 12: ora["b12"]}
    cc = dict(c)
    cc[1], cc[12] = 0.0, 1.0
    for q in range(1, 9):  # order conditions sum b_i c_i^(q-1) = 1/q up to order 8
        assert abs(sum(b[i] * cc[i] ** (q - 1) for i in b) - 1.0 / q) < 1e-13, q
    er = {i: ora[f"er{i}"] for i in (1, 6, 7, 8, 9, 10, 11, 12)}
    assert abs(sum(er.values())) < 1e-13  # error weights annihilate constants


def test_dopr54_coefficients_consistent():
    b = [Fraction(35, 384), 0, Fraction(500, 1113), Fraction(125, 192), Fraction(-2187, 6784),
         Fraction(11, 84)]
    assert sum(b) == 1
    bstar = [Fraction(5179, 57600), 0, Fraction(7571, 16695), Fraction(393, 640),
             Fraction(-92097, 339200), Fraction(187, 2100), Fraction(1, 40)]
    be = [b[i] - bstar[i] for i in range(6)] + [-bstar[6]]
    assert sum(be) == 0


@pytest.mark.parametrize(("start", "end", "expected"), [(0.0, 1.0, [0.0, 0.5, 1.0]),
                                                         (1.0, -1.0, [1.0, 0.0, -1.0])])
def test_output_grid_endpoints(start, end, expected):
    assert list(oracle_lib.sample_times(start, end, 3)) == expected


def test_output_grid_differs_from_linspace_but_hits_endpoints():
    g = oracle_lib.sample_times(0.0, 10.0, 1001)
    lin = np.linspace(0.0, 10.0, 1001)
    assert g[0] == 0.0 and g[-1] == 10.0
    assert np.count_nonzero(g != lin) > 0
    np.testing.assert_allclose(g, lin, rtol=0, atol=1e-14)
