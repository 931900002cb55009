"""Writes the golden fixtures under tests/golden/.

* dopr54_reference.fixture / dopr54_reference_tail.txt: the reference's
  canonical DOPR54 known-answer case, transcribed as data.  Inputs follow
  scripts/generate_galpy_fixtures.py:347-373 (y0 = (1,0,0,0,1,0),
  t = numpy.linspace(0, 10, 1001), rtol = atol = the "%.16g" round trip of
  ln 1e-12, dt_one = -9999.99); the expected final state bits are
  tests/galpy_integrator_tests.rs:35-42.
* case_*.fixture: trajectories of the oracle (the restated reference CPU
  integrators) in the galpy dump format, on initial conditions drawn like the
  reference's seeded corpus (scripts/generate_galpy_fixtures.py:383-401,
  seed 0xA57D0F54; galpy cylindrical (R, vR, vT, z, vz, phi) converted to
  Cartesian), unit galpy Kepler force, rtol = atol = 1e-12.
* mw_dop853_case0.fixture, plummer_dopr54_case0.fixture: the same for the
  Milky-Way and Plummer(1, 0.1) forces (named in a "# potential" comment).

Run from the repository root: python tests/golden/make_golden.py
"""

from __future__ import annotations

import math
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parent.parent
sys.path.insert(0, str(REPO))
sys.path.insert(0, str(REPO / "tests"))

from astrodrift_amd import _native  # noqa: E402
from astrodrift_amd.fixtures import GalpyFixture  # noqa: E402

import oracle_lib  # noqa: E402

SEED = 0xA57D0F54
CORPUS_STEPS = 1000
KAT_TAIL = ["bfead9ac890cbf36", "bfe1689ef5f2f535", "0000000000000000", "3fe1689ef5f2da37",
            "bfead9ac890ca90a", "0000000000000000"]


def seeded_cases(count: int):
    rng = np.random.default_rng(SEED)
    for _ in range(count):
        R = rng.uniform(0.75, 2.25)
        phi = rng.uniform(-np.pi, np.pi)
        z = rng.uniform(-0.35, 0.35)
        vR = rng.uniform(-0.35, 0.35)
        vT = rng.uniform(0.65, 1.25) / np.sqrt(R)
        vz = rng.uniform(-0.25, 0.25)
        end = rng.uniform(0.75, 1.50)
        c, s = math.cos(phi), math.sin(phi)
        y0 = np.array([R * c, R * s, z, vR * c - vT * s, vR * s + vT * c, vz])
        yield (R, vR, vT, z, vz, phi), y0, end, np.linspace(0.0, end, CORPUS_STEPS)


def oracle_fixture(method, y0, times, tol_ln, spec=None, comments=()):
    res, _, err = oracle_lib.orbit(method, y0, times, tol_ln, tol_ln, spec)
    assert err == 0
    return GalpyFixture.from_run(times, y0, res, tol_ln, tol_ln, comments=comments)


def main() -> None:
    kat_tol = float(f"{math.log(1e-12):.16g}")
    t = np.linspace(0.0, 10.0, 1001)
    GalpyFixture(6, 1001, -9999.99, kat_tol, kat_tol, t, np.array([1.0, 0, 0, 0, 1.0, 0]),
                 comments=["# reference canonical DOPR54 input (input-only)"]).write(
        HERE / "dopr54_reference.fixture")
    (HERE / "dopr54_reference_tail.txt").write_text(" ".join(KAT_TAIL) + "\n")

    tol = math.log(1e-12)
    # canonical DOP853 dump: orbit_cyl (1, 0, 1, 0, 0, 0), t = linspace(0, 10, 1001)
    oracle_fixture(1, np.array([1.0, 0, 0, 0, 1.0, 0]), t, tol,
                   comments=["# oracle DOP853, galpy Kepler, orbit_cyl 1 0 1 0 0 0"]).write(
        HERE / "dop853_reference.fixture")
    for case, (cyl, y0, end, times) in enumerate(seeded_cases(2)):
        head = [f"# seed 0x{SEED:x}", f"# case {case}",
                "# orbit_cyl " + " ".join(f"{v:.17g}" for v in cyl), f"# t_end {end:.17g}"]
        for method, name in ((0, "dopr54"), (1, "dop853")):
            oracle_fixture(method, y0, times, tol, comments=[f"# oracle {name.upper()}"] + head
                           ).write(HERE / f"{name}_case_{case:03}.fixture")
    cyl, y0, end, _ = next(seeded_cases(1))
    times = np.linspace(0.0, 5.0, 101)
    oracle_fixture(1, y0, times, tol, _native.PotentialSpec.milky_way(),
                   comments=["# oracle DOP853", "# potential mw"]).write(
        HERE / "mw_dop853_case0.fixture")
    oracle_fixture(0, y0, times, math.log(1e-10), _native.PotentialSpec.plummer(1.0, 0.1),
                   comments=["# oracle DOPR54", "# potential plummer 1.0 0.1"]).write(
        HERE / "plummer_dopr54_case0.fixture")


if __name__ == "__main__":
    main()
