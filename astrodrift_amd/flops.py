"""Algorithmic FP64 operation counts of the integrators.

Counts are source-level operations of the reference algorithms
(src/integrators/galpy/dop853.rs, dopr54.rs and the force formulas of
shared/src/potential.rs): + - * / and sqrt count 1 each, every libm
transcendental call (log, exp, pow, and the incomplete gamma function of the
bulge) counts 1, comparisons / fabs / min / max count 0.  They were measured by
running the reference arithmetic with an operation-counting scalar type;
tests/test_flops.py re-derives them.

Per event:
  DOP853  attempt 979 (12 stages, 8th-order update, blended error norm,
          step control), accepted-step extra 36 (FSAL bookkeeping, rcont1-4),
          dense-output step 888 (rcont5-8 and the three extra stage
          combinations), sample 87 (the 7-level dense polynomial), init 109
  DOPR54  attempt 479 (6 stages, update, log-space error norm: 15 log, 8 exp,
          2 pow), sub-step clamp 4, init 867 (RK4 step-doubling estimate)
  RHS     galpy Kepler 23, Kepler 12, Plummer 11, Milky Way 65
Right-hand-side evaluations: DOP853 2 + 11 per attempt + 1 per accept + 3 per
dense step; DOPR54 7 + 6 per attempt (plus 12 per estimate iteration, counted
in its init).

The kernels evaluate DOP853's three dense stages only on accepted steps that
emit a sample (the reference does so on every accepted step); the
``dense_steps`` counter records how many did, and these counts charge only
that executed work.
"""

from __future__ import annotations

import numpy as np

from ._native import (
    POTENTIAL_GALPY_KEPLER,
    POTENTIAL_KEPLER,
    POTENTIAL_MW,
    POTENTIAL_PLUMMER,
)

DOP853_ATTEMPT = 979
DOP853_ACCEPT = 36
DOP853_DENSE = 888
DOP853_SAMPLE = 87
DOP853_INIT = 109

DOPR54_ATTEMPT = 479
DOPR54_SUBSTEP = 4
DOPR54_INIT = 867

RHS_FLOPS = {
    POTENTIAL_GALPY_KEPLER: 23,
    POTENTIAL_KEPLER: 12,
    POTENTIAL_PLUMMER: 11,
    POTENTIAL_MW: 65,
}


def rhs_flops(potential) -> int:
    return RHS_FLOPS[POTENTIAL_GALPY_KEPLER if potential is None else potential.kind]


def dop853_flops(stats: np.ndarray, nt: int, potential=None) -> float:
    """Algorithmic FLOPs of one DOP853 launch from its per-orbit stats
    (structured drift_orbit_stats array or an (n, 4) integer array)."""
    acc, rej, dense = _fields(stats)
    att = acc + rej
    rhs = 2 * len(acc) + 11 * att.sum() + acc.sum() + 3 * dense.sum()
    return float(
        DOP853_ATTEMPT * att.sum()
        + DOP853_ACCEPT * acc.sum()
        + DOP853_DENSE * dense.sum()
        + DOP853_SAMPLE * (nt - 1) * len(acc)
        + DOP853_INIT * len(acc)
        + rhs_flops(potential) * rhs
    )


def dopr54_flops(stats: np.ndarray, potential=None) -> float:
    acc, rej, _ = _fields(stats)
    att = acc + rej
    rhs = 7 * len(acc) + 6 * att.sum()
    return float(
        (DOPR54_ATTEMPT + DOPR54_SUBSTEP) * att.sum()
        + DOPR54_INIT * len(acc)
        + rhs_flops(potential) * rhs
    )


def _fields(stats):
    if stats.dtype.names:
        return (stats["accepted"].astype(np.int64), stats["rejected"].astype(np.int64),
                stats["dense_steps"].astype(np.int64))
    s = np.asarray(stats).astype(np.int64)
    return s[:, 0], s[:, 1], s[:, 2]
