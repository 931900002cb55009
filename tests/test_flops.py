"""The algorithmic FLOP constants in astrodrift_amd/flops.py against the
operation-counting run of the oracle (the reference arithmetic)."""

from __future__ import annotations

import math

import numpy as np
import pytest

from astrodrift_amd import _native, flops, workloads

import oracle_lib

SPECS = {
    "galpy": None,
    "kepler": _native.PotentialSpec.kepler(1.0),
    "plummer": _native.PotentialSpec.plummer(1.0, 0.1),
    "mw": _native.PotentialSpec.milky_way(),
}


def _stats_array(st):
    a = np.zeros(1, dtype=_native.STATS_DTYPE)
    a["accepted"], a["rejected"], a["dense_steps"] = st.accepted, st.rejected, st.dense_steps
    return a


@pytest.mark.parametrize("pot", list(SPECS))
def test_rhs_constants(pot):
    w = workloads.mw_disk(n=4, seed=1, t_end=5.0, steps=11)
    ops, st = oracle_lib.count_ops(1, w.state0[0], w.times, math.log(1e-10), math.log(1e-10),
                                   SPECS[pot])
    assert sum(ops["rhs"].values()) == flops.rhs_flops(SPECS[pot]) * st.rhs_evals


@pytest.mark.parametrize("pot", list(SPECS))
@pytest.mark.parametrize("seed", [1, 2])
def test_dop853_totals(pot, seed):
    w = workloads.mw_disk(n=8, seed=seed, t_end=8.0, steps=81)
    spec = SPECS[pot]
    ops, st = oracle_lib.count_ops(1, w.state0[3], w.times, math.log(1e-11), math.log(1e-11), spec)
    assert st.rhs_evals == 2 + 11 * st.attempts + st.accepted + 3 * st.dense_steps
    assert ops["dense"] and sum(ops["dense"].values()) == (
        flops.DOP853_DENSE * st.dense_steps + flops.rhs_flops(spec) * 0)
    assert sum(ops["sample"].values()) == flops.DOP853_SAMPLE * st.samples
    assert sum(ops["accept"].values()) == flops.DOP853_ACCEPT * st.accepted
    total = sum(sum(v.values()) for v in ops.values())
    est = flops.dop853_flops(_stats_array(st), w.nt, spec)
    assert abs(est - total) <= 2e-3 * total, (est, total)


@pytest.mark.parametrize("pot", ["galpy", "mw"])
def test_dopr54_totals(pot):
    w = workloads.kepler_disk(n=4)
    spec = SPECS[pot]
    ops, st = oracle_lib.count_ops(0, w.state0[1], w.times, math.log(1e-10), math.log(1e-10), spec)
    total = sum(sum(v.values()) for v in ops.values())
    est = flops.dopr54_flops(_stats_array(st), spec)
    # the estimate's iteration count varies per orbit; the bound covers it
    assert abs(est - total) <= 1e-2 * total, (est, total)
    assert ops["attempt"]["trans"] == 25 * st.attempts
