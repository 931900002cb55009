"""The reference's Python API contract, on CPU (no integration is launched).

Mirrors tests/test_python_{time_grids,errors,selectors,containers,smoke}.py of
the reference where they do not need a GPU.
"""

from __future__ import annotations

import math
import re
import warnings

import numpy as np
import pytest

import astrodrift_amd as dft
from astrodrift_amd import api


@pytest.mark.parametrize(
    "ts",
    (
        (0.0, 1.0, 2),
        (1.0, -1.0, 7),
        np.linspace(0.0, 1.0, 7),
        np.linspace(0.0, 1.0e-12, 7),
        np.linspace(1.0e12, 1.0e12 + 1.0e6, 7),
        np.linspace(1.0, -1.0, 7),
        (0.0, 1.0, 1025),  # above the reference's 1024 limit, lifted here
    ),
)
def test_valid_time_grids_are_accepted(ts):
    dft.Config(ts=ts)


def test_read_only_time_array_is_accepted():
    times = np.linspace(0.0, 1.0, 7)
    times.flags.writeable = False
    dft.Config(ts=times)


@pytest.mark.parametrize(
    ("ts", "message"),
    (
        ([0.0, 1.0e-12, 3.0e-12], "uniformly spaced"),
        ([0.0, 0.25, 0.500_000_000_001, 0.75, 1.0], "uniformly spaced"),
        (np.logspace(0.0, 1.0, 7), "uniformly spaced"),
        ([0.0, 0.5, 0.5, 1.0], "strictly increasing or strictly decreasing"),
        ([0.0, 0.5, 0.25, 1.0], "strictly increasing or strictly decreasing"),
        ((float("nan"), 1.0, 3), "finite"),
        ([0.0, float("inf")], "finite"),
        ((0.0, 1.0, 1), "between 2 and"),
        ([0.0], "between 2 and"),
        ((1.0, 1.0, 2), "non-zero"),
        ([1.0, 1.0], "non-zero"),
        (np.linspace(0.0, 1.0, 7)[::2], "contiguous"),
        ((0.0, 1.0), "3 elements"),
    ),
)
def test_invalid_time_grids_are_rejected(ts, message):
    with pytest.raises(ValueError, match=message):
        dft.Config(ts=ts)


def test_output_grid_samples_match_reference_formula():
    g = api.OutputGrid(0.0, 10.0, 1001)
    t = g.times()
    assert t[0] == 0.0 and t[-1] == 10.0
    assert t[137] == 0.0 + (10.0 - 0.0) * 137.0 / 1000.0
    # differs from numpy.linspace in some entries (the reference reconstructs)
    assert np.count_nonzero(t != np.linspace(0.0, 10.0, 1001)) > 0


def test_exception_hierarchy():
    assert issubclass(dft.IntegrationError, dft.DriftError)
    assert issubclass(dft.DriftError, RuntimeError)


def test_unsupported_configuration_uses_not_implemented_error():
    with pytest.raises(NotImplementedError) as raised:
        dft.Config(engine=dft.Engine.CPU, method=dft.Method.DOPR54,
                   implementation=dft.Implementation.SCIPY)
    assert not isinstance(raised.value, dft.DriftError)


@pytest.mark.parametrize("engine", list(dft.Engine))
@pytest.mark.parametrize("method", [dft.Method.DOPR54, dft.Method.DOP853])
@pytest.mark.parametrize("implementation", (dft.Implementation.SCIPY, dft.Implementation.DRIFT))
def test_unimplemented_integrator_is_rejected_before_execution(engine, method, implementation):
    message = rf"{engine.value} \+ {implementation.value} {method.value} is not implemented"
    with pytest.raises(NotImplementedError, match=message):
        dft.Config(engine=engine, method=method, implementation=implementation)


def test_input_validation_still_raises_plain_value_error():
    with pytest.raises(ValueError) as raised:
        dft.Config(tolerance=-1.0)
    assert not isinstance(raised.value, dft.DriftError)


@pytest.mark.parametrize("tol", [(1e-9, 1e-12), 1e-10])
def test_tolerances_and_log_conversion(tol):
    c = dft.Config(tolerance=tol)
    r, a = (tol, tol) if isinstance(tol, float) else tol
    assert (c.rtol, c.atol) == (r, a)
    assert dft.log_tolerance(r, a) == (math.log(r), math.log(a))


@pytest.mark.parametrize(("selector", "value", "expected"), (
    (dft.Engine, "CPU", dft.Engine.CPU),
    (dft.Method, "DOP853", dft.Method.DOP853),
    (dft.Implementation, "GALPY", dft.Implementation.GALPY),
))
def test_selectors_support_value_construction(selector, value, expected):
    assert selector(value) is expected


@pytest.mark.parametrize("keyword", ({"engine": "CPU"}, {"method": "DOPR54"},
                                     {"implementation": "GALPY"}))
def test_config_rejects_raw_strings(keyword):
    with pytest.raises(TypeError, match="expected a drift.* member"):
        dft.Config(**keyword)


@pytest.mark.parametrize(("factory", "args"), ((dft.Potential.kepler, ()),
                                               (dft.Potential.plummer, ()),
                                               (dft.Potential.plummer, (1.0,))))
def test_potential_parameters_are_required(factory, args):
    with pytest.raises(TypeError):
        factory(*args)


@pytest.mark.parametrize(("potential", "expected"), (
    (dft.Potential.kepler(2.5), "Potential.kepler(amp=2.5)"),
    (dft.Potential.plummer(3.5, 0.25), "Potential.plummer(amp=3.5, radius=0.25)"),
    (dft.Potential.bovy(), "Potential.bovy()"),
))
def test_potential_reprs_use_public_constructors(potential, expected):
    assert repr(potential) == expected


def test_container_reprs():
    k = dft.Potential.kepler(1.0)
    assert repr(dft.background(k)) == (
        "Container(kind='background', potential=Potential.kepler(amp=1.0))")
    assert repr(dft.test_particles(np.zeros((2, 6)))) == (
        "Container(kind='test_particles', num_particles=2)")
    assert repr(dft.particles(k, np.zeros((1, 6)))) == (
        "Container(kind='particles', num_particles=1, potential=Potential.kepler(amp=1.0))")


@pytest.mark.parametrize("state", (np.zeros((2, 6)), np.zeros(12)))
def test_initial_state_accepts_particle_records(state):
    assert dft.test_particles(state).num_particles == 2


@pytest.mark.parametrize(("state", "message"), (
    (np.zeros((2, 5)), "shape"),
    (np.zeros(7), "shape"),
    (np.zeros((0, 6)), "at least one particle"),
    (np.full((1, 6), np.nan), "finite"),
))
def test_invalid_initial_states(state, message):
    with pytest.raises(ValueError, match=message):
        dft.test_particles(state)


def test_particle_limit_is_lifted():
    assert dft.test_particles(np.zeros((5000, 6))).num_particles == 5000


def test_bovy_cannot_be_attached_to_particles():
    with pytest.raises(NotImplementedError):
        dft.particles(dft.Potential.bovy(), np.zeros((1, 6)))


def test_empty_config_returns_no_results():
    assert dft.Config().run() == []


def test_config_construction_does_not_print(capsys):
    dft.Config()
    captured = capsys.readouterr()
    assert captured.out == "" and captured.err == ""


def test_info_and_repr_summary():
    sim = dft.Config()
    k = dft.Potential.kepler(1.0)
    sim.add(dft.particles(k, np.zeros((1, 6))), dft.background(k))
    summary = sim.info()
    assert repr(sim) == summary
    assert summary.startswith("Config(engine=CPU, method=DOPR54")
    assert "implementation=GALPY" in summary
    assert "containers=2" in summary and "dependencies=1" in summary
    assert len(summary) < 300


def test_particle_group_requires_a_registered_force_source():
    sim = dft.Config()
    iso = dft.test_particles(np.zeros((1, 6)))
    gmc = dft.particles(dft.Potential.kepler(1.0), np.zeros((1, 6)))
    sim.add(iso, gmc)
    with pytest.raises(ValueError, match="registered with add"):
        sim.run()


def test_add_rejects_cycles_and_empty_requirements():
    a = dft.test_particles(np.zeros((1, 6)))
    b = dft.particles(dft.Potential.kepler(1.0), np.zeros((1, 6)))
    sim = dft.Config()
    sim.add(a, b)
    with pytest.raises(ValueError, match="cycle"):
        sim.add(b, a)
    with pytest.raises(ValueError, match="at least one dependency"):
        sim.add(a)


def test_add_enforces_the_model_capacity():
    sim = dft.Config()
    bg = dft.background(dft.Potential.kepler(1.0))
    for _ in range(10):
        sim.add(dft.test_particles(np.zeros((1, 6))), bg)
    with pytest.raises(ValueError, match="at most 11"):
        sim.add(dft.test_particles(np.zeros((1, 6))), bg)


def test_dependency_alias_is_deprecated():
    sim = dft.Config()
    with warnings.catch_warnings(record=True) as rec:
        warnings.simplefilter("always")
        sim.dependency(dft.test_particles(np.zeros((1, 6))),
                       dft.background(dft.Potential.kepler(1.0)))
    assert any(issubclass(w.category, DeprecationWarning) for w in rec)


def test_cpu_engine_is_not_provided_by_this_build():
    sim = dft.Config(engine=dft.Engine.CPU)
    sim.add(dft.test_particles(np.zeros((1, 6))), dft.background(dft.Potential.kepler(1.0)))
    with pytest.raises(NotImplementedError, match="GPU engine only"):
        sim.run()


def test_model_force_resolution():
    k = dft.Potential.plummer(2.0, 0.3)
    sim = dft.Config(engine=dft.Engine.GPU, forces="model")
    node = dft.test_particles(np.zeros((1, 6)))
    sim.add(node, dft.background(k))
    spec = sim._force_for(node)
    assert spec.kind == 2 and spec.amp == 2.0 and spec.b == 0.3
    assert dft.Config(engine=dft.Engine.GPU)._force_for(node) is None
    with pytest.raises(ValueError):
        dft.Config(forces="bogus")


def test_model_force_resolution_of_several_backgrounds():
    sim = dft.Config(engine=dft.Engine.GPU, forces="model", ts=(0.0, 4.0, 5))
    node = dft.test_particles(np.zeros((1, 6)))
    sim.add(node, dft.background(dft.Potential.bovy()), dft.background(dft.Potential.kepler(0.5)),
            dft.background(dft.Potential.plummer(0.2, 0.1)))
    spec = sim._force_for(node)
    # the bovy model is the static force; the point masses become origins at rest
    assert spec.kind == 3 and spec.origins is not None
    np.testing.assert_array_equal(spec.origins.params, [[0.5, 0.0], [0.2, 0.1]])
    assert not spec.origins.tracks.any()
    assert (spec.origins.t_start, spec.origins.t_end, spec.origins.n_segments) == (0.0, 4.0, 4)
    lone = dft.Config(engine=dft.Engine.GPU, forces="model")
    free = dft.test_particles(np.zeros((1, 6)))
    lone.add(free, dft.test_particles(np.zeros((1, 6))))
    assert lone._force_for(free).kind == 4  # no potential among the sources: free motion
    two = dft.Config(engine=dft.Engine.GPU, forces="model")
    two.add(free, dft.background(dft.Potential.bovy()), dft.background(dft.Potential.bovy()))
    with pytest.raises(NotImplementedError, match="at most one bovy"):
        two._force_for(free)


def test_model_integration_order_puts_force_sources_first():
    sim = dft.Config(engine=dft.Engine.GPU, forces="model")
    bg = dft.background(dft.Potential.bovy())
    stream = dft.test_particles(np.zeros((2, 6)))
    sat = dft.particles(dft.Potential.plummer(1e-3, 0.01), np.zeros((1, 6)))
    sim.add(stream, bg, sat)
    sim.add(sat, bg)
    assert [c._identity for c in sim._integration_order()] == [sat._identity, stream._identity]
    ref = dft.Config(engine=dft.Engine.GPU)
    ref.add(stream, bg, sat)
    ref.add(sat, bg)
    assert [c._identity for c in ref._integration_order()] == [stream._identity, sat._identity]


def test_arbitrary_times_are_opt_in():
    times = np.logspace(-1.0, 1.0, 9)
    with pytest.raises(ValueError, match="uniformly spaced"):
        dft.Config(ts=times)
    cfg = dft.Config(ts=times, arbitrary_ts=True)
    np.testing.assert_array_equal(cfg.grid.times(), times)
    assert not cfg.grid.is_affine()
    assert dft.Config(ts=np.linspace(0.0, 1.0, 7), arbitrary_ts=True).grid.is_affine()
    back = dft.Config(ts=[3.0, 2.5, 0.5, -1.0], arbitrary_ts=True)
    np.testing.assert_array_equal(back.grid.times(), [3.0, 2.5, 0.5, -1.0])
    for bad, message in (([0.0, 0.5, 0.25, 1.0], "strictly"), ([0.0, float("inf")], "finite"),
                         ([1.0], "between 2")):
        with pytest.raises(ValueError, match=message):
            dft.Config(ts=bad, arbitrary_ts=True)


def test_dopr54_refuses_arbitrary_times_before_launching():
    sim = dft.Config(engine=dft.Engine.GPU, method=dft.Method.DOPR54,
                     ts=[0.0, 0.1, 0.5, 2.0], arbitrary_ts=True)
    sim.add(dft.test_particles(np.array([[1.0, 0, 0, 0, 1.0, 0]])),
            dft.background(dft.Potential.kepler(1.0)))
    with pytest.raises(NotImplementedError, match="DOP853"):
        sim.run()


def test_durations_print_as_the_reference_records_do():
    """Rust's `{:.3?}` of a Duration: the largest unit the value reaches."""
    from astrodrift_amd.api import _rust_duration

    assert _rust_duration(1.5) == "1.500s"
    assert _rust_duration(0.0123456) == "12.346ms"
    assert _rust_duration(2.5e-5) == "25.000µs"
    assert _rust_duration(3e-8) == "30.000ns"
