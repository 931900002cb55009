"""astrodrift_amd -- MI355X-native batched Dormand-Prince orbit integration.

Drop-in for the GPU engine of drift (lbparticles/astrodrift): the same
``Config`` / ``add`` / ``run`` API, with every integration executed by the
gfx950 kernels in ``libdrift_hip.so`` (C ABI: include/drift_hip.h).

    import astrodrift_amd as dft
    sim = dft.Config(engine=dft.Engine.GPU, method=dft.Method.DOP853)
    sim.add(dft.test_particles(istate), dft.background(dft.Potential.kepler(1.0)))
    trajectories = sim.run()  # [ndarray (nt, n, 6)]
"""

from ._native import (
    NativeLibraryError,
    PotentialSpec,
    launch_device,
    launch_host,
    load_library,
)
from .api import (
    Config,
    Container,
    DriftError,
    IntegrationError,
    OutputGrid,
    Potential,
    background,
    log_tolerance,
    particles,
    test_particles,
)
from .selectors import Engine, Implementation, Method

__all__ = [
    "Potential",
    "Container",
    "Config",
    "Implementation",
    "Method",
    "Engine",
    "DriftError",
    "IntegrationError",
    "particles",
    "test_particles",
    "background",
    "OutputGrid",
    "PotentialSpec",
    "NativeLibraryError",
    "launch_host",
    "launch_device",
    "load_library",
    "log_tolerance",
]
