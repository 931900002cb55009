"""Selector enums of the public API (python/drift/selectors.py:4-26).

The reference uses ``enum.StrEnum`` (Python >= 3.11); a ``str`` mix-in Enum
gives the same value construction, comparison and formatting on Python 3.10.
"""

from enum import Enum


class _StrEnum(str, Enum):
    def __str__(self) -> str:
        return str(self.value)

    def __format__(self, spec: str) -> str:
        return format(str(self.value), spec)


class Engine(_StrEnum):
    """Where run() integrates: on the host (CPU) or on the MI355X (GPU).  This
    package runs GPU configurations only; CPU ones validate but do not run."""

    CPU = "CPU"
    GPU = "GPU"


class Method(_StrEnum):
    """Embedded Runge-Kutta pair with adaptive steps: DOPR54, the 5th-order
    pair with a 4th-order error estimate, or DOP853, the 8th-order pair with
    5th- and 3rd-order estimates and 7th-order dense output."""

    DOPR54 = "DOPR54"
    DOP853 = "DOP853"


class Implementation(_StrEnum):
    """Whose version of the method runs; GALPY (the step control of galpy's C
    integrators) is the one with kernels."""

    GALPY = "GALPY"
    SCIPY = "SCIPY"
    DRIFT = "DRIFT"


__all__ = ["Engine", "Implementation", "Method"]
