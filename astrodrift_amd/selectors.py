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
    """Execution backend."""

    CPU = "CPU"
    GPU = "GPU"


class Method(_StrEnum):
    """Adaptive integration method.

    DOPR54 is Dormand-Prince 5(4); DOP853 is Dormand-Prince 8(5,3).
    """

    DOPR54 = "DOPR54"
    DOP853 = "DOP853"


class Implementation(_StrEnum):
    """Numerical implementation of the selected integration method."""

    GALPY = "GALPY"
    SCIPY = "SCIPY"
    DRIFT = "DRIFT"


__all__ = ["Engine", "Implementation", "Method"]
