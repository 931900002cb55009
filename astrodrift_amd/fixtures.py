"""galpy ``.fixture`` dumps: reader, writer, ULP metrics and a GPU runner.

Format written by the reference's galpy instrumentation
(tests/fixtures/galpy_native/galpy-v1.11.2.patch) and parsed by its test
support (tests/galpy_support/mod.rs:185-307):

    dim 6
    nt <nt>
    dt_one <%.17g>
    rtol <%.17g>          natural logarithm, as galpy's C integrators take it
    atol <%.17g>
    t_hex <nt x %016x>    exact output times (a decimal "t" row is also read)
    yo_hex <6 x %016x>    exact initial state (a decimal "yo" row is also read)
    nargs <int>
    states_hex
    state_hex <step> <6 x %016x>   one row per output time (optional)

Hex rows win over decimal rows.  A fixture with no state rows is an
input-only reference (the legacy DOPR54 canonical case).
"""

from __future__ import annotations

import struct
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import _native


def _hex_bits(x: float) -> str:
    return f"{struct.unpack('<Q', struct.pack('<d', float(x)))[0]:016x}"


def _from_hex(h: str) -> float:
    return struct.unpack("<d", struct.pack("<Q", int(h, 16)))[0]


@dataclass
class GalpyFixture:
    dim: int
    nt: int
    dt_one: float
    rtol: float
    atol: float
    times: np.ndarray
    initial_state: np.ndarray
    nargs: int = 0
    expected_state_bits: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint64))
    comments: list[str] = field(default_factory=list)

    # -- parsing (tests/galpy_support/mod.rs:195-255) -----------------------
    @classmethod
    def parse(cls, text: str, name: str = "<fixture>") -> "GalpyFixture":
        dim = nt = None
        dt_one = rtol = atol = None
        times: list[float] = []
        y0: list[float] = []
        nargs = 0
        bits: list[int] = []
        comments = []
        for line in text.splitlines():
            parts = line.split()
            if not parts:
                continue
            key, rest = parts[0], parts[1:]
            if key.startswith("#"):
                comments.append(line)
            elif key == "dim":
                dim = int(rest[0])
            elif key == "nt":
                nt = int(rest[0])
            elif key == "dt_one":
                dt_one = float(rest[0])
            elif key == "rtol":
                rtol = float(rest[0])
            elif key == "atol":
                atol = float(rest[0])
            elif key == "t" and not times:
                times = [float(v) for v in rest]
            elif key == "t_hex":
                times = [_from_hex(v) for v in rest]
            elif key == "yo" and not y0:
                y0 = [float(v) for v in rest]
            elif key == "yo_hex":
                y0 = [_from_hex(v) for v in rest]
            elif key == "nargs":
                nargs = int(rest[0])
            elif key == "state_hex":
                bits.extend(int(v, 16) for v in rest[1:])
        missing = [k for k, v in (("dim", dim), ("nt", nt), ("dt_one", dt_one), ("rtol", rtol),
                                  ("atol", atol)) if v is None]
        if missing:
            raise ValueError(f"{name}: {', '.join(missing)} missing in fixture")
        fx = cls(dim, nt, dt_one, rtol, atol, np.array(times, dtype=np.float64),
                 np.array(y0, dtype=np.float64), nargs, np.array(bits, dtype=np.uint64), comments)
        fx.validate(name)
        return fx

    @classmethod
    def read(cls, path) -> "GalpyFixture":
        path = Path(path)
        return cls.parse(path.read_text(), str(path))

    def validate(self, name: str = "<fixture>") -> None:
        """GalpyFixture::validate, tests/galpy_support/mod.rs:257-295."""
        if self.dim != 6:
            raise ValueError(f"{name}: state dimension {self.dim} != 6")
        if self.times.size != self.nt:
            raise ValueError(f"{name}: {self.times.size} output times != nt {self.nt}")
        if self.initial_state.size != self.dim:
            raise ValueError(f"{name}: initial state length {self.initial_state.size} != "
                             f"dimension {self.dim}")
        if self.expected_state_bits.size and self.expected_state_bits.size != self.nt * self.dim:
            raise ValueError(f"{name}: partial expected state: length "
                             f"{self.expected_state_bits.size} != nt * dimension "
                             f"{self.nt * self.dim}")

    # -- writing (the instrumentation's output format) ----------------------
    def format(self) -> str:
        lines = list(self.comments)
        lines += [f"dim {self.dim}", f"nt {self.nt}", f"dt_one {self.dt_one:.17g}",
                  f"rtol {self.rtol:.17g}", f"atol {self.atol:.17g}",
                  "t_hex " + " ".join(_hex_bits(v) for v in self.times),
                  "yo_hex " + " ".join(_hex_bits(v) for v in self.initial_state),
                  f"nargs {self.nargs}", "states_hex"]
        if self.expected_state_bits.size:
            rows = self.expected_state_bits.reshape(self.nt, self.dim)
            for step, row in enumerate(rows):
                lines.append(f"state_hex {step} " + " ".join(f"{int(b):016x}" for b in row))
        return "\n".join(lines) + "\n"

    def write(self, path) -> None:
        Path(path).write_text(self.format())

    @classmethod
    def from_run(cls, times, y0, states, rtol_ln, atol_ln, dt_one=_native.INITIAL_STEP_SENTINEL,
                 comments=()) -> "GalpyFixture":
        states = np.ascontiguousarray(states, dtype=np.float64)
        return cls(6, len(times), float(dt_one), float(rtol_ln), float(atol_ln),
                   np.asarray(times, dtype=np.float64), np.asarray(y0, dtype=np.float64), 0,
                   states.view(np.uint64).reshape(-1).copy(), list(comments))

    def expected_states(self) -> np.ndarray:
        return self.expected_state_bits.view(np.float64).reshape(self.nt, self.dim)


def ordered_bits(bits: np.ndarray) -> np.ndarray:
    """Monotone integer image of IEEE doubles (tests/galpy_support/mod.rs:330-337)."""
    b = np.asarray(bits, dtype=np.uint64)
    sign = (b >> np.uint64(63)) != 0
    return np.where(sign, ~b, b | np.uint64(1 << 63))


def ulp_distance(a, b) -> np.ndarray:
    ab = np.asarray(a, dtype=np.float64).view(np.uint64)
    bb = np.asarray(b, dtype=np.float64).view(np.uint64)
    oa, ob = ordered_bits(ab), ordered_bits(bb)
    return np.where(oa > ob, oa - ob, ob - oa)


def error_summary(actual: np.ndarray, expected: np.ndarray) -> dict:
    """The reference's ErrorSummary: mismatches, worst absolute and ULP error."""
    actual = np.asarray(actual, dtype=np.float64).reshape(-1)
    expected = np.asarray(expected, dtype=np.float64).reshape(-1)
    ulps = ulp_distance(actual, expected)
    return {
        "values": int(actual.size),
        "mismatched": int(np.count_nonzero(ulps)),
        "max_abs": float(np.max(np.abs(actual - expected))),
        "max_ulp": int(ulps.max()) if ulps.size else 0,
    }


def run_fixture_gpu(method: int, fixture: GalpyFixture, potential=None, device: int = 0):
    """Integrates a fixture's inputs on the GPU (LogTolerance::from_logarithmic,
    tests/galpy_integrator_tests.rs:78) and returns the (nt, 6) trajectory."""
    out, stats, code = _native.launch_host(method, fixture.initial_state.reshape(1, 6),
                                           fixture.times, fixture.rtol, fixture.atol, potential,
                                           dt_one=fixture.dt_one, device=device)
    return out[:, 0, :], stats[0], code
