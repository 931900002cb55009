"""Builds libdrift_hip.so (gfx950) in-tree with hipcc.

The library holds the HIP kernels and the C ABI declared in
include/drift_hip.h.  It is compiled for gfx950 only; the host part links
against the HIP runtime.  Run ``python -m astrodrift_amd.build``.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PACKAGE_DIR = Path(__file__).resolve().parent
REPO_DIR = PACKAGE_DIR.parent
CSRC = PACKAGE_DIR / "csrc"
LIB_PATH = PACKAGE_DIR / "libdrift_hip.so"
SOURCES = [
    CSRC / "drift_hip.hip",
    CSRC / "ragged.hip",
    CSRC / "serving.cpp",
    CSRC / "spherical_table.cpp",
    CSRC / "origins.cpp",
]
HEADERS = sorted(CSRC.glob("*.hpp")) + [REPO_DIR / "include" / "drift_hip.h"]

ARCH = "gfx950"


def hipcc() -> str:
    found = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(found).exists():
        raise RuntimeError("hipcc not found; the ROCm toolchain is required to build")
    return found


def compile_command(out: Path = LIB_PATH, fp_contract: str = "fast") -> list[str]:
    """hipcc command line.  fp_contract="off" is a diagnostic build: the
    compiler then fuses no multiply-add by itself (the kernels' explicit fma()
    calls stay fused), which separates rounding differences that come from
    contraction from those of the step control; the default contracts."""
    if fp_contract not in ("fast", "off"):
        raise ValueError("fp_contract must be 'fast' or 'off'")
    return [
        hipcc(),
        f"--offload-arch={ARCH}",
        "-O3",
        f"-ffp-contract={fp_contract}",
        "-std=c++17",
        "-fPIC",
        "-shared",
        # no packed FP32 instructions: some of their forms have produced
        # silently wrong results on this hardware
        "-Xclang",
        "-target-feature",
        "-Xclang",
        "-packed-fp32-ops",
        f"-I{REPO_DIR / 'include'}",
        *[str(s) for s in SOURCES],
        "-o",
        str(out),
    ]


def build(verbose: bool = False, fp_contract: str = "fast") -> Path:
    """Compiles the library unconditionally and returns its path."""
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = compile_command(tmp, fp_contract)
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"hipcc failed with exit code {proc.returncode}")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser(description="build libdrift_hip.so for gfx950")
    ap.add_argument("--fp-contract", choices=("fast", "off"), default="fast",
                    help="off: diagnostic build without compiler-formed multiply-adds "
                         "(rebuild without the flag afterwards)")
    args = ap.parse_args()
    print(build(verbose=True, fp_contract=args.fp_contract))
