"""Build recipe of libgsolve_hip: hipcc for gfx950, in-tree (lib/libgsolve_hip.so).

Flags:
  -ffp-contract=off                 the reference's Julia arithmetic does not fuse a*b + c
  -fno-gpu-flush-denormals-to-zero  keep subnormals, as the reference's scaling rules assume
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                    no packed FP32 instructions in the generated code
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "lib" / "libgsolve_hip.so"
SOURCES = ["gsolve.hip", "batch_reduce.hip"]
ARCH = os.environ.get("GSOLVE_OFFLOAD_ARCH", "gfx950")


def hipcc_command(out: Path) -> list[str]:
    return [
        "hipcc", f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
        "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-function",
        "-o", str(out), *[str(CSRC / s) for s in SOURCES],
    ]


def build(verbose: bool = False) -> Path:
    """Compile libgsolve_hip.so (always; the caller decides when a rebuild is needed)."""
    LIB.parent.mkdir(parents=True, exist_ok=True)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = hipcc_command(tmp)
    res = subprocess.run(cmd, cwd=CSRC, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in (res.stdout + res.stderr).splitlines() if noise not in l and "(ignoring feature)" not in l)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{err}")
    if verbose and err.strip():
        print(err, file=sys.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
