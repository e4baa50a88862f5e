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
SOURCES = ["gsolve.hip", "small_reuse.hip", "small_dogleg.hip", "small_f32.hip", "small_deflated.hip", "small_stationarity.hip", "small_reparam.hip", "small_stationarity_f32.hip", "small_reparam_f32.hip", "batch_reduce.hip", "lsq.hip", "mlp.hip", "mid_solver.hip", "mid_reuse.hip", "dense_solver.hip", "small_dense.hip", "dense_n12.hip", "dense_r12.hip", "dense_n16.hip", "dense_n24.hip", "dense_n32.hip", "dense_n48.hip", "dense_n64.hip", "dense_r16.hip", "dense_r24.hip", "dense_r32.hip", "dense_r48.hip", "dense_r64.hip", "dense_d12.hip", "dense_d16.hip", "dense_d24.hip", "dense_d32.hip", "dense_d48.hip", "dense_d64.hip", "dense_f16.hip", "dense_f24.hip", "dense_f32.hip", "dense_f48.hip", "dense_f64.hip"]
ARCH = os.environ.get("GSOLVE_OFFLOAD_ARCH", "gfx950")
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
OBJ = HERE / "lib" / "obj"


def _filter(text: str) -> str:
    noise = "is not a recognized feature for this target"
    return "\n".join(l for l in text.splitlines() if noise not in l and "(ignoring feature)" not in l)


def build(verbose: bool = False) -> Path:
    """Compile libgsolve_hip.so (always; the caller decides when a rebuild is needed): one hipcc per
    source, in parallel, then one link."""
    OBJ.mkdir(parents=True, exist_ok=True)
    procs = []
    for src in SOURCES:
        obj = OBJ / (Path(src).stem + ".o")
        cmd = ["hipcc", *FLAGS, "-c", "-o", str(obj), str(CSRC / src)]
        procs.append((src, obj, subprocess.Popen(cmd, cwd=CSRC, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    errors, objs = [], []
    for src, obj, p in procs:
        out, _ = p.communicate()
        out = _filter(out)
        if p.returncode != 0:
            errors.append(f"hipcc {src} failed ({p.returncode}):\n{out}")
        elif verbose and out.strip():
            print(out, file=sys.stderr)
        objs.append(str(obj))
    if errors:
        raise RuntimeError("\n".join(errors))
    tmp = LIB.with_suffix(".so.tmp")
    res = subprocess.run(["hipcc", *FLAGS, "-shared", "-o", str(tmp), *objs], cwd=CSRC, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc link failed ({res.returncode}):\n{_filter(res.stdout + res.stderr)}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
