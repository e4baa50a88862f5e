"""Throughput of the batched FP64 Newton solve on one MI355X.

Workload (the configuration the headline metric is quoted on): one implicit-midpoint step of the
Hénon–Heiles system for 2^20 independent states, n = 4 unknowns each, h = 0.1, states on energy
shells E ~ U[0.01, 0.15]; Newton + Static + static LU in FP64 from x₀ = x_k, f_abstol = 1e-12,
f_reltol = 0, status form of solve_batch (72-byte SolverStatus per system). A "step" is one
solve_batch call over a fresh copy of the batch; every copy is made before the clock starts, as the
reference's device benchmark does (benchmark/devices.jl:372-404).

value = converged systems/s: systems with code SUCCESS whose FP64 residual, rechecked on the host
after the timed region, satisfies ‖F(x)‖₂ ≤ 1e-12; divided by the wall time of the K timed steps.

It prints one JSON line (rank 0, one GPU).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

NSYS = 1 << 20
H = 0.1
TOL = 1e-12
# algorithmic HBM bytes per system of the status form: x_k 32 (record) + X in 32 + X out 32 + S 72
BYTES_PER_SYSTEM = 32 + 32 + 32 + 72
HBM_PEAK_GBS = 8000.0


def cpu_baseline(gs, workloads, xk, min_seconds: float = 8.0):
    """The oracle (C++ restatement of the reference's R1 loop, OpenMP over systems) on the host cores:
    a bounded sample of the same workload, repeated until `min_seconds` of wall time."""
    from oracle import oracle as orc

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    sample = xk[: 1 << 18]
    opts = gs.Options(f_abstol=TOL, f_reltol=0.0)
    orc.solve_batch(gs.Newton(), gs.HenonHeilesMidpoint(), sample.copy(), gs.BatchParams(H, sample), opts,
                    nthreads=threads)  # warm-up
    reps, t0 = 0, time.perf_counter()
    while True:
        x, S = orc.solve_batch(gs.Newton(), gs.HenonHeilesMidpoint(), sample.copy(), gs.BatchParams(H, sample),
                               opts, nthreads=threads)
        reps += 1
        el = time.perf_counter() - t0
        if el >= min_seconds:
            break
    conv = int((S["code"] == 0).sum())
    return {
        "value": conv * reps / el,
        "unit": "converged systems/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{len(sample)} systems of the same workload (seed 0) x {reps} repetitions, {el:.1f} s wall, "
                  f"OpenMP schedule(static) over systems",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the systems are independent and this build runs on a single "
              "MI355X; --gpus above 1 is not supported", file=sys.stderr)
        return 2

    import torch

    import __graft_entry__

    gs = __graft_entry__.load_package()
    import importlib

    workloads = importlib.import_module(gs.__name__ + ".workloads")
    lib = gs.load_library()
    dev = torch.device("cuda:0")
    xk = workloads.henon_heiles_states(NSYS, seed=0)
    Xk = torch.as_tensor(xk, device=dev)
    K, W = args.steps, args.warmup
    starts = [Xk.clone() for _ in range(K + W)]
    S = [gs.status_buffer(NSYS, dev) for _ in range(K)]
    S_warm = gs.status_buffer(NSYS, dev)
    prob = gs.NonlinearProblem(gs.HenonHeilesMidpoint())
    method = gs.Newton()
    opts = gs.Options(f_abstol=TOL, f_reltol=0.0)
    p = gs.BatchParams(H, Xk)

    lib.gs_set_kernel_timing(1)
    for i in range(W):
        gs.solve_batch(starts[K + i], S_warm, prob, method, p, opts)
    torch.cuda.synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    for i in range(K):
        gs.solve_batch(starts[i], S[i], prob, method, p, opts)
        kernel_ms.append(lib.gs_last_kernel_ms())
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    lib.gs_set_kernel_timing(0)

    # host FP64 recheck of every step's result, outside the timed region
    converged = 0
    for i in range(K):
        st = gs.statuses(S[i])
        x = starts[i].cpu().numpy()
        r = np.sqrt((workloads.henon_heiles_midpoint_residual(x, xk, H) ** 2).sum(axis=1))
        converged += int(((st["code"] == 0) & (r <= TOL)).sum())
    iters = gs.statuses(S[-1])["iterations"]

    value = converged / elapsed
    kms = float(np.mean(kernel_ms))
    achieved = BYTES_PER_SYSTEM * NSYS / (kms * 1e-3) / 1e9
    traffic = None
    tfile = ROOT / "profiles" / "traffic_henon_heiles_midpoint_2p20.json"
    if tfile.exists():
        traffic = json.loads(tfile.read_text()).get("bytes_per_launch")
    line = {
        "metric": "converged systems/s at residual ≤1e-12 (FP64 check), learned+mixed vs plain, 1 GPU",
        "value": value,
        "unit": "converged systems/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": elapsed / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (numpy PCG64 seed 0, Hénon–Heiles energy shells)",
        "config": {
            "workload": "henon_heiles_implicit_midpoint n=4 FP64, batch 2^20, h=0.1, "
                        "Newton+Static+LU, abstol 1e-12, reltol 0, status form",
            "method": "plain FP64 Newton (lane per system)",
            "batch": NSYS,
            "mean_iterations": float(iters.mean()),
            "kernel_ms": kms,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
        },
        "cpu_baseline": None if args.no_cpu_baseline else cpu_baseline(gs, workloads, xk),
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
