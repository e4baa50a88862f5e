"""The launch-shape autotuner of the least-squares batch, with no counterpart in the reference.

The batch runs one persistent kernel per size bucket (K = 1..4 exponential terms, n = 2K); each has
two knobs, the threads per workgroup and the resident workgroups per CU (0: as many as the kernel's
occupancy allows). `tune_lsq` times the real entry point on the caller's own batch and sets, bucket
by bucket (coordinate descent, the other buckets at their current shapes), the candidate with the
smallest time; the buckets run concurrently, so each is timed in the mix it runs in.

    best = tune_lsq(X0, data)     # then solve_batch(...) runs with the tuned shapes
"""
from __future__ import annotations

import ctypes
import time

DEFAULT_CANDIDATES = ((64, 0), (64, 1), (64, 2), (128, 0), (128, 1), (256, 0))


def get_shape(k: int, nfit: int = 1):
    """(threads, blocks_per_cu, workgroups for nfit fits) of bucket K = k."""
    from . import _abi, _raise_on

    lib = _abi.load_library()
    out = (ctypes.c_int64 * 3)()
    _raise_on(lib.gs_lsq_get_shape(int(k), int(nfit), out), lib)
    return int(out[0]), int(out[1]), int(out[2])


def set_shape(k: int, threads: int, blocks_per_cu: int) -> None:
    from . import _abi, _raise_on

    lib = _abi.load_library()
    _raise_on(lib.gs_lsq_set_shape(int(k), int(threads), int(blocks_per_cu)), lib)


def tune_lsq(X0, data, method=None, options=None, candidates=DEFAULT_CANDIDATES, repeats: int = 2) -> dict:
    """Tune the four buckets' shapes on the batch (X0, data) and leave the library set to the best.
    X0 is not modified (every trial solves a copy). Returns {k: (threads, blocks_per_cu, seconds)}."""
    import torch

    from . import LeastSquaresProblem, ExpSumFits, LevenbergMarquardt, Newton, solve_batch, status_buffer

    if repeats < 1:
        raise ValueError("repeats must be ≥ 1")
    method = method if method is not None else Newton(globalization=LevenbergMarquardt())
    prob = LeastSquaresProblem(ExpSumFits())
    X = torch.empty_like(X0)
    S = status_buffer(X0.shape[0], X0.device)

    def run() -> float:
        best = float("inf")
        for _ in range(repeats):
            X.copy_(X0)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            solve_batch(X, S, prob, method, data, options)
            best = min(best, time.perf_counter() - t0)
        return best

    run()  # first call: workspaces, graph capture, code loading
    result = {}
    for k in (1, 2, 3, 4):
        current = get_shape(k)[:2]
        timings = {}
        for cand in dict.fromkeys((current,) + tuple(candidates)):
            set_shape(k, *cand)
            timings[cand] = run()
        choice = min(timings, key=timings.get)
        set_shape(k, *choice)
        result[k] = (choice[0], choice[1], timings[choice])
    return result
