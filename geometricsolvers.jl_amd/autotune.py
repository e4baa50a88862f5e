"""The launch-shape autotuner, with no counterpart in the reference. Every tuner times the real entry
point on the caller's own batch (the best of `repeats` synchronised calls per candidate, each on a
fresh copy of the start) and leaves the fastest shape in place; results never depend on a shape.

- `tune_lsq`: the least-squares batch runs one persistent kernel per size bucket (K = 1..4
  exponential terms, n = 2K); each has two knobs, the threads per workgroup and the resident
  workgroups per CU (0: as many as the kernel's occupancy allows). The buckets are tuned one by one
  (coordinate descent, the others at their current shapes); they run concurrently, so each is timed
  in the mix it runs in.
- `tune_dense`: the dense mid-size family (`DenseSin`, n = 16 … 64) takes a wave segment per system;
  the segment's width (16, 32 or 64 lanes: four, two or one systems per wave) is the choice.
- `tune_workgroup`: the lane-per-system families (n ≤ 8) take one work-item per system; the
  workgroup size is the choice, returned for the caller to pass as `workgroupsize`.

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


# ---- the dense mid-size family: lanes per system ----------------------------------------------------

DENSE_LANES = {12: (16, 32), 16: (16, 32), 24: (32, 64), 32: (32, 64), 48: (64,), 64: (64,)}


def dense_get_shape(n: int) -> int:
    """The lanes per system the next dense launch at size n uses."""
    from . import _abi

    lib = _abi.load_library()
    r = lib.gs_dense_get_shape(int(n))
    if r < 0:
        from . import _raise_on

        _raise_on(r, lib)
    return int(r)


def dense_set_shape(n: int, lanes: int) -> None:
    """Set the lanes per system of the dense launches at size n (0: the default)."""
    from . import _abi, _raise_on

    lib = _abi.load_library()
    _raise_on(lib.gs_dense_set_shape(int(n), int(lanes)), lib)


def _timer(call, X, X0, repeats):
    import torch

    def run() -> float:
        best = float("inf")
        for _ in range(repeats):
            X.copy_(X0)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            call()
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        return best

    return run


def tune_dense(X0, prob, method, p, options=None, repeats: int = 2) -> dict:
    """Tune the lanes per system of the dense family's launches at the batch's size on the batch
    (X0, prob, method, p) and leave the library set to the fastest. X0 is not modified. Returns
    {"n": n, "lanes": chosen, "seconds": {lanes: time}}."""
    import torch

    from . import DenseSin, solve_batch, status_buffer

    if repeats < 1:
        raise ValueError("repeats must be ≥ 1")
    fam = prob.F
    if not isinstance(fam, DenseSin):
        raise TypeError("tune_dense takes a NonlinearProblem over DenseSin")
    n = fam.n
    if n not in DENSE_LANES:
        # n ≤ 8 runs one lane per system: no segment width to choose
        return {"n": n, "lanes": None, "seconds": {}}
    X = torch.empty_like(X0)
    S = status_buffer(X0.shape[0], X0.device, "float32" if X0.dtype == torch.float32 else "float64")
    run = _timer(lambda: solve_batch(X, S, prob, method, p, options), X, X0, repeats)
    timings = {}
    for lanes in DENSE_LANES[n]:
        dense_set_shape(n, lanes)
        run()  # first launch of this shape: code loading
        timings[lanes] = run()
    choice = min(timings, key=timings.get)
    dense_set_shape(n, choice)
    return {"n": n, "lanes": choice, "seconds": timings}


WORKGROUP_CANDIDATES = (64, 128, 256)


def tune_workgroup(X0, prob, method, p, options=None, candidates=WORKGROUP_CANDIDATES, repeats: int = 2) -> dict:
    """Time the lane-per-system solve of (X0, prob, method, p) at each workgroup size and return
    {"workgroupsize": fastest, "seconds": {size: time}}; the caller passes the size to solve_batch.
    X0 is not modified."""
    import torch

    from . import solve_batch, status_buffer

    if repeats < 1:
        raise ValueError("repeats must be ≥ 1")
    X = torch.empty_like(X0)
    S = status_buffer(X0.shape[0], X0.device, "float32" if X0.dtype == torch.float32 else "float64")
    timings = {}
    for wg in candidates:
        run = _timer(lambda: solve_batch(X, S, prob, method, p, options, workgroupsize=int(wg)), X, X0, repeats)
        run()
        timings[int(wg)] = run()
    choice = min(timings, key=timings.get)
    return {"workgroupsize": choice, "seconds": timings}
