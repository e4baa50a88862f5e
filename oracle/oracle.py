"""CPU ORACLE — TEST INFRASTRUCTURE ONLY.

ctypes binding of oracle/_build/liboracle.so, the C++ restatement of the reference's R1 solve
(oracle.hpp). Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may use it, as
the checker or as the timed CPU baseline; the product path never does.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"

_LIB = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib():
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        vp = ctypes.c_void_p
        L.orc_solve_batch_f64.restype = ctypes.c_int
        L.orc_solve_batch_f64.argtypes = [vp, vp, ctypes.c_int64, vp, vp, ctypes.c_int64, vp, vp, vp, vp, ctypes.c_int]
        L.orc_solve_batch_f32.restype = ctypes.c_int
        L.orc_solve_batch_f32.argtypes = L.orc_solve_batch_f64.argtypes
        L.orc_solve_batch_deflated_f64.restype = ctypes.c_int
        L.orc_solve_batch_deflated_f64.argtypes = [vp, vp, vp, ctypes.c_int64, vp, vp, ctypes.c_int64, vp, vp, vp, vp,
                                                   ctypes.c_int]
        L.orc_solve_batch_stationarity_f64.restype = ctypes.c_int
        L.orc_solve_batch_stationarity_f64.argtypes = [vp, vp, vp, ctypes.c_int64, vp, vp, ctypes.c_int64, vp, vp, vp,
                                                       vp, ctypes.c_int]
        L.orc_solve_batch_reparametrized_f64.restype = ctypes.c_int
        L.orc_solve_batch_reparametrized_f64.argtypes = L.orc_solve_batch_deflated_f64.argtypes
        L.orc_solve_batch_stationarity_f32.restype = ctypes.c_int
        L.orc_solve_batch_stationarity_f32.argtypes = L.orc_solve_batch_stationarity_f64.argtypes
        L.orc_solve_batch_reparametrized_f32.restype = ctypes.c_int
        L.orc_solve_batch_reparametrized_f32.argtypes = L.orc_solve_batch_deflated_f64.argtypes
        L.orc_hh_action_eval.restype = None
        L.orc_hh_action_eval.argtypes = [ctypes.c_double, ctypes.c_double, vp, vp, vp]
        L.orc_pow_int.restype = None
        L.orc_pow_int.argtypes = [vp, ctypes.c_int64, vp, ctypes.c_int64]
        L.orc_standard_cell.restype = ctypes.c_int
        L.orc_standard_cell.argtypes = [ctypes.c_int] * 4 + [vp] * 5
        L.orc_standard_u0.restype = ctypes.c_int
        L.orc_standard_u0.argtypes = [ctypes.c_int, ctypes.c_int, vp]
        L.orc_lu_factor.restype = ctypes.c_int
        L.orc_lu_factor.argtypes = [ctypes.c_int, vp, ctypes.c_int, vp, vp, vp, vp]
        L.orc_lu_solve.restype = ctypes.c_int
        L.orc_lu_solve.argtypes = [ctypes.c_int, vp, vp, ctypes.c_int, vp]
        L.orc_rnorm.restype = ctypes.c_double
        L.orc_rnorm.argtypes = [vp, ctypes.c_int]
        L.orc_rhs_exponent4.restype = ctypes.c_int
        L.orc_rhs_exponent4.argtypes = [vp, ctypes.c_int]
        L.orc_headroom.restype = ctypes.c_int
        L.orc_headroom.argtypes = [ctypes.c_int]
        L.orc_backtrack_step.restype = ctypes.c_double
        L.orc_backtrack_step.argtypes = [ctypes.c_double] * 7
        L.orc_max_threads.restype = ctypes.c_int
        L.orc_solve_expsum_batch.restype = ctypes.c_int
        L.orc_solve_expsum_batch.argtypes = [vp, vp, ctypes.c_int64, vp, vp, vp, vp, vp, vp, ctypes.c_int]
        L.orc_solve_mgh.restype = ctypes.c_int
        L.orc_solve_mgh.argtypes = [ctypes.c_int, vp, vp, vp, vp, vp]
        L.orc_solve_linear_lsq.restype = ctypes.c_int
        L.orc_solve_linear_lsq.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp, vp, vp, vp, vp]
        L.orc_lm_damping_diagonal.restype = ctypes.c_int
        L.orc_lm_damping_diagonal.argtypes = [ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, vp]
        L.orc_fpu_gl3_eval.restype = None
        L.orc_fpu_gl3_eval.argtypes = [ctypes.c_double, ctypes.c_double, vp, vp, vp, vp]
        L.orc_exp_family.restype = None
        L.orc_exp_family.argtypes = [vp, vp, ctypes.c_int64]
        L.orc_givens.restype = None
        L.orc_givens.argtypes = [ctypes.c_double, ctypes.c_double, vp]
        L.orc_solve_fixture.restype = ctypes.c_int
        L.orc_solve_fixture.argtypes = [ctypes.c_int, vp, vp, vp, vp, vp]
        L.orc_dogleg_pw_step.restype = ctypes.c_int
        L.orc_dogleg_pw_step.argtypes = [ctypes.c_int, vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, vp]
        L.orc_solve_fixture_counts.restype = ctypes.c_int
        L.orc_solve_fixture_counts.argtypes = [ctypes.c_int, vp, vp, vp, vp, vp, vp]
        L.orc_solve_fixture_t.restype = ctypes.c_int
        L.orc_solve_fixture_t.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp, vp, vp, vp, vp]
        L.orc_fixture_eval.restype = ctypes.c_int
        L.orc_fixture_eval.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp, vp, vp]
        L.orc_fixture_step.restype = ctypes.c_int
        L.orc_fixture_step.argtypes = [ctypes.c_int, vp, vp, vp, vp]
        L.orc_refine_fixture.restype = ctypes.c_int
        L.orc_refine_fixture.argtypes = [ctypes.c_int, vp, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, vp]
        L.orc_read_mixed_stall.restype = None
        L.orc_read_mixed_stall.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                           ctypes.c_double, ctypes.c_double, vp]
        L.orc_dense_eval.restype = ctypes.c_int
        L.orc_dense_eval.argtypes = [ctypes.c_int, vp, vp, vp, vp, vp]
        L.orc_fsincos.restype = None
        L.orc_fsincos.argtypes = [vp, vp, vp, ctypes.c_int64, ctypes.c_int]
        L.orc_rmaxnorm.restype = ctypes.c_double
        L.orc_rmaxnorm.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        _LIB = L
    return _LIB


def _ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


def host_records(family, p, nsys: int, real=np.float64):
    """The parameter records of a batch on the host: (array, byte stride, shared scalars); `real` is
    the records' float type (the batch's working precision)."""
    name = type(family).__name__
    if name == "FpuGl3":
        rec = np.ascontiguousarray(np.asarray(p.per_system, dtype=real).reshape(nsys, 32))
        return rec, rec.strides[0], (float(p.shared), float(family.beta))
    if name == "HenonHeilesMidpoint":
        rec = np.ascontiguousarray(np.asarray(p.per_system, dtype=real).reshape(nsys, 4))
        return rec, rec.strides[0], (float(p.shared),)
    dtype = np.int64 if name == "Hetero2" else real
    if hasattr(p, "data"):
        rec = np.ascontiguousarray(np.asarray(p.data, dtype=dtype).reshape(nsys, family.p_len))
        return rec, rec.strides[0], ()
    rec = np.ascontiguousarray(np.asarray(p, dtype=dtype).reshape(1, family.p_len))
    return rec, 0, ()


def solve_batch(method, family, X: np.ndarray, p, options, order=None, nthreads: int = 0, deflated=None,
                stationarity=None, b=None, reparametrized=None):
    """The oracle's batched solve on host arrays: returns (X, statuses) with X overwritten.

    `method`, `options`: objects with `to_c()` giving the ABI structs of include/gsolve.h;
    `family`: a registered family (id, n, p_len); `order`: 1-based permutation or None;
    `deflated`: a Deflated (its roots, power and shift; its problem is `family`'s) or None;
    `stationarity`: a StationarityProblem over the action `family` (X rows are full states), with
    the datum `b` overriding its own, or None; `reparametrized`: a Reparametrized over `family`
    (X holds the chart coordinates y₀ and receives x* = g(y*)), or None.
    """
    if stationarity is not None:
        return _solve_stationarity(method, family, X, p, options, order, nthreads, stationarity, b)
    f32 = np.asarray(X).dtype == np.float32
    real = np.float32 if f32 else np.float64
    X = np.ascontiguousarray(X, dtype=real)
    nsys = X.shape[0]
    rec, stride, shared = host_records(family, p, nsys, real)
    cm = method.to_c()
    co = options.to_c()
    fam = _family_struct(family, shared, real)
    S = np.zeros(nsys, dtype=_status_dtype_f32() if f32 else _status_dtype())
    o = None if order is None else np.ascontiguousarray(np.asarray(order, dtype=np.int64))
    if reparametrized is not None:
        cc = reparametrized.to_c(family.n)
        entry = lib().orc_solve_batch_reparametrized_f32 if f32 else lib().orc_solve_batch_reparametrized_f64
        rc = entry(
            ctypes.byref(cm), ctypes.byref(fam), ctypes.byref(cc), nsys, _ptr(X), _ptr(rec), stride, _ptr(S), None,
            None if o is None else _ptr(o), ctypes.byref(co), nthreads)
    elif deflated is not None:
        cd = deflated.to_c(family.n)
        rc = lib().orc_solve_batch_deflated_f64(
            ctypes.byref(cm), ctypes.byref(fam), ctypes.byref(cd), nsys, _ptr(X), _ptr(rec), stride, _ptr(S), None,
            None if o is None else _ptr(o), ctypes.byref(co), nthreads)
    else:
        entry = lib().orc_solve_batch_f32 if f32 else lib().orc_solve_batch_f64
        rc = entry(
            ctypes.byref(cm), ctypes.byref(fam), nsys, _ptr(X), _ptr(rec), stride, _ptr(S), None,
            None if o is None else _ptr(o), ctypes.byref(co), nthreads)
    if rc != 0:
        raise RuntimeError(f"oracle failed with {rc}")
    return X, S


def _solve_stationarity(method, family, X, p, options, order, nthreads, prob, b):
    f32 = np.asarray(X).dtype == np.float32
    real = np.float32 if f32 else np.float64
    X = np.ascontiguousarray(X, dtype=real)
    nsys = X.shape[0]
    rec, stride, shared = host_records(family, p, nsys, real)
    cm, co = method.to_c(), options.to_c()
    fam = _family_struct(family, shared, real)
    cs, keep = prob.to_c(family.n, b, nsys, dtype=real)
    S = np.zeros(nsys, dtype=_status_dtype_f32() if f32 else _status_dtype())
    o = None if order is None else np.ascontiguousarray(np.asarray(order, dtype=np.int64))
    entry = lib().orc_solve_batch_stationarity_f32 if f32 else lib().orc_solve_batch_stationarity_f64
    rc = entry(
        ctypes.byref(cm), ctypes.byref(fam), ctypes.byref(cs), nsys, _ptr(X), _ptr(rec), stride, _ptr(S), None,
        None if o is None else _ptr(o), ctypes.byref(co), nthreads)
    del keep
    if rc != 0:
        raise RuntimeError(f"oracle failed with {rc}")
    return X, S


def hh_action_eval(h: float, lam: float, x):
    """∇S and ∇²S of the Hénon–Heiles two-step discrete action at x (6 entries)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    g, H = np.zeros(6), np.zeros((6, 6))
    lib().orc_hh_action_eval(h, lam, _ptr(x), _ptr(g), _ptr(H))
    return g, H


def _status_dtype():
    return np.dtype({
        "names": ["code", "f_calls", "iterations", "fnorm", "stepnorm", "step_failures", "promoted",
                  "stationary", "jvp_calls", "rank", "nullnorm", "gradratio", "rate", "contracted", "reused"],
        "formats": ["u1", "<i2", "<i4", "<f8", "<f8", "<i4", "u1", "u1", "<i2", "<i4", "<f8", "<f8", "<f8", "u1", "u1"],
        "offsets": [0, 2, 4, 8, 16, 24, 28, 29, 30, 32, 40, 48, 56, 64, 65],
        "itemsize": 72,
    })


def _status_dtype_f32():
    return np.dtype({
        "names": ["code", "f_calls", "iterations", "fnorm", "stepnorm", "step_failures", "promoted",
                  "stationary", "jvp_calls", "rank", "nullnorm", "gradratio", "rate", "contracted", "reused"],
        "formats": ["u1", "<i2", "<i4", "<f4", "<f4", "<i4", "u1", "u1", "<i2", "<i4", "<f4", "<f4", "<f4", "u1", "u1"],
        "offsets": [0, 2, 4, 8, 12, 16, 20, 21, 22, 24, 28, 32, 36, 40, 41],
        "itemsize": 44,
    })


class _Fam(ctypes.Structure):
    _fields_ = [("id", ctypes.c_int32), ("n", ctypes.c_int32), ("m", ctypes.c_int32),
                ("p_len", ctypes.c_int32), ("shared", ctypes.c_double * 8), ("data", ctypes.c_void_p)]


def _family_struct(family, shared, real=np.float64) -> _Fam:
    f = _Fam()
    f.id, f.n, f.m, f.p_len = family.id, family.n, family.n, family.p_len
    for k, v in enumerate(shared):
        f.shared[k] = v
    if hasattr(family, "host_data"):
        data = np.ascontiguousarray(family.host_data(), dtype=real)
        f.data = data.ctypes.data
        f._keep = data  # the struct points into it for the call
    return f


CODE_NAMES = ("SUCCESS", "STALLED", "MAXITERS", "SINGULAR", "NONFINITE", "LINESEARCH_FAILED")


def standard_cell(problem: int, scale: int, column: int, eltype: int):
    """One cell of the standard set: (code name, iterations, f_evals, fnorm, x)."""
    n = lib().orc_standard_u0(problem, eltype, None)
    code, it, nf, fn = ctypes.c_int(), ctypes.c_int(), ctypes.c_long(), ctypes.c_double()
    x = np.zeros(n)
    lib().orc_standard_cell(problem, scale, column, eltype, ctypes.byref(code), ctypes.byref(it),
                            ctypes.byref(nf), ctypes.byref(fn), _ptr(x))
    return CODE_NAMES[code.value], it.value, nf.value, fn.value, x


def lu_factor(J: np.ndarray, f32: bool = False):
    """factorize!! StaticLU of a square matrix (n ≤ 8): (factors, ipiv (1-based), info, e)."""
    J = np.asfortranarray(J, dtype=np.float64)
    n = J.shape[0]
    fac = np.zeros(n * n)
    ipiv = np.zeros(n, dtype=np.int32)
    info, e = ctypes.c_int32(), ctypes.c_int32()
    rc = lib().orc_lu_factor(n, _ptr(J.reshape(-1, order="F").copy()), int(f32), _ptr(fac), _ptr(ipiv),
                             ctypes.byref(info), ctypes.byref(e))
    assert rc == 0
    return fac.reshape(n, n, order="F"), ipiv, info.value, e.value


def lu_solve(J: np.ndarray, r: np.ndarray, f32: bool = False) -> np.ndarray:
    J = np.asarray(J, dtype=np.float64)
    n = J.shape[0]
    x = np.zeros(n)
    flat = np.ascontiguousarray(J.reshape(-1, order="F"))
    rr = np.ascontiguousarray(r, dtype=np.float64)
    assert lib().orc_lu_solve(n, _ptr(flat), _ptr(rr), int(f32), _ptr(x)) == 0
    return x


def rnorm(a) -> float:
    a = np.ascontiguousarray(a, dtype=np.float64)
    return lib().orc_rnorm(_ptr(a), len(a))


def solve_fixture(fixture: int, method, options, x, p=(0.0, 0.0), counts: bool = False, f32: bool = False):
    """One solve of a fixture of oracle/families.hpp (101 y.^2 − q, 102 2y − q, 103 x² − 4 with J = Inf,
    104 f_floor_r1, 105 the cubic of Check 1, 106 a x with J = b, 107 x with the jumping Jacobian,
    108 FloorResidual, 109 CycleStep, 110 JumpStep; 9 the quadratic-map family); with `counts`, also
    (residual calls, Jacobian calls). `f32` solves in Float32 (x, p and the status pass as doubles)."""
    x = np.ascontiguousarray(np.atleast_1d(np.asarray(x, dtype=np.float64))).copy()
    pp = np.ascontiguousarray(np.atleast_1d(np.asarray(p, dtype=np.float64)))
    S = np.zeros(1, dtype=_status_dtype())
    c = np.zeros(2, dtype=np.int64)
    cm, co = method.to_c(), options.to_c()
    rc = lib().orc_solve_fixture_t(fixture, int(f32), ctypes.byref(cm), ctypes.byref(co), _ptr(x), _ptr(pp), _ptr(S),
                                   _ptr(c))
    if rc != 0:
        raise RuntimeError(f"oracle fixture failed with {rc}")
    if counts:
        return x, S[0], (int(c[0]), int(c[1]))
    return x, S[0]


def fixture_eval(fixture: int, x, p=(0.0, 0.0), f32: bool = False):
    """(F(x), J(x)) of a fixture in Float64 (or Float32, returned as doubles)."""
    x = np.ascontiguousarray(np.atleast_1d(np.asarray(x, dtype=np.float64)))
    pp = np.ascontiguousarray(np.atleast_1d(np.asarray(p, dtype=np.float64)))
    n = len(x)
    F = np.zeros(n)
    J = np.zeros(n * n)
    assert lib().orc_fixture_eval(fixture, int(f32), _ptr(pp), _ptr(x), _ptr(F), _ptr(J)) == 0
    return F, J.reshape(n, n, order="F")


def fixture_step(fixture: int, method, x, p=(0.0, 0.0)) -> dict:
    """One step!! of `method` from x on a fixture, after start!!: x₊ and the step's record."""
    x = np.ascontiguousarray(np.atleast_1d(np.asarray(x, dtype=np.float64))).copy()
    pp = np.ascontiguousarray(np.atleast_1d(np.asarray(p, dtype=np.float64)))
    out = np.zeros(8)
    cm = method.to_c()
    assert lib().orc_fixture_step(fixture, ctypes.byref(cm), _ptr(x), _ptr(pp), _ptr(out)) == 0
    return dict(x=x, code=int(out[0]), failures=int(out[1]), promoted=bool(out[2]), info_promoted=bool(out[3]),
                fnorm=out[4], stepnorm=out[5], radius=out[6], f_calls=int(out[7]))


def refine_fixture(fixture: int, p, x, d, fac: str = "f32", refinement=(1.4901161193847656e-08, 5), zero_fx=False,
                   fnorm=None):
    """refine!! on a fixture at x from the direction d: `fac` is "f32" or "f64" (J(x) factorised in
    that precision) or "bad" (the Float32 matrix −I); `refinement` = (η, maxiter) or None for
    NoRefinement(); `fnorm` the caller's ‖F(x)‖ or None. Returns (d, failed)."""
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    pp = np.ascontiguousarray(np.atleast_1d(np.asarray(p, dtype=np.float64)))
    dd = np.ascontiguousarray(np.asarray(d, dtype=np.float64)).copy()
    kind = {"f32": 0, "f64": 1, "bad": 2}[fac]
    eta, maxiter = refinement if refinement is not None else (0.0, 0)
    failed = ctypes.c_int()
    assert lib().orc_refine_fixture(fixture, _ptr(pp), _ptr(x), _ptr(dd), kind, int(refinement is not None),
                                    float(eta), int(maxiter), int(zero_fx), int(fnorm is not None),
                                    float(fnorm) if fnorm is not None else 0.0, ctypes.byref(failed)) == 0
    return dd, bool(failed.value)


def read_mixed_stall(promoted: bool, slow_steps: int, fnorm: float, stepnorm: float, fnorm_old: float,
                     lin_rel: float, failures: int = 0) -> dict:
    """read_mixed_stall on a state and a SUCCESS step: the state and step it returns."""
    out = np.zeros(6)
    lib().orc_read_mixed_stall(int(promoted), int(slow_steps), float(fnorm), float(stepnorm), int(failures),
                               float(fnorm_old), float(lin_rel), _ptr(out))
    return dict(promoted=bool(out[0]), slow_steps=int(out[1]), failures=int(out[2]), info_promoted=bool(out[3]),
                fnorm=out[4], stepnorm=out[5])


def rmaxnorm(a, f32: bool = False) -> float:
    """rmaxnorm of a vector of length ≤ 4 in Float64 (or Float32)."""
    a = np.ascontiguousarray(np.atleast_1d(np.asarray(a, dtype=np.float64)))
    return lib().orc_rmaxnorm(_ptr(a), len(a), int(f32))


def solve_expsum_batch(method, options, X8, K, m, y_offset, Y, nthreads: int = 0):
    """The exponential-sum fits (LevenbergMarquardt), X8 an (nfit, 8) array (first 2K slots used)."""
    X8 = np.ascontiguousarray(X8, dtype=np.float64).copy()
    K = np.ascontiguousarray(K, dtype=np.int32)
    m = np.ascontiguousarray(m, dtype=np.int32)
    off = np.ascontiguousarray(y_offset, dtype=np.int64)
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    S = np.zeros(len(X8), dtype=_status_dtype())
    cm, co = method.to_c(), options.to_c()
    rc = lib().orc_solve_expsum_batch(ctypes.byref(cm), ctypes.byref(co), len(X8), _ptr(X8), _ptr(K), _ptr(m),
                                      _ptr(off), _ptr(Y), _ptr(S), nthreads)
    if rc != 0:
        raise RuntimeError(f"oracle failed with {rc}")
    return X8, S


def solve_mgh(problem: int, method, options, x0):
    """One Moré–Garbow–Hillstrom least-squares problem (11, 12, 8, 6): (x, status, f_evals)."""
    x = np.ascontiguousarray(x0, dtype=np.float64).copy()
    S = np.zeros(1, dtype=_status_dtype())
    nf = ctypes.c_long()
    cm, co = method.to_c(), options.to_c()
    assert lib().orc_solve_mgh(problem, ctypes.byref(cm), ctypes.byref(co), _ptr(x), _ptr(S), ctypes.byref(nf)) == 0
    return x, S[0], nf.value


def solve_linear_lsq(A, b, method, options, x0):
    """Newton + LevenbergMarquardt on F(x) = A x − b (A m × n, n ≤ 4) from x0: (x, status)."""
    A = np.asarray(A, dtype=np.float64)
    m, n = A.shape
    Ac = np.ascontiguousarray(A.T).reshape(-1)  # column-major
    bb = np.ascontiguousarray(b, dtype=np.float64)
    x = np.array(x0, dtype=np.float64)
    S = np.zeros(1, dtype=_status_dtype())
    cm, co = method.to_c(), options.to_c()
    assert lib().orc_solve_linear_lsq(n, m, _ptr(Ac), _ptr(bb), ctypes.byref(cm), ctypes.byref(co), _ptr(x),
                                      _ptr(S)) == 0
    return x, S[0]


def lm_damping_diagonal(J, diagonal: bool = True):
    """lm_damping_diagonal (levenbergmarquardt.jl:318-347) of an m × n matrix (n ≤ 4)."""
    J = np.asarray(J, dtype=np.float64)
    m, n = J.shape
    Jc = np.ascontiguousarray(J.T).reshape(-1)
    out = np.empty(n)
    assert lib().orc_lm_damping_diagonal(n, m, _ptr(Jc), 1 if diagonal else 0, _ptr(out)) == 0
    return out


def fpu_gl3_eval(h: float, beta: float, qp, z):
    """(F, J) of the FPU-β Gauss–Legendre s = 3 stage system (families.hpp FpuGl3)."""
    qp = np.ascontiguousarray(qp, dtype=np.float64)
    z = np.ascontiguousarray(z, dtype=np.float64)
    r = np.empty(48)
    J = np.empty(48 * 48)
    lib().orc_fpu_gl3_eval(h, beta, _ptr(qp), _ptr(z), _ptr(r), _ptr(J))
    return r, J.reshape(48, 48, order="F")


def exp_family(x):
    """The exponential of the exponential-sum family (a fixed algorithm, see lsq_capi.cpp)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    lib().orc_exp_family(_ptr(x), _ptr(out), x.size)
    return out


def givens(f: float, g: float):
    out = np.zeros(3)
    lib().orc_givens(f, g, _ptr(out))
    return tuple(out)


def max_threads() -> int:
    return lib().orc_max_threads()


def dogleg_pw_step(kind: int, method, p0: float = 0.0, p1: float = 0.0, x0: float = 0.0) -> dict:
    """One DogLeg step from x0 on the one-dimensional residual `kind` of the reference's DogLeg tests
    (0: p0 at 0 and p1 elsewhere; 1: −10 at 0, −10.5 at x ≥ 5, x − 10 otherwise; 2: x − 10), J = 1."""
    out = np.zeros(7)
    cm = method.to_c()
    rc = lib().orc_dogleg_pw_step(kind, ctypes.byref(cm), float(p0), float(p1), float(x0), _ptr(out))
    if rc != 0:
        raise RuntimeError(f"oracle DogLeg step failed with {rc}")
    return dict(x=out[0], radius=out[1], code=int(out[2]), failures=int(out[3]), stepnorm=out[4], f_calls=int(out[5]),
                limited=bool(out[6]))


def pow_int(x, n: int):
    """Julia's Float64 x^n for an integer n (Base.Math.pow_body), elementwise."""
    x = np.ascontiguousarray(np.atleast_1d(np.asarray(x, dtype=np.float64)))
    out = np.empty_like(x)
    lib().orc_pow_int(_ptr(x), int(n), _ptr(out), x.size)
    return out


def fsincos(x, f32: bool = False):
    """The sine and cosine of the dense family (families.hpp fsincos), elementwise."""
    x = np.ascontiguousarray(np.atleast_1d(np.asarray(x, dtype=np.float64)))
    sn = np.empty_like(x)
    cs = np.empty_like(x)
    lib().orc_fsincos(_ptr(x), _ptr(sn), _ptr(cs), x.size, int(f32))
    return sn, cs


def dense_eval(A, b, x):
    """(F(x), J(x)) of the dense family at one system: A (n × n), b and x (n)."""
    A = np.asarray(A, dtype=np.float64)
    n = A.shape[0]
    Ac = np.ascontiguousarray(A.T).reshape(-1)
    bb = np.ascontiguousarray(b, dtype=np.float64)
    xx = np.ascontiguousarray(x, dtype=np.float64)
    F = np.zeros(n)
    J = np.zeros(n * n)
    assert lib().orc_dense_eval(n, _ptr(Ac), _ptr(bb), _ptr(xx), _ptr(F), _ptr(J)) == 0
    return F, J.reshape(n, n, order="F")
