"""The batched nonlinear solve of GeometricSolvers.jl on MI355X (gfx950).

This package is the host side of libgsolve_hip for Python callers. It mirrors the reference's R2
interface (src/nonlinear/batched.jl:1-332): the same names, argument meanings, validation order and
error messages, with `solve_batch` standing for Julia's `solve_batch!` (it overwrites `X` in place and
returns `S`, or the six outcome counts). Device memory and streams come from PyTorch; the solve is
always the HIP kernel: a missing library raises, there is no CPU fallback.

    import torch
    gs = load_package()                          # see __graft_entry__.load_package
    prob = gs.NonlinearProblem(gs.HenonHeilesMidpoint())
    X = Xk.clone()                                # (nsys, 4) float64 on cuda
    S = gs.status_buffer(len(X), X.device)
    gs.solve_batch(X, S, prob, gs.Newton(), gs.BatchParams(0.1, Xk), f_abstol=1e-12, f_reltol=0.0)
    st = gs.statuses(S)                           # numpy structured array, SolverStatus{Float64}
"""
from __future__ import annotations

import ctypes
import enum
import math
from dataclasses import dataclass, field
from typing import Any, Optional

import numpy as np

from . import _abi
from ._abi import STATUS_DTYPE, load_library

__all__ = [
    "ArgumentError", "DimensionMismatch", "ReturnCode", "SUCCESS", "STALLED", "MAXITERS",
    "SINGULAR", "NONFINITE", "LINESEARCH_FAILED", "STATUS_DTYPE", "L2Norm", "MaxNorm", "Options",
    "LUFactorization", "Static", "Backtracking", "NoRefinement", "Refinement", "Newton",
    "NonlinearProblem", "Coupled2", "Hetero2", "ScalarSquare", "TanhLinear3",
    "HenonHeilesMidpoint", "PerSystem", "BatchParams", "parameters", "solve_batch",
    "status_buffer", "statuses", "isconverged", "batch_summary", "BatchSummary", "load_library",
]

_SQRT_EPS = math.sqrt(np.finfo(np.float64).eps)
_EPS = float(np.finfo(np.float64).eps)


class ArgumentError(ValueError):
    """Julia's ArgumentError."""


class DimensionMismatch(ValueError):
    """Julia's DimensionMismatch."""


class ReturnCode(enum.IntEnum):
    """ReturnCode (src/base/returncodes.jl:14), one byte."""

    SUCCESS = 0
    STALLED = 1
    MAXITERS = 2
    SINGULAR = 3
    NONFINITE = 4
    LINESEARCH_FAILED = 5


SUCCESS, STALLED, MAXITERS, SINGULAR, NONFINITE, LINESEARCH_FAILED = list(ReturnCode)


# --------------------------------------------------------------------------------------------
# Options (src/base/options.jl:8-236)
# --------------------------------------------------------------------------------------------


class L2Norm:
    """The Euclidean StoppingNorm (options.jl:16)."""

    def __eq__(self, other):
        return isinstance(other, L2Norm)


class MaxNorm:
    """The max StoppingNorm (options.jl:25)."""

    def __eq__(self, other):
        return isinstance(other, MaxNorm)


def _nonnegative(name: str, tol: float) -> None:
    if not tol >= 0:
        raise ArgumentError(f"{name} must be non-negative, got {tol}")


@dataclass(frozen=True)
class Options:
    """Options{Float64} (options.jl:58-71). Build with `Options.make(**kw)` or `Options()`."""

    f_abstol: float = 0.0
    f_reltol: float = _SQRT_EPS
    x_abstol: float = 0.0
    x_reltol: float = 2 * _EPS
    min_iterations: int = 0
    max_iterations: int = 100
    max_stalls: int = 2
    g_tol: float = _SQRT_EPS
    norm: Any = field(default_factory=L2Norm)
    kappa: float = 0.1
    roundoff: bool = False
    rate_test: bool = True

    def __post_init__(self):
        # Options(T; kwargs...) validation (options.jl:220-235)
        for name in ("f_abstol", "f_reltol", "x_abstol", "x_reltol", "g_tol"):
            _nonnegative(name, getattr(self, name))
        if self.min_iterations < 0:
            raise ArgumentError(f"min_iterations must be non-negative, got {self.min_iterations}")
        if self.max_iterations < self.min_iterations:
            raise ArgumentError(
                f"max_iterations = {self.max_iterations} is below min_iterations = {self.min_iterations}")
        if self.max_stalls < 1:
            raise ArgumentError(f"max_stalls must be at least 1, got {self.max_stalls}")
        # check_kappa (options.jl:105-109)
        if not (0.0 < self.kappa < 1.0):
            raise ArgumentError("kappa must be finite and in (0, 1)")
        if not isinstance(self.norm, (L2Norm, MaxNorm)):
            raise ArgumentError("norm must be L2Norm() or MaxNorm()")

    def to_c(self) -> _abi.GsOptionsF64:
        c = self.__dict__.get("_c")
        if c is not None:
            return c
        o = _abi.GsOptionsF64()
        o.f_abstol, o.f_reltol = self.f_abstol, self.f_reltol
        o.x_abstol, o.x_reltol = self.x_abstol, self.x_reltol
        o.min_iterations, o.max_iterations = self.min_iterations, self.max_iterations
        o.max_stalls = self.max_stalls
        o.norm = _abi.GS_NORM_MAX if isinstance(self.norm, MaxNorm) else _abi.GS_NORM_L2
        o.g_tol, o.kappa = self.g_tol, self.kappa
        o.roundoff, o.rate_test = int(bool(self.roundoff)), int(bool(self.rate_test))
        object.__setattr__(self, "_c", o)  # the options are frozen: build the struct once
        return o


def resolve_options(options: Optional[Options], kwargs: dict) -> Options:
    """resolve_options (src/nonlinear/solver.jl:391-398): keywords or an Options, never both."""
    if options is None:
        return Options(**kwargs)
    if kwargs:
        raise ArgumentError("pass either `options` or keyword arguments, not both")
    if not isinstance(options, Options):
        raise ArgumentError("options must be an Options")
    return options


# --------------------------------------------------------------------------------------------
# Methods (src/nonlinear/newton.jl:138-250, src/globalization/*.jl, src/nonlinear/refinement.jl)
# --------------------------------------------------------------------------------------------


@dataclass(frozen=True)
class LUFactorization:
    """LUFactorization(TF) (src/linear/methods.jl:37-47): TF None = working precision, or 'float32'."""

    precision: Optional[str] = None

    def __post_init__(self):
        if self.precision not in (None, "float32", np.float32):
            raise ArgumentError("LUFactorization takes no precision or Float32 on this device")


@dataclass(frozen=True)
class Static:
    """Static(α = 1.0) (src/globalization/static.jl:8-30)."""

    alpha: float = 1.0

    def __post_init__(self):
        if not (self.alpha > 0 and math.isfinite(self.alpha)):
            raise ArgumentError("the step of Static must be positive and finite")


@dataclass(frozen=True)
class Backtracking:
    """Backtracking(; c₁ = 1e-4, p = 0.5, maxiter = 100) (src/globalization/backtracking.jl:286-306)."""

    c1: float = 1e-4
    p: float = 0.5
    maxiter: int = 100

    def __post_init__(self):
        if not (0 < self.c1 < 1):
            raise ArgumentError("Backtracking needs 0 < c₁ < 1")
        if not (0 < self.p < 1):
            raise ArgumentError("Backtracking needs 0 < p < 1")
        if not (1 <= self.maxiter <= 2**31 - 1):
            raise ArgumentError("Backtracking needs 1 ≤ maxiter ≤ typemax(Int32)")


@dataclass(frozen=True)
class NoRefinement:
    """NoRefinement() (src/nonlinear/solver.jl:7)."""


@dataclass(frozen=True)
class Refinement:
    """Refinement(; η = √eps, maxiter = 5) (src/nonlinear/refinement.jl:27-49)."""

    eta: float = _SQRT_EPS
    maxiter: int = 5

    def __post_init__(self):
        if self.maxiter < 1:
            raise ArgumentError("Refinement maxiter must be ≥ 1")
        if not (math.isfinite(self.eta) and 0 < self.eta < 1):
            raise ArgumentError("Refinement η must be finite and in (0, 1)")


@dataclass(frozen=True)
class Newton:
    """Newton(; linsolve, globalization, refinement, regularization, best) (newton.jl:138-250).

    `reuse`, `keep_factor` and a `jacobian` selector have no device form: the families carry their
    analytic Jacobians, and a factor kept across steps is not part of this kernel.
    """

    linsolve: LUFactorization = field(default_factory=LUFactorization)
    globalization: Any = field(default_factory=Static)
    refinement: Any = field(default_factory=NoRefinement)
    regularization: float = 0.0
    best: bool = False
    reuse: Optional[int] = None
    keep_factor: bool = False

    def __post_init__(self):
        if isinstance(self.refinement, bool):
            object.__setattr__(self, "refinement", Refinement() if self.refinement else NoRefinement())
        if not (self.regularization >= 0 and math.isfinite(self.regularization)):
            raise ArgumentError("Newton regularization must be finite and ≥ 0")
        if self.keep_factor and self.reuse is None:
            raise ArgumentError(
                "Newton keep_factor = true needs reuse = k: only a factor that its steps reuse can survive a solve")
        if self.reuse is not None and self.reuse < 1:
            raise ArgumentError("Newton reuse must be ≥ 1 or nothing")
        if not isinstance(self.linsolve, LUFactorization):
            raise ArgumentError("the device solve takes LUFactorization() or LUFactorization(Float32)")
        if not isinstance(self.globalization, (Static, Backtracking)):
            raise ArgumentError("the device solve takes Static() or Backtracking() with LU")
        if not isinstance(self.refinement, (NoRefinement, Refinement)):
            raise ArgumentError("refinement must be NoRefinement() or Refinement()")

    def to_c(self) -> _abi.GsMethod:
        if self.reuse is not None:
            raise ArgumentError("Newton(; reuse = k) has no device form in libgsolve_hip")
        c = self.__dict__.get("_c")
        if c is not None:
            return c
        m = _abi.GsMethod()
        m.linsolve = _abi.GS_LU
        m.factor_precision = _abi.GS_FACTOR_F32 if self.linsolve.precision is not None else _abi.GS_FACTOR_WORKING
        g = self.globalization
        m.globalization = _abi.GS_BACKTRACKING if isinstance(g, Backtracking) else _abi.GS_STATIC
        m.static_alpha = g.alpha if isinstance(g, Static) else 1.0
        bt = g if isinstance(g, Backtracking) else Backtracking()
        m.bt_c1, m.bt_p, m.bt_maxiter = bt.c1, bt.p, bt.maxiter
        r = self.refinement
        m.refinement = _abi.GS_REFINEMENT if isinstance(r, Refinement) else _abi.GS_NO_REFINEMENT
        rr = r if isinstance(r, Refinement) else Refinement()
        m.ref_eta, m.ref_maxiter = rr.eta, rr.maxiter
        m.regularization = self.regularization
        m.best = int(bool(self.best))
        object.__setattr__(self, "_c", m)  # the method is frozen: build the struct once
        return m


# --------------------------------------------------------------------------------------------
# Problems and parameters (src/nonlinear/problem.jl:81-110, batched.jl:1-47)
# --------------------------------------------------------------------------------------------


class PerSystem:
    """PerSystem(data) (batched.jl:8-10): system i receives data[i]."""

    def __init__(self, data):
        self.data = data


class BatchParams:
    """BatchParams(shared, per_system) (batched.jl:34-37): system i receives (shared, per_system[i])."""

    def __init__(self, shared, per_system):
        self.shared = shared
        self.per_system = per_system


def parameters(p, i: int):
    """parameters(p, i) (batched.jl:20-21, 47), with a 0-based index i."""
    if isinstance(p, PerSystem):
        return p.data[i]
    if isinstance(p, BatchParams):
        return (p.shared, p.per_system[i])
    return p


class Family:
    """A registered problem family: the residual and Jacobian that the device code implements."""

    id: int = 0
    n: int = 0
    p_len: int = 0
    p_dtype = None  # torch dtype of a parameter record

    def marshal(self, p, nsys: int, device):
        """The parameter records of a batch: (tensor, byte stride, shared scalars)."""
        raise NotImplementedError


def _torch():
    import torch

    return torch


def _records(data, nsys: int, p_len: int, dtype, device, what: str):
    torch = _torch()
    if (isinstance(data, torch.Tensor) and data.device == device and data.dtype == dtype
            and data.dim() == 2 and data.shape[1] == p_len and data.is_contiguous()):
        return data
    t = data if isinstance(data, torch.Tensor) else torch.as_tensor(np.asarray(data))
    t = t.to(device=device, dtype=dtype).reshape(len(t), -1) if t.dim() > 0 else t
    if t.dim() != 2 or t.shape[1] != p_len:
        raise ArgumentError(f"{what} records must have {p_len} entries each")
    return t.contiguous()


class _SimpleFamily(Family):
    """A family whose parameter is PerSystem(records) or one shared record."""

    def marshal(self, p, nsys, device):
        torch = _torch()
        if isinstance(p, PerSystem):
            t = _records(p.data, nsys, self.p_len, self.p_dtype, device, type(self).__name__)
            return t, t.stride(0) * t.element_size(), ()
        if isinstance(p, BatchParams):
            raise ArgumentError(f"{type(self).__name__} takes PerSystem or shared parameters")
        t = torch.as_tensor(np.asarray(p, dtype=np.int64 if self.p_dtype == torch.int64 else np.float64))
        t = t.reshape(1, -1).to(device=device, dtype=self.p_dtype).contiguous()
        if t.shape[1] != self.p_len:
            raise ArgumentError(f"{type(self).__name__} takes {self.p_len} parameters per system")
        return t, 0, ()


class Coupled2(_SimpleFamily):
    """x + y²/8 − p (test/nonlinear/batched.jl:92-105), n = 2, p = (p₁, p₂)."""

    id, n, p_len = _abi.GS_FAM_COUPLED2, 2, 2

    @property
    def p_dtype(self):
        return _torch().float64


class Hetero2(_SimpleFamily):
    """The six-outcome fixture of the reference's batched tests (batched.jl:54-90), p = an Int code."""

    id, n, p_len = _abi.GS_FAM_HETERO2, 2, 1

    @property
    def p_dtype(self):
        return _torch().int64


class ScalarSquare(_SimpleFamily):
    """x² − p (batched.jl:107-110), n = 1."""

    id, n, p_len = _abi.GS_FAM_SCALAR_SQUARE, 1, 1

    @property
    def p_dtype(self):
        return _torch().float64


class TanhLinear3(_SimpleFamily):
    """A_p x + 0.1 tanh.(x) − b_p, n = 3; a record is vec(A_p) (column-major, 9) then b_p (3)."""

    id, n, p_len = _abi.GS_FAM_TANH_LINEAR3, 3, 12

    @property
    def p_dtype(self):
        return _torch().float64


class HenonHeilesMidpoint(Family):
    """The implicit-midpoint step x₊ = x_k + h f((x_k + x₊)/2) of Hénon–Heiles, n = 4.

    Parameter: BatchParams(h, X_k) — the step size shared, the previous state x_k per system.
    """

    id, n, p_len = _abi.GS_FAM_HENON_HEILES_MIDPOINT, 4, 4

    def marshal(self, p, nsys, device):
        torch = _torch()
        if not isinstance(p, BatchParams):
            raise ArgumentError("HenonHeilesMidpoint takes BatchParams(h, X_k)")
        h = float(p.shared)
        t = _records(p.per_system, nsys, 4, torch.float64, device, "HenonHeilesMidpoint")
        return t, t.stride(0) * t.element_size(), (h,)


class NonlinearProblem:
    """NonlinearProblem(F) (src/nonlinear/problem.jl:81-110) over a registered device family."""

    def __init__(self, F: Family, J=None):
        if not isinstance(F, Family):
            raise ArgumentError("the device solve takes a registered family as the residual F")
        self.F = F
        self.J = J if J is not None else F  # each family carries its analytic Jacobian


# --------------------------------------------------------------------------------------------
# Status buffers and the batched solve (batched.jl:49-332)
# --------------------------------------------------------------------------------------------


def status_buffer(nsys: int, device):
    """A device vector of `nsys` SolverStatus{Float64} (72 bytes each), as a (nsys, 72) uint8 tensor."""
    return _torch().empty((nsys, STATUS_DTYPE.itemsize), dtype=_torch().uint8, device=device)


def statuses(S) -> np.ndarray:
    """The statuses of a status buffer as a host numpy array of dtype STATUS_DTYPE."""
    host = S.detach().to("cpu").contiguous().numpy()
    return host.reshape(-1).view(STATUS_DTYPE)


def isconverged(x) -> bool | np.ndarray:
    """isconverged (src/base/status.jl:276-277) of a status array or of a counts dict."""
    if isinstance(x, dict):
        return x["SUCCESS"] == sum(x.values())
    return np.asarray(x["code"]) == SUCCESS


_CODE_NAMES = ("SUCCESS", "STALLED", "MAXITERS", "SINGULAR", "NONFINITE", "LINESEARCH_FAILED")


def _stream_ptr(X):
    """The raw hipStream_t of PyTorch's current stream on the device of X."""
    torch = _torch()
    raw = getattr(torch._C, "_cuda_getCurrentRawStream", None)
    if raw is not None:
        return ctypes.c_void_p(raw(X.device.index if X.device.index is not None else torch.cuda.current_device()))
    return ctypes.c_void_p(torch.cuda.current_stream(X.device).cuda_stream)


def _check_x(X, nfam: int):
    torch = _torch()
    if not isinstance(X, torch.Tensor):
        raise ArgumentError("X must be a torch tensor of shape (nsys, n)")
    if X.dtype != torch.float64:
        raise ArgumentError(
            f"SolverStatus real type Float64 does not match batch real element type {X.dtype}")
    if X.dim() != 2 or X.shape[1] != nfam:
        raise DimensionMismatch(f"X must have shape (nsys, {nfam}), got {tuple(X.shape)}")
    if not X.is_contiguous():
        raise ArgumentError("X must be contiguous (a Vector{SVector{n,Float64}} is AoS)")


def _validate_batch(X, p, order):
    """_validate_batch (batched.jl:113-138), with 0-based or 1-based `order` accepted as Julia's 1:N."""
    nsys = X.shape[0]
    if isinstance(p, PerSystem) and len(p.data) != nsys:
        raise DimensionMismatch(
            f"PerSystem parameter length {len(p.data)} does not match batch length {nsys}")
    if isinstance(p, BatchParams) and len(p.per_system) != nsys:
        raise DimensionMismatch(
            f"BatchParams per-system length {len(p.per_system)} does not match batch length {nsys}")
    if order is None:
        return None
    o = np.asarray(order.detach().cpu().numpy() if hasattr(order, "detach") else order, dtype=np.int64).reshape(-1)
    if len(o) != nsys:
        raise DimensionMismatch(f"order length {len(o)} does not match batch length {nsys}")
    seen = np.zeros(nsys + 1, dtype=bool)
    ok = bool(np.all((o >= 1) & (o <= nsys)))
    if ok:
        seen[o] = True
        ok = bool(seen[1:].all())
    if not ok:
        raise ArgumentError(f"order is not a permutation of 1:{nsys}")
    return o


def _validate_workgroupsize(ws):
    """_validate_workgroupsize (batched.jl:205-210)."""
    if ws is None:
        return 0
    if not ws >= 1:
        raise ArgumentError(f"workgroupsize must be a positive integer or `nothing`, got {ws}")
    if ws > 1024:
        raise ArgumentError(f"workgroupsize {ws} exceeds the device maximum of 1024 work-items")
    return int(ws)


def solve_batch(X, *args, order=None, options: Optional[Options] = None, workgroupsize=None, **kwargs):
    """solve_batch!(X, S, prob, method, p = nothing[, options]; order, options, workgroupsize, kwargs...)
    solve_batch!(X, prob, method, p = nothing[, options]; ...)  (batched.jl:257-332)

    Solve every system of `X` (a (nsys, n) float64 HIP tensor, overwritten with the final iterates)
    in one kernel launch. The status form writes each system's SolverStatus into `S` (from
    `status_buffer`) and returns `S`; the counts form returns the six outcome counts as a dict.
    `order` is a permutation of 1:nsys (1-based, as Julia's): work-item i solves system order[i].
    The call is synchronous, as `_launch!` is (batched.jl:193).
    """
    if args and isinstance(args[0], NonlinearProblem):
        S = None
        rest = args
    else:
        if len(args) < 3:
            raise ArgumentError("solve_batch(X, S, prob, method, p = None) or solve_batch(X, prob, method, p = None)")
        S, rest = args[0], args[1:]
    prob, method = rest[0], rest[1]
    p = rest[2] if len(rest) > 2 else None
    if len(rest) > 3:
        if options is not None:
            raise ArgumentError("pass the options positionally or as a keyword, not both")
        options = rest[3]
    if len(rest) > 4:
        raise ArgumentError("too many positional arguments")
    if not isinstance(prob, NonlinearProblem):
        raise ArgumentError("prob must be a NonlinearProblem over a registered family")
    if not isinstance(method, Newton):
        raise ArgumentError("method must be a Newton")
    fam = prob.F
    _check_x(X, fam.n)
    order_h = _validate_batch(X, p, order)
    nsys = X.shape[0]
    torch = _torch()
    if S is not None:
        if S.dtype != torch.uint8 or S.dim() != 2 or S.shape[1] != STATUS_DTYPE.itemsize:
            raise ArgumentError("S must be a status buffer (status_buffer(nsys, device))")
        if S.shape[0] != nsys:
            raise DimensionMismatch(
                f"status vector length {S.shape[0]} does not match batch length {nsys}")
        if S.device != X.device:
            raise ArgumentError(
                f"status vector backend {S.device} does not match batch backend {X.device}")
        if not S.is_contiguous():
            raise ArgumentError("S must be contiguous")
    ws = _validate_workgroupsize(workgroupsize)
    opt = resolve_options(options, kwargs)
    cm = method.to_c()
    if X.device.type != "cuda":
        raise ArgumentError(
            "X must be a device array (a HIP tensor): libgsolve_hip solves on the GPU and has no host path")
    if nsys == 0:
        return S if S is not None else {k: 0 for k in _CODE_NAMES}
    lib = load_library()
    P, p_stride, shared = fam.marshal(p, nsys, X.device)
    cf = _abi.GsFamily()
    cf.id, cf.n, cf.m, cf.p_len = fam.id, fam.n, fam.n, fam.p_len
    for k, v in enumerate(shared):
        cf.shared[k] = v
    order_dev = None
    if order_h is not None:
        order_dev = torch.as_tensor(order_h, dtype=torch.int64).to(X.device)
    codes = None
    if S is None:
        codes = torch.empty(nsys, dtype=torch.uint8, device=X.device)
    copt = opt.to_c()
    rc = lib.gs_solve_batch_f64(
        ctypes.byref(cm), ctypes.byref(cf), nsys, ctypes.c_void_p(X.data_ptr()),
        ctypes.c_void_p(P.data_ptr()), p_stride,
        ctypes.c_void_p(S.data_ptr()) if S is not None else None,
        ctypes.c_void_p(codes.data_ptr()) if codes is not None else None,
        ctypes.c_void_p(order_dev.data_ptr()) if order_dev is not None else None,
        ctypes.byref(copt), ws, _stream_ptr(X))
    _raise_on(rc, lib)
    if S is not None:
        return S
    counts = (ctypes.c_int64 * 6)()
    _raise_on(lib.gs_tally_codes(ctypes.c_void_p(codes.data_ptr()), nsys, counts, _stream_ptr(X)), lib)
    return {k: int(counts[i]) for i, k in enumerate(_CODE_NAMES)}


def _raise_on(rc: int, lib) -> None:
    if rc == _abi.GS_OK:
        return
    msg = _abi.last_error(lib)
    if rc == _abi.GS_EDIM:
        raise DimensionMismatch(msg)
    if rc in (_abi.GS_EINVAL, _abi.GS_EFAMILY):
        raise ArgumentError(msg)
    raise RuntimeError(f"libgsolve_hip: {msg}")


@dataclass(frozen=True)
class BatchSummary:
    """BatchSummary{Float64} (batched.jl:63-69)."""

    fnorm: float
    all_converged: bool
    failure_count: int
    max_iterations: int
    total_iterations: int


def batch_summary(S) -> BatchSummary:
    """batch_summary(S) (batched.jl:80-89), reduced on the device."""
    nsys = S.shape[0]
    if nsys == 0:
        return BatchSummary(0.0, True, 0, 0, 0)
    lib = load_library()
    fn = ctypes.c_double()
    out = (ctypes.c_int64 * 3)()
    _raise_on(lib.gs_batch_summary_f64(ctypes.c_void_p(S.data_ptr()), nsys, ctypes.byref(fn), out,
                                       _stream_ptr(S)), lib)
    return BatchSummary(fn.value, out[0] == 0, int(out[0]), int(out[1]), int(out[2]))
