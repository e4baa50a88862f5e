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
import functools
import enum
import math
from dataclasses import dataclass, field
from typing import Any, Optional

import numpy as np

from . import _abi
from ._abi import STATUS_DTYPE, STATUS_DTYPE_F32, load_library

__all__ = [
    "ArgumentError", "DimensionMismatch", "ReturnCode", "SUCCESS", "STALLED", "MAXITERS",
    "SINGULAR", "NONFINITE", "LINESEARCH_FAILED", "STATUS_DTYPE", "STATUS_DTYPE_F32", "L2Norm", "MaxNorm", "Options",
    "LUFactorization", "Static", "Backtracking", "DogLeg", "LevenbergMarquardt", "NoRefinement", "Refinement", "Newton",
    "NonlinearProblem", "Deflated", "StationarityProblem", "HenonHeilesAction", "Reparametrized", "LinearChart",
    "PolarChart", "Coupled2", "Hetero2", "ScalarSquare", "TanhLinear3", "QuadMap2", "DenseSin",
    "HenonHeilesMidpoint", "FpuGl3", "PerSystem", "BatchParams", "parameters", "solve_batch",
    "LeastSquaresProblem", "ExpSumFits", "ExpSumData",
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
    """Options{R} (options.jl:58-71, 207-235) with R = Float64 or, with `eltype="float32"`, Float32:
    the defaults √eps(R), 2 eps(R), √eps(R) and R(1)/10 are R's, and every tolerance is held as its
    R value (R(x), as the constructor converts)."""

    f_abstol: float = 0.0
    f_reltol: Optional[float] = None
    x_abstol: float = 0.0
    x_reltol: Optional[float] = None
    min_iterations: int = 0
    max_iterations: int = 100
    max_stalls: int = 2
    g_tol: Optional[float] = None
    norm: Any = field(default_factory=L2Norm)
    kappa: Optional[float] = None
    roundoff: bool = False
    rate_test: bool = True
    eltype: str = "float64"

    def __post_init__(self):
        if self.eltype not in ("float64", "float32"):
            raise ArgumentError("Options eltype must be 'float64' or 'float32'")
        f32 = self.eltype == "float32"
        eps = float(np.finfo(np.float32).eps) if f32 else _EPS
        sqrt_eps = float(np.sqrt(np.float32(eps))) if f32 else _SQRT_EPS
        defaults = {"f_reltol": sqrt_eps, "x_reltol": 2 * eps, "g_tol": sqrt_eps,
                    "kappa": float(np.float32(1) / np.float32(10)) if f32 else 0.1}
        for name, v in defaults.items():
            if getattr(self, name) is None:
                object.__setattr__(self, name, v)
        if f32:
            for name in ("f_abstol", "f_reltol", "x_abstol", "x_reltol", "g_tol", "kappa"):
                object.__setattr__(self, name, float(np.float32(getattr(self, name))))
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


def resolve_options(options: Optional[Options], kwargs: dict, eltype: str = "float64") -> Options:
    """resolve_options (src/nonlinear/solver.jl:391-398): keywords or an Options, never both; an
    Options of another real type than the batch's is refused (check_options, options.jl:241-245)."""
    if options is None:
        return Options(eltype=eltype, **kwargs)
    if kwargs:
        raise ArgumentError("pass either `options` or keyword arguments, not both")
    if not isinstance(options, Options):
        raise ArgumentError("options must be an Options")
    if options.eltype != eltype:
        R, T = ("Float32" if e == "float32" else "Float64" for e in (options.eltype, eltype))
        raise ArgumentError(f"options are Options{{{R}}}, but the iterate has element type {T}: pass Options({T}; …)")
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
class DogLeg:
    """DogLeg(; radius = None) (src/globalization/dogleg.jl:1-70): the trust-region step of Procedure
    4.1 of Pawlowski, Simonis, Walker and Shadid (2008). `radius` is the initial radius, stored
    clamped to [1e-6, 1e10]; None starts it at the first Newton step's norm."""

    radius: Optional[float] = None

    def __post_init__(self):
        if self.radius is not None:
            r = float(self.radius)
            if not (r > 0 and math.isfinite(r)):
                raise ArgumentError("the radius of DogLeg must be positive and finite")
            object.__setattr__(self, "radius", min(max(r, 1e-6), 1e10))


@dataclass(frozen=True)
class LevenbergMarquardt:
    """LevenbergMarquardt(; μ₀, μ_min, μ_max, μ_decrease, μ_increase, ρ_accept, scaling)
    (src/globalization/levenbergmarquardt.jl:48-98) on the stacked Givens step; `scaling` is
    "none" (D = I) or "diagonal" (Marquardt's column norms of J, floored at √eps of the largest,
    levenbergmarquardt.jl:318-347)."""

    mu0: float = 1e-6
    mu_min: float = max(1e-15, 4 * _EPS)
    mu_max: float = 1e12
    mu_decrease: float = 1.0 / 3.0
    mu_increase: float = 4.0
    rho_accept: float = 0.0
    scaling: str = "none"

    def __post_init__(self):
        if not (all(math.isfinite(v) for v in (self.mu0, self.mu_min, self.mu_max))
                and 0 < self.mu_min <= self.mu0 <= self.mu_max):
            raise ArgumentError(
                "LevenbergMarquardt requires finite damping bounds 0 < μ_min ≤ μ₀ ≤ μ_max < Inf; "
                f"got μ_min = {self.mu_min}, μ₀ = {self.mu0}, μ_max = {self.mu_max}")
        if not (math.isfinite(self.mu_decrease) and 0 < self.mu_decrease < 1):
            raise ArgumentError(f"LevenbergMarquardt requires 0 < μ_decrease < 1; got μ_decrease = {self.mu_decrease}")
        if not (math.isfinite(self.mu_increase) and self.mu_increase > 1):
            raise ArgumentError(f"LevenbergMarquardt requires μ_increase > 1; got μ_increase = {self.mu_increase}")
        if not (math.isfinite(self.rho_accept) and 0 <= self.rho_accept < 1):
            raise ArgumentError(f"LevenbergMarquardt requires 0 ≤ ρ_accept < 1; got ρ_accept = {self.rho_accept}")
        if self.scaling not in ("none", "diagonal"):
            raise ArgumentError(
                f"LevenbergMarquardt requires scaling ∈ (:none, :diagonal); got scaling = :{self.scaling}")


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
    """Newton(; linsolve, globalization, refinement, regularization, best, reuse, keep_factor)
    (newton.jl:138-250).

    `reuse = k` holds the factor for up to k consecutive steps (newton.jl:1781-1874) in the
    lane-per-system kernels. `keep_factor` keeps the factor across the solves of one state; a batch
    solves each system once from a fresh state, so it changes no result here. A `jacobian` selector
    has no device form: the families carry their analytic Jacobians.
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
        if not isinstance(self.globalization, (Static, Backtracking, LevenbergMarquardt, DogLeg)):
            raise ArgumentError(
                "the device solve takes Static(), Backtracking(), DogLeg() or LevenbergMarquardt() with LU")
        if isinstance(self.globalization, DogLeg) and self.reuse is not None:
            raise ArgumentError("Newton with a DogLeg globalization takes no reuse: its model needs the "
                                "Jacobian of the current iterate")
        if not isinstance(self.refinement, (NoRefinement, Refinement)):
            raise ArgumentError("refinement must be NoRefinement() or Refinement()")
        if isinstance(self.globalization, LevenbergMarquardt):
            if self.reuse is not None:
                raise ArgumentError("Newton with a LevenbergMarquardt globalization takes no reuse")
            if isinstance(self.refinement, Refinement):
                raise ArgumentError(
                    "Newton with a LevenbergMarquardt globalization takes no refinement: a refinement corrects "
                    "the solve of J d = -F, which the damped step does not take; use refinement = NoRefinement()")

    def to_c(self) -> _abi.GsMethod:
        c = self.__dict__.get("_c")
        if c is not None:
            return c
        m = _abi.GsMethod()
        m.linsolve = _abi.GS_LU
        m.factor_precision = _abi.GS_FACTOR_F32 if self.linsolve.precision is not None else _abi.GS_FACTOR_WORKING
        g = self.globalization
        m.globalization = (_abi.GS_BACKTRACKING if isinstance(g, Backtracking) else
                           _abi.GS_LEVENBERG_MARQUARDT if isinstance(g, LevenbergMarquardt) else
                           _abi.GS_DOGLEG if isinstance(g, DogLeg) else _abi.GS_STATIC)
        m.dl_radius = float(g.radius) if isinstance(g, DogLeg) and g.radius is not None else 0.0
        lm = g if isinstance(g, LevenbergMarquardt) else LevenbergMarquardt()
        m.lm_mu0, m.lm_mu_min, m.lm_mu_max = lm.mu0, lm.mu_min, lm.mu_max
        m.lm_mu_decrease, m.lm_mu_increase, m.lm_rho_accept = lm.mu_decrease, lm.mu_increase, lm.rho_accept
        m.lm_scaling = _abi.GS_LM_SCALING_DIAGONAL if lm.scaling == "diagonal" else _abi.GS_LM_SCALING_NONE
        m.static_alpha = g.alpha if isinstance(g, Static) else 1.0
        bt = g if isinstance(g, Backtracking) else Backtracking()
        m.bt_c1, m.bt_p, m.bt_maxiter = bt.c1, bt.p, bt.maxiter
        r = self.refinement
        m.refinement = _abi.GS_REFINEMENT if isinstance(r, Refinement) else _abi.GS_NO_REFINEMENT
        rr = r if isinstance(r, Refinement) else Refinement()
        m.ref_eta, m.ref_maxiter = rr.eta, rr.maxiter
        m.regularization = self.regularization
        m.best = int(bool(self.best))
        m.reuse = int(self.reuse) if self.reuse is not None else 0
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

    def marshal(self, p, nsys: int, device, real=None):
        """The parameter records of a batch: (tensor, byte stride, shared scalars); `real` is the torch
        float type of the records (the batch's working precision, float64 by default)."""
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

    def marshal(self, p, nsys, device, real=None):
        torch = _torch()
        dtype = self.p_dtype if self.p_dtype == torch.int64 or real is None else real
        if isinstance(p, PerSystem):
            t = _records(p.data, nsys, self.p_len, dtype, device, type(self).__name__)
            return t, t.stride(0) * t.element_size(), ()
        if isinstance(p, BatchParams):
            raise ArgumentError(f"{type(self).__name__} takes PerSystem or shared parameters")
        t = torch.as_tensor(np.asarray(p, dtype=np.int64 if dtype == torch.int64 else np.float64))
        t = t.reshape(1, -1).to(device=device, dtype=dtype).contiguous()
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


class QuadMap2(_SimpleFamily):
    """u = A x, F_i = (u_i + c_i u_i²) − b_i, n = 2: the systems f_mod_r1 and f_ill_r1 of the reference's
    mixed-precision tests (test/nonlinear/refinement.jl:40-96), which take c = (0.1, −0.2); a record is
    vec(A) (column-major, 4), then b (2), then c (2)."""

    id, n, p_len = _abi.GS_FAM_QUAD_MAP2, 2, 8

    @property
    def p_dtype(self):
        return _torch().float64


# The orders of the dense family: up to 8 one lane per system, from 12 a wave segment per system.
DENSE_ORDERS = (2, 4, 6, 8, 12, 16, 24, 32, 48, 64)


class DenseSin(_SimpleFamily):
    """TenEResidual(A) of the reference's device benchmark (benchmark/device_problems.jl:133-175):
    F(x, b) = A x + sin.(x) ./ 10 − b, n in DENSE_ORDERS, with A (n × n) the family's own, shared by the
    batch, and p = PerSystem(B), one b per system. The Jacobian is A + Diagonal(cos.(x) ./ 10). Its sine
    and cosine are one fixed algorithm (Cody–Waite reduction, Taylor polynomials), the same on the
    device and in the oracle. A Float32 batch (the device grid's element type, devices.jl:47) reads A
    rounded to Float32, as tenE_matrix(n, Float32) holds it."""

    id = _abi.GS_FAM_DENSE_SIN

    def __init__(self, A):
        a = np.asarray(A.detach().cpu().numpy() if hasattr(A, "detach") else A, dtype=np.float64)
        if a.ndim != 2 or a.shape[0] != a.shape[1] or a.shape[0] not in DENSE_ORDERS:
            raise ArgumentError("DenseSin takes a square A of order 2, 4, 6, 8, 12, 16, 24, 32, 48 or 64")
        self.A = a
        self.n = self.p_len = a.shape[0]
        self._dev = {}

    @property
    def p_dtype(self):
        return _torch().float64

    def host_data(self) -> np.ndarray:
        """A column-major, as the library reads it."""
        return np.ascontiguousarray(self.A.T).reshape(-1)

    def device_data(self, device, dtype=None):
        """A on `device` (column-major) in `dtype` (torch.float64 unless given), kept per device and type."""
        torch = _torch()
        dtype = torch.float64 if dtype is None else dtype
        key = (str(device), str(dtype))
        t = self._dev.get(key)
        if t is None:
            t = torch.as_tensor(self.host_data(), device=device).to(dtype).contiguous()
            self._dev[key] = t
        return t


class HenonHeilesMidpoint(Family):
    """The implicit-midpoint step x₊ = x_k + h f((x_k + x₊)/2) of Hénon–Heiles, n = 4.

    Parameter: BatchParams(h, X_k) — the step size shared, the previous state x_k per system.
    """

    id, n, p_len = _abi.GS_FAM_HENON_HEILES_MIDPOINT, 4, 4

    def marshal(self, p, nsys, device, real=None):
        torch = _torch()
        if not isinstance(p, BatchParams):
            raise ArgumentError("HenonHeilesMidpoint takes BatchParams(h, X_k)")
        h = float(p.shared)
        t = _records(p.per_system, nsys, 4, real or torch.float64, device, "HenonHeilesMidpoint")
        return t, t.stride(0) * t.element_size(), (h,)


class FpuGl3(Family):
    """The Gauss–Legendre s = 3 stage system of a 16-site FPU-β chain with fixed ends, n = 48.

    Position-only (Nyström) form of the collocation equations, Q_i = q + h c_i p + h² Σ_j (A²)_ij g(Q_j),
    in the stage increments z_i = Q_i − q (the reference's StageResidual, src/nonlinear/problem.jl:
    183-189, stacks the 96 first-order increments of (q, p); this form eliminates the momenta).
    V(r) = r²/2 + β r⁴/4 per bond. Parameter: BatchParams(h, QP) with QP (nsys, 32) = (q, p) per system;
    the start is z = 0. One wave per system on the device, each lane holding one row of the factor.
    """

    id, n, p_len = _abi.GS_FAM_FPU_GL3, 48, 32

    def __init__(self, beta: float = 1.0):
        self.beta = float(beta)

    def marshal(self, p, nsys, device, real=None):
        torch = _torch()
        if not isinstance(p, BatchParams):
            raise ArgumentError("FpuGl3 takes BatchParams(h, QP) with QP the (q, p) record of every system")
        t = _records(p.per_system, nsys, 32, torch.float64, device, "FpuGl3")
        return t, t.stride(0) * t.element_size(), (float(p.shared), self.beta)


class HenonHeilesAction(_SimpleFamily):
    """The two-step discrete action of the variational midpoint rule for Hénon–Heiles, an action for
    StationarityProblem: on the full state x = (q₋, q, q₊) (6 entries),
    S(x) = L_d(q₋, q) + L_d(q, q₊), L_d(a, b) = |b − a|²/(2h) − h V((a + b)/2),
    V(q) = (q₁² + q₂²)/2 + λ (q₁² q₂ − q₂³/3). With E = (3, 4) and U = (5, 6) the stationarity
    equations are the discrete Euler–Lagrange equations for q₊. Parameter p = (h, λ), shared or
    PerSystem; ∇S and ∇²S are closed forms on the device."""

    id, n, p_len = _abi.GS_ACT_HENON_HEILES, 6, 2

    @property
    def p_dtype(self):
        return _torch().float64


def _index_set(idx, name):
    """_normalize_indices (stationarity.jl:60-72): None or ':' is Colon; a tuple of integers is kept."""
    if idx is None or (isinstance(idx, str) and idx == ":"):
        return None
    if isinstance(idx, (tuple, list)):
        if len(idx) == 0:
            raise ArgumentError("index set must be non-empty, got ()")
        if not all(isinstance(i, (int, np.integer)) and not isinstance(i, bool) for i in idx):
            raise ArgumentError(f"invalid index set type for {name}")
        return tuple(int(i) for i in idx)
    raise ArgumentError(f"invalid index set type {type(idx).__name__}")


class StationarityProblem:
    """StationarityProblem(S; E = :, U = :, b = nothing) (src/nonlinear/stationarity.jl:1-118): the
    residual (∇S(x, p))_E + b on the unknowns x[U] of each system's full state, with the Jacobian
    (∇²S)_{E,U}. `S` is a registered action (HenonHeilesAction); E and U are None (:) or tuples of
    1-based indices of one length; b is None, a number, a sequence of |E| numbers, or a PerSystem
    of those. A batch (solve_batch with this problem) overwrites the entries U of each row of X and
    keeps the others."""

    def __init__(self, S, E=None, U=None, b=None):
        if not isinstance(S, HenonHeilesAction):
            raise ArgumentError("the device StationarityProblem takes a registered action (HenonHeilesAction)")
        En, Un = _index_set(E, "E"), _index_set(U, "U")
        if (En is None) != (Un is None):
            raise ArgumentError("StationarityProblem requires both E and U to be :")
        if En is not None:
            if len(En) != len(Un):
                raise ArgumentError(f"length(E) == length(U) required, got {len(En)} and {len(Un)}")
            if not all(i >= 1 for i in En):
                raise ArgumentError(f"indices in E must be >= 1, got {En}")
            if not all(i >= 1 for i in Un):
                raise ArgumentError(f"indices in U must be >= 1, got {Un}")
            if len(set(En)) != len(En):
                raise ArgumentError("index set E has duplicate indices")
            if len(set(Un)) != len(Un):
                raise ArgumentError("index set U has duplicate indices")
        if callable(b) and not isinstance(b, PerSystem):
            raise ArgumentError("a callable datum b(p) has no device form: pass its values as PerSystem")
        self.S, self.E, self.U, self.b = S, En, Un, b

    def n_equations(self, M: int) -> int:
        return M if self.E is None else len(self.E)

    def to_c(self, M: int, b_kw, nsys: int, device=None, dtype=np.float64):
        """gs_stationarity for full states of M entries and the datum `b_kw` (or the problem's own):
        the checks of _validate_stationarity_solve / _validate_stationarity_batch
        (stationarity.jl:133-157, 600-618). Returns (struct, the datum array to keep alive)."""
        m = self.n_equations(M)
        for name, idx in (("E", self.E), ("U", self.U)):
            if idx is not None and not all(1 <= i <= M for i in idx):
                raise ArgumentError(f"index in {name} exceeds {M}")
        if callable(b_kw) and not isinstance(b_kw, PerSystem):
            raise ArgumentError("a callable datum b(p) has no device form: pass its values as PerSystem")
        bsrc = self.b if b_kw is None else b_kw
        c = _abi.GsStationarity()
        c.n_eq = 0 if self.E is None else len(self.E)
        for k in range(c.n_eq):
            c.e[k], c.u[k] = self.E[k], self.U[k]
        keep = None
        if bsrc is None:
            c.b_kind = _abi.GS_DATUM_NONE
        elif isinstance(bsrc, PerSystem):
            data = bsrc.data
            torch = _torch()
            if device is None:
                arr = np.ascontiguousarray(np.asarray(data.cpu() if isinstance(data, torch.Tensor) else data,
                                                      dtype=dtype).reshape(len(data), -1))
                ptr, stride, rows, width = arr.ctypes.data, arr.strides[0], arr.shape[0], arr.shape[1]
                keep = arr
            else:
                tdt = torch.float32 if np.dtype(dtype) == np.float32 else torch.float64
                t = data if isinstance(data, torch.Tensor) else torch.as_tensor(np.asarray(data, dtype=dtype))
                t = t.to(device=device, dtype=tdt)
                t = (t.reshape(len(t), -1) if t.dim() > 0 else t).contiguous()
                ptr, stride, rows, width = t.data_ptr(), t.stride(0) * t.element_size(), t.shape[0], t.shape[1]
                keep = t
            if rows != nsys:
                raise DimensionMismatch(f"PerSystem datum has {rows} entries, expected {nsys} to match the batch")
            if width != m:
                raise ArgumentError(f"datum b must have length {m}, got {width}")
            c.b_kind, c.b_data, c.b_stride = _abi.GS_DATUM_PER_SYSTEM, ptr, stride
        else:
            v = np.atleast_1d(np.asarray(bsrc, dtype=np.float64))
            if np.ndim(bsrc) == 0:
                v = np.full(m, float(bsrc))
            elif v.size != m:
                raise ArgumentError(f"datum b must have length {m}, got {v.size}")
            c.b_kind = _abi.GS_DATUM_SHARED
            for k in range(m):
                c.b_shared[k] = float(v[k])
        return c, keep


class LinearChart:
    """The chart x = B y + c (B an n × n matrix, c an n-vector, shared by the batch), ∂g = B."""

    def __init__(self, B, c=None):
        B = np.atleast_2d(np.asarray(B, dtype=np.float64))
        if B.ndim != 2 or B.shape[0] != B.shape[1] or B.shape[0] > 4:
            raise ArgumentError("a LinearChart takes a square B of order 1 to 4")
        c = np.zeros(B.shape[0]) if c is None else np.atleast_1d(np.asarray(c, dtype=np.float64))
        if c.shape != (B.shape[0],):
            raise ArgumentError(f"a LinearChart takes c of length {B.shape[0]}")
        if not (np.isfinite(B).all() and np.isfinite(c).all()):
            raise ArgumentError("the linear chart needs a finite B and c")
        self.B, self.c = B, c

    def __call__(self, y):
        return self.B @ np.asarray(y, dtype=np.float64) + self.c


class PolarChart:
    """The polar chart of the reference's tests (test/nonlinear/reparametrized.jl:31-38):
    x = (r cos θ, r sin θ), y = (r, θ), ∂g = [cos θ, −r sin θ; sin θ, r cos θ]."""

    def __call__(self, y):
        return np.array([y[0] * math.cos(y[1]), y[0] * math.sin(y[1])])


class Reparametrized:
    """Reparametrized(prob, g, ∂g = nothing; covariant = false) (src/nonlinear/reparametrized.jl:1-60)
    over a registered chart: Newton iterates the chart coordinates y with F̃(y) = F(g(y)) and
    J̃ = J(g(y)) ∂g(y); a batch takes y₀ in X and receives x* = g(y*). `g` is a LinearChart or the
    PolarChart (n = 2), each carrying its Jacobian; `covariant = True` has no device form."""

    def __init__(self, prob, g, dg=None, covariant: bool = False, curvature: bool = True):
        if isinstance(prob, Reparametrized):
            raise ArgumentError(
                "Reparametrized cannot wrap a Reparametrized; compose the charts into one Reparametrized, with "
                "the chart g₁∘g₂ and the chart Jacobian ∂g₁(g₂)·∂g₂")
        if not isinstance(prob, NonlinearProblem):
            raise ArgumentError("the device Reparametrized wraps a NonlinearProblem of a registered family")
        if not isinstance(g, (LinearChart, PolarChart)):
            raise ArgumentError("the device Reparametrized takes a registered chart (LinearChart or PolarChart)")
        if covariant:
            raise ArgumentError("the device Reparametrized takes covariant = false (F̃ = F∘g)")
        self.prob, self.g, self.dg, self.covariant, self.curvature = prob, g, dg, False, curvature

    def to_c(self, n: int):
        c = _abi.GsChart()
        if isinstance(self.g, PolarChart):
            if n != 2:
                raise ArgumentError("the polar chart maps 2 coordinates: families 1 and 2")
            c.kind = _abi.GS_CHART_POLAR
            return c
        if self.g.B.shape[0] != n:
            raise DimensionMismatch(f"the chart maps {self.g.B.shape[0]} coordinates, the family has {n}")
        c.kind = _abi.GS_CHART_LINEAR
        for i in range(n):
            c.c[i] = float(self.g.c[i])
            for j in range(n):
                c.B[i * n + j] = float(self.g.B[i, j])
        return c


class Deflated:
    """Deflated(prob, roots; power = 2, shift = 1) (src/nonlinear/deflation.jl:15-235): the problem whose
    residual is M(x) F(x), M(x) = ∏ᵢ (‖x − rᵢ‖^(−p) + σ), with the product-rule Jacobian
    M J + F ∇Mᵀ, so that the known roots repel Newton; the status reports, and the residual test
    reads, the inner ‖F‖. `prob` is a NonlinearProblem of a lane-per-system family; `roots` a
    non-empty tuple of roots (numbers for n = 1, sequences of n reals otherwise), shared by every
    system of a batch. On the device: at most 4 roots and an integer power."""

    def __init__(self, prob, roots, power: float = 2, shift: float = 1):
        if not isinstance(prob, NonlinearProblem):
            raise ArgumentError("Deflated needs a NonlinearProblem or a StationarityProblem")
        power, shift = float(power), float(shift)
        if not (power > 0.5 and math.isfinite(power)):
            raise ArgumentError(
                "Deflated power must be finite and > 1/2: at or below 1/2 the known roots do not repel Newton")
        if not (shift > 0 and math.isfinite(shift)):
            raise ArgumentError("Deflated shift must be finite and > 0")
        if not isinstance(roots, (tuple, list)):
            raise ArgumentError(
                "Deflated roots must be a non-empty Tuple of roots or an AbstractVector of real vectors")
        if len(roots) == 0:
            raise ArgumentError("Deflated needs at least one root")
        rs = []
        for r in roots:
            v = np.atleast_1d(np.asarray(r, dtype=np.float64))
            if v.ndim != 1 or v.size == 0:
                raise ArgumentError("Deflated roots must be real numbers or real vectors")
            rs.append(v)
        if any(v.size != rs[0].size for v in rs):
            raise ArgumentError("Deflated needs roots of one length")
        if not all(np.isfinite(v).all() for v in rs):
            raise ArgumentError("Deflated needs roots with finite real entries")
        self.prob, self.roots, self.power, self.shift = prob, tuple(rs), power, shift

    def to_c(self, n: int):
        """gs_deflation for an iterate of n entries (the checks of a solve's start, deflation.jl:551-575)."""
        if any(v.size != n for v in self.roots):
            raise ArgumentError(
                f"Deflated needs roots of the length of the iterate: the start has {n} entries and a root has "
                f"{next(v.size for v in self.roots if v.size != n)}")
        if len(self.roots) > 4:
            raise ArgumentError("the device deflation takes at most 4 roots")
        if self.power != math.floor(self.power) or self.power > 64:
            raise ArgumentError("the device deflation takes an integer power in 1:64")
        d = _abi.GsDeflation()
        d.nroots, d.power, d.shift = len(self.roots), self.power, self.shift
        for r, v in enumerate(self.roots):
            for i, x in enumerate(v):
                d.roots[r * n + i] = float(x)
        return d


class NonlinearProblem:
    """NonlinearProblem(F) (src/nonlinear/problem.jl:81-110) over a registered device family."""

    def __init__(self, F: Family, J=None):
        if not isinstance(F, Family):
            raise ArgumentError("the device solve takes a registered family as the residual F")
        self.F = F
        self.J = J if J is not None else F  # each family carries its analytic Jacobian


class LeastSquaresProblem:
    """LeastSquaresProblem(F, J) (src/nonlinear/leastsquares.jl:1-40) over a registered least-squares
    family: minimise ½‖F(x, p)‖² over x, with the family's analytic Jacobian."""

    def __init__(self, F: Family, J=None):
        if not isinstance(F, ExpSumFits):
            raise ArgumentError("the device least-squares solve takes a registered least-squares family (ExpSumFits)")
        self.F = F
        self.J = J if J is not None else F


class ExpSumFits(Family):
    """Exponential-sum fits y(t) ≈ Σ_{k≤K} a_k exp(−λ_k t) at t_j = 5j/(m−1), j = 0..m−1.

    Unknowns x = (a_1..a_K, λ_1..λ_K), K ∈ 1..4, in the first 2K of 8 slots of a row of X.
    F_j = (Σ_k a_k e^(−λ_k t_j)) − y_j (a left fold over k), J_jk = e^(−λ_k t_j),
    J_j,K+k = ((−a_k) t_j) e^(−λ_k t_j). The exponential is a fixed algorithm (Cody–Waite reduction
    and a degree-13 Taylor polynomial with fused multiply-adds, within 1 ulp), so the device and the
    CPU oracle evaluate F and J to the same bits. Parameter: an ExpSumData.
    """

    id, n, p_len = _abi.GS_FAM_EXPSUM_LSQ, 8, 0
    max_points = 4096


class ExpSumData:
    """The per-fit data of an ExpSumFits batch: K (int32), m (int32) and y_offset (int64) per fit, and
    the concatenated samples Y (float64); fit i's samples are Y[y_offset[i] : y_offset[i] + m[i]].
    All four are HIP tensors on the device of X. It plays the part of a PerSystem (batched.jl:8-10)."""

    def __init__(self, K, m, y_offset, Y):
        torch = _torch()
        for name, t, dt in (("K", K, torch.int32), ("m", m, torch.int32), ("y_offset", y_offset, torch.int64),
                            ("Y", Y, torch.float64)):
            if not isinstance(t, torch.Tensor) or t.dtype != dt or t.dim() != 1:
                raise ArgumentError(f"ExpSumData {name} must be a 1-d torch tensor of {dt}")
            if not t.is_contiguous():
                raise ArgumentError(f"ExpSumData {name} must be contiguous")
        if not (len(K) == len(m) == len(y_offset)):
            raise DimensionMismatch(
                f"ExpSumData lengths differ: K {len(K)}, m {len(m)}, y_offset {len(y_offset)}")
        if not (K.device == m.device == y_offset.device == Y.device):
            raise ArgumentError("ExpSumData tensors must share one device")
        self.K, self.m, self.y_offset, self.Y = K, m, y_offset, Y

    def __len__(self):
        return len(self.K)


# --------------------------------------------------------------------------------------------
# Status buffers and the batched solve (batched.jl:49-332)
# --------------------------------------------------------------------------------------------


def status_buffer(nsys: int, device, eltype: str = "float64"):
    """A device vector of `nsys` SolverStatus{Float64} (72 bytes each) or, with eltype="float32",
    SolverStatus{Float32} (44 bytes each), as an (nsys, size) uint8 tensor."""
    size = (STATUS_DTYPE_F32 if eltype == "float32" else STATUS_DTYPE).itemsize
    return _torch().empty((nsys, size), dtype=_torch().uint8, device=device)


def statuses(S) -> np.ndarray:
    """The statuses of a status buffer as a host numpy array of dtype STATUS_DTYPE (or
    STATUS_DTYPE_F32 for a buffer of Float32 statuses)."""
    host = S.detach().to("cpu").contiguous().numpy()
    dt = STATUS_DTYPE_F32 if S.shape[-1] == STATUS_DTYPE_F32.itemsize else STATUS_DTYPE
    return host.reshape(-1).view(dt)


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


def _eltype(X) -> str:
    """The working precision of a batch: 'float64' or 'float32' (the R of its SolverStatus{R})."""
    torch = _torch()
    return "float32" if X.dtype == torch.float32 else "float64"


def _check_x(X, nfam: int, allow_f32: bool = False):
    torch = _torch()
    if not isinstance(X, torch.Tensor):
        raise ArgumentError("X must be a torch tensor of shape (nsys, n)")
    if X.dtype != torch.float64 and not (allow_f32 and X.dtype == torch.float32):
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
    if args and isinstance(args[0], (NonlinearProblem, LeastSquaresProblem, Deflated, StationarityProblem,
                                     Reparametrized)):
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
    if isinstance(prob, LeastSquaresProblem):
        return _solve_lsq(X, S, prob, method, p, order, options, workgroupsize, kwargs)
    if isinstance(prob, StationarityProblem):
        return _solve_stationarity(X, S, prob, method, p, order, options, workgroupsize, kwargs)
    deflation = reparam = None
    if isinstance(prob, Reparametrized):
        reparam, prob = prob, prob.prob
    if isinstance(prob, Deflated):
        deflation, prob = prob, prob.prob
        if isinstance(method, Newton) and method.best:
            raise ArgumentError(
                "Deflated takes no Newton with best = true: the best iterate would be ranked by the deflated "
                "residual, not by the inner residual that status.fnorm reports")
    if not isinstance(prob, NonlinearProblem):
        raise ArgumentError("prob must be a NonlinearProblem over a registered family")
    if not isinstance(method, Newton):
        raise ArgumentError("method must be a Newton")
    if isinstance(method.globalization, LevenbergMarquardt):
        raise ArgumentError(
            "the device solve of a NonlinearProblem takes Static(), Backtracking() or DogLeg(); "
            "LevenbergMarquardt() is the least-squares batch's (a LeastSquaresProblem)")
    fam = prob.F
    _check_x(X, fam.n, allow_f32=True)
    elt = _eltype(X)
    cdefl = cchart = None
    if reparam is not None:
        cchart = reparam.to_c(fam.n)
        if elt == "float32" and cchart.kind != _abi.GS_CHART_LINEAR:
            raise ArgumentError("the Float32 Reparametrized batch takes the linear chart")
    if deflation is not None:
        if elt != "float64":
            raise ArgumentError(
                "Deflated needs an iterate in the element type of its roots: the roots give Float64 and the "
                "start is Float32")
        cdefl = deflation.to_c(fam.n)
    order_h = _validate_batch(X, p, order)
    nsys = X.shape[0]
    torch = _torch()
    if S is not None:
        if S.dtype != torch.uint8 or S.dim() != 2 or S.shape[1] not in (STATUS_DTYPE.itemsize, STATUS_DTYPE_F32.itemsize):
            raise ArgumentError("S must be a status buffer (status_buffer(nsys, device))")
        if S.shape[1] != (STATUS_DTYPE_F32 if elt == "float32" else STATUS_DTYPE).itemsize:
            R = "Float64" if elt == "float32" else "Float32"
            raise ArgumentError(
                f"SolverStatus real type {R} does not match batch real element type {'Float32' if elt == 'float32' else 'Float64'}")
        if S.shape[0] != nsys:
            raise DimensionMismatch(
                f"status vector length {S.shape[0]} does not match batch length {nsys}")
        if S.device != X.device:
            raise ArgumentError(
                f"status vector backend {S.device} does not match batch backend {X.device}")
        if not S.is_contiguous():
            raise ArgumentError("S must be contiguous")
    ws = _validate_workgroupsize(workgroupsize)
    opt = resolve_options(options, kwargs, elt)
    cm = method.to_c()
    if X.device.type != "cuda":
        raise ArgumentError(
            "X must be a device array (a HIP tensor): libgsolve_hip solves on the GPU and has no host path")
    if nsys == 0:
        return S if S is not None else {k: 0 for k in _CODE_NAMES}
    lib = load_library()
    P, p_stride, shared = fam.marshal(p, nsys, X.device, X.dtype)
    cf = _abi.GsFamily()
    cf.id, cf.n, cf.m, cf.p_len = fam.id, fam.n, fam.n, fam.p_len
    for k, v in enumerate(shared):
        cf.shared[k] = v
    fdata = fam.device_data(X.device, X.dtype) if hasattr(fam, "device_data") else None
    if fdata is not None:
        cf.data = fdata.data_ptr()
    order_dev = None
    if order_h is not None:
        order_dev = torch.as_tensor(order_h, dtype=torch.int64).to(X.device)
    codes = None
    if S is None:
        codes = torch.empty(nsys, dtype=torch.uint8, device=X.device)
    copt = opt.to_c()
    if cchart is not None:
        entry = functools.partial(
            lib.gs_solve_batch_reparametrized_f32 if elt == "float32" else lib.gs_solve_batch_reparametrized_f64,
            ctypes.byref(cm), ctypes.byref(cf), ctypes.byref(cchart))
    elif cdefl is not None:
        entry = functools.partial(lib.gs_solve_batch_deflated_f64, ctypes.byref(cm), ctypes.byref(cf),
                                  ctypes.byref(cdefl))
    else:
        entry = functools.partial(lib.gs_solve_batch_f32 if elt == "float32" else lib.gs_solve_batch_f64,
                                  ctypes.byref(cm), ctypes.byref(cf))
    rc = entry(
        nsys, ctypes.c_void_p(X.data_ptr()),
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


def _solve_stationarity(X, S, prob, method, p, order, options, workgroupsize, kwargs):
    """solve_batch on a StationarityProblem (stationarity.jl:642-698): each row of X is a full state;
    the datum comes from the keyword `b` or the problem's own."""
    b_kw = kwargs.pop("b", None)
    if not isinstance(method, Newton):
        raise ArgumentError("method must be a Newton")
    if isinstance(method.globalization, LevenbergMarquardt):
        raise ArgumentError("the device StationarityProblem batch takes Static(), Backtracking() or DogLeg(); "
                            "LevenbergMarquardt has no stationarity form")
    act = prob.S
    _check_x(X, act.n, allow_f32=True)
    elt = _eltype(X)
    order_h = _validate_batch(X, p, order)
    if isinstance(b_kw if b_kw is not None else prob.b, PerSystem):
        _validate_batch(X, b_kw if b_kw is not None else prob.b, None)
    nsys = X.shape[0]
    if S is not None:
        _check_status_buffer(S, nsys, X, STATUS_DTYPE_F32 if elt == "float32" else STATUS_DTYPE)
    ws = _validate_workgroupsize(workgroupsize)
    opt = resolve_options(options, kwargs, elt)
    cm = method.to_c()
    if X.device.type != "cuda":
        raise ArgumentError(
            "X must be a device array (a HIP tensor): libgsolve_hip solves on the GPU and has no host path")
    cs, keep = prob.to_c(act.n, b_kw, nsys, X.device, np.float32 if elt == "float32" else np.float64)
    if nsys == 0:
        return S if S is not None else {k: 0 for k in _CODE_NAMES}
    lib = load_library()
    torch = _torch()
    P, p_stride, shared = act.marshal(p, nsys, X.device, X.dtype)
    cf = _abi.GsFamily()
    cf.id, cf.n, cf.m, cf.p_len = act.id, act.n, act.n, act.p_len
    order_dev = None if order_h is None else torch.as_tensor(order_h, dtype=torch.int64).to(X.device)
    codes = torch.empty(nsys, dtype=torch.uint8, device=X.device) if S is None else None
    copt = opt.to_c()
    entry = lib.gs_solve_batch_stationarity_f32 if elt == "float32" else lib.gs_solve_batch_stationarity_f64
    rc = entry(
        ctypes.byref(cm), ctypes.byref(cf), ctypes.byref(cs), nsys, ctypes.c_void_p(X.data_ptr()),
        ctypes.c_void_p(P.data_ptr()), p_stride,
        ctypes.c_void_p(S.data_ptr()) if S is not None else None,
        ctypes.c_void_p(codes.data_ptr()) if codes is not None else None,
        ctypes.c_void_p(order_dev.data_ptr()) if order_dev is not None else None,
        ctypes.byref(copt), ws, _stream_ptr(X))
    del keep
    _raise_on(rc, lib)
    if S is not None:
        return S
    counts = (ctypes.c_int64 * 6)()
    _raise_on(lib.gs_tally_codes(ctypes.c_void_p(codes.data_ptr()), nsys, counts, _stream_ptr(X)), lib)
    return {k: int(counts[i]) for i, k in enumerate(_CODE_NAMES)}


def _check_status_buffer(S, nsys: int, X, dtype=None) -> None:
    torch = _torch()
    size = (STATUS_DTYPE if dtype is None else dtype).itemsize
    if S.dtype != torch.uint8 or S.dim() != 2 or S.shape[1] != size:
        raise ArgumentError("S must be a status buffer (status_buffer(nsys, device))")
    if S.shape[0] != nsys:
        raise DimensionMismatch(f"status vector length {S.shape[0]} does not match batch length {nsys}")
    if S.device != X.device:
        raise ArgumentError(f"status vector backend {S.device} does not match batch backend {X.device}")
    if not S.is_contiguous():
        raise ArgumentError("S must be contiguous")


def _solve_lsq(X, S, prob, method, p, order, options, workgroupsize, kwargs):
    """solve_batch on a LeastSquaresProblem: the per-fit R1 solve (solver.jl:148-189, 277-378) with
    Newton(; globalization = LevenbergMarquardt()) for every fit of the batch, in one call."""
    if not isinstance(method, Newton):
        raise ArgumentError("method must be a Newton")
    if not isinstance(method.globalization, LevenbergMarquardt):
        raise ArgumentError(
            "the device least-squares batch takes Newton(; globalization = LevenbergMarquardt())")
    _check_x(X, prob.F.n)
    nfit = X.shape[0]
    if not isinstance(p, ExpSumData):
        raise ArgumentError("an ExpSumFits batch takes its data as p = ExpSumData(K, m, y_offset, Y)")
    if len(p) != nfit:
        raise DimensionMismatch(f"ExpSumData length {len(p)} does not match batch length {nfit}")
    if order is not None:
        raise ArgumentError("the least-squares batch orders its fits itself (bucketed by size); order must be None")
    if workgroupsize is not None:
        raise ArgumentError(
            "the least-squares batch takes its launch shapes from tune_lsq (or the library defaults); "
            "workgroupsize must be None")
    if S is not None:
        _check_status_buffer(S, nfit, X)
    opt = resolve_options(options, kwargs)
    cm = method.to_c()
    if X.device.type != "cuda":
        raise ArgumentError(
            "X must be a device array (a HIP tensor): libgsolve_hip solves on the GPU and has no host path")
    if p.K.device != X.device:
        raise ArgumentError(f"ExpSumData backend {p.K.device} does not match batch backend {X.device}")
    if nfit == 0:
        return S if S is not None else {k: 0 for k in _CODE_NAMES}
    lib = load_library()
    own = S is None
    if own:
        S = status_buffer(nfit, X.device)
    copt = opt.to_c()
    stream = _stream_ptr(X)
    rc = lib.gs_solve_expsum_lsq_f64(
        ctypes.byref(cm), nfit, ctypes.c_void_p(X.data_ptr()), ctypes.c_void_p(p.K.data_ptr()),
        ctypes.c_void_p(p.m.data_ptr()), ctypes.c_void_p(p.y_offset.data_ptr()), ctypes.c_void_p(p.Y.data_ptr()),
        len(p.Y), ctypes.c_void_p(S.data_ptr()), ctypes.byref(copt), stream)
    _raise_on(rc, lib)
    if not own:
        return S
    codes = S[:, 0].contiguous()
    counts = (ctypes.c_int64 * 6)()
    _raise_on(lib.gs_tally_codes(ctypes.c_void_p(codes.data_ptr()), nfit, counts, stream), lib)
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


from . import autotune  # noqa: E402  (the least-squares launch-shape autotuner)
from .autotune import tune_lsq, tune_dense, tune_workgroup, dense_set_shape, dense_get_shape  # noqa: E402
from . import learned  # noqa: E402  (the learned start; imports the names above lazily)
from .learned import LearnedStart, generate_solutions, solve_batch_learned  # noqa: E402
