"""ctypes binding of libgsolve_hip (include/gsolve.h).

The structs here mirror the C declarations field by field; `load_library` checks their sizes
against the library's own `gs_abi_sizes` when it loads, as the Julia extension asserts
`sizeof`/`fieldoffset` of `SolverStatus{Float64}` at load time.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

ABI_VERSION = 2
LIB_DIR = Path(__file__).resolve().parent / "lib"
LIB_PATH = LIB_DIR / "libgsolve_hip.so"

# gs_error
GS_OK, GS_EINVAL, GS_EDIM, GS_EFAMILY, GS_EHIP = 0, -1, -2, -3, -4
# gs_globalization / gs_factor_precision / gs_refinement / gs_norm / gs_linsolve
GS_STATIC, GS_BACKTRACKING, GS_LEVENBERG_MARQUARDT, GS_DOGLEG = 0, 1, 2, 3
GS_FACTOR_WORKING, GS_FACTOR_F32 = 0, 1
GS_NO_REFINEMENT, GS_REFINEMENT = 0, 1
GS_NORM_L2, GS_NORM_MAX = 0, 1
GS_LM_SCALING_NONE, GS_LM_SCALING_DIAGONAL = 0, 1
GS_LU = 0
# gs_family_id
GS_FAM_COUPLED2 = 1
GS_FAM_HETERO2 = 2
GS_FAM_SCALAR_SQUARE = 3
GS_FAM_TANH_LINEAR3 = 4
GS_FAM_HENON_HEILES_MIDPOINT = 5
GS_FAM_FPU_GL3 = 6
GS_FAM_EXPSUM_LSQ = 7
GS_FAM_QUAD_MAP2 = 9
GS_FAM_DENSE_SIN = 10
GS_ACT_HENON_HEILES = 8

GS_DATUM_NONE, GS_DATUM_SHARED, GS_DATUM_PER_SYSTEM = 0, 1, 2
GS_CHART_LINEAR, GS_CHART_POLAR = 1, 2


class GsStatusF64(ctypes.Structure):
    """gs_status_f64: the byte layout of Julia's SolverStatus{Float64} (src/base/status.jl:101-117)."""

    _fields_ = [
        ("code", ctypes.c_uint8),
        ("pad0_", ctypes.c_uint8),
        ("f_calls", ctypes.c_int16),
        ("iterations", ctypes.c_int32),
        ("fnorm", ctypes.c_double),
        ("stepnorm", ctypes.c_double),
        ("step_failures", ctypes.c_int32),
        ("promoted", ctypes.c_uint8),
        ("stationary", ctypes.c_uint8),
        ("jvp_calls", ctypes.c_int16),
        ("rank", ctypes.c_int32),
        ("pad1_", ctypes.c_int32),
        ("nullnorm", ctypes.c_double),
        ("gradratio", ctypes.c_double),
        ("rate", ctypes.c_double),
        ("contracted", ctypes.c_uint8),
        ("reused", ctypes.c_uint8),
        ("pad2_", ctypes.c_uint8 * 6),
    ]


class GsOptionsF64(ctypes.Structure):
    _fields_ = [
        ("f_abstol", ctypes.c_double),
        ("f_reltol", ctypes.c_double),
        ("x_abstol", ctypes.c_double),
        ("x_reltol", ctypes.c_double),
        ("min_iterations", ctypes.c_int32),
        ("max_iterations", ctypes.c_int32),
        ("max_stalls", ctypes.c_int32),
        ("norm", ctypes.c_int32),
        ("g_tol", ctypes.c_double),
        ("kappa", ctypes.c_double),
        ("roundoff", ctypes.c_uint8),
        ("rate_test", ctypes.c_uint8),
        ("pad_", ctypes.c_uint8 * 6),
    ]


class GsMethod(ctypes.Structure):
    _fields_ = [
        ("linsolve", ctypes.c_int32),
        ("factor_precision", ctypes.c_int32),
        ("globalization", ctypes.c_int32),
        ("refinement", ctypes.c_int32),
        ("static_alpha", ctypes.c_double),
        ("bt_c1", ctypes.c_double),
        ("bt_p", ctypes.c_double),
        ("bt_maxiter", ctypes.c_int32),
        ("ref_maxiter", ctypes.c_int32),
        ("ref_eta", ctypes.c_double),
        ("regularization", ctypes.c_double),
        ("lm_mu0", ctypes.c_double),
        ("lm_mu_min", ctypes.c_double),
        ("lm_mu_max", ctypes.c_double),
        ("lm_mu_decrease", ctypes.c_double),
        ("lm_mu_increase", ctypes.c_double),
        ("lm_rho_accept", ctypes.c_double),
        ("best", ctypes.c_int32),
        ("reuse", ctypes.c_int32),
        ("dl_radius", ctypes.c_double),
        ("lm_scaling", ctypes.c_int32),
        ("pad_", ctypes.c_int32),
    ]


class GsFamily(ctypes.Structure):
    _fields_ = [
        ("id", ctypes.c_int32),
        ("n", ctypes.c_int32),
        ("m", ctypes.c_int32),
        ("p_len", ctypes.c_int32),
        ("shared", ctypes.c_double * 8),
        ("data", ctypes.c_void_p),
    ]


class GsDeflation(ctypes.Structure):
    """gs_deflation: the shared roots, power and shift of a Deflated batch."""
    _fields_ = [
        ("nroots", ctypes.c_int32),
        ("pad_", ctypes.c_int32),
        ("power", ctypes.c_double),
        ("shift", ctypes.c_double),
        ("roots", ctypes.c_double * 32),
    ]


class GsStationarity(ctypes.Structure):
    """gs_stationarity: the index sets (1-based; n_eq = 0 for E = U = :) and the datum of a
    StationarityProblem batch."""
    _fields_ = [
        ("n_eq", ctypes.c_int32),
        ("b_kind", ctypes.c_int32),
        ("e", ctypes.c_int32 * 8),
        ("u", ctypes.c_int32 * 8),
        ("b_shared", ctypes.c_double * 8),
        ("b_data", ctypes.c_void_p),
        ("b_stride", ctypes.c_int64),
    ]


class GsChart(ctypes.Structure):
    """gs_chart: a registered chart of a Reparametrized batch (linear: B row-major and c)."""
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("pad_", ctypes.c_int32),
        ("B", ctypes.c_double * 16),
        ("c", ctypes.c_double * 4),
    ]


# SolverStatus{Float64} as a numpy structured dtype (72 bytes, Julia's field offsets).
STATUS_DTYPE = np.dtype(
    {
        "names": [
            "code", "f_calls", "iterations", "fnorm", "stepnorm", "step_failures", "promoted",
            "stationary", "jvp_calls", "rank", "nullnorm", "gradratio", "rate", "contracted",
            "reused",
        ],
        "formats": [
            "u1", "<i2", "<i4", "<f8", "<f8", "<i4", "u1", "u1", "<i2", "<i4", "<f8", "<f8",
            "<f8", "u1", "u1",
        ],
        "offsets": [0, 2, 4, 8, 16, 24, 28, 29, 30, 32, 40, 48, 56, 64, 65],
        "itemsize": 72,
    }
)

# SolverStatus{Float32} (44 bytes, Julia's field offsets).
STATUS_DTYPE_F32 = np.dtype(
    {
        "names": list(STATUS_DTYPE.names),
        "formats": [
            "u1", "<i2", "<i4", "<f4", "<f4", "<i4", "u1", "u1", "<i2", "<i4", "<f4", "<f4",
            "<f4", "u1", "u1",
        ],
        "offsets": [0, 2, 4, 8, 12, 16, 20, 21, 22, 24, 28, 32, 36, 40, 41],
        "itemsize": 44,
    }
)

GS_MLP_IN, GS_MLP_HIDDEN, GS_MLP_OUT = 32, 128, 16
GS_MLP_NPARAM, GS_MLP_NPACKED = 39312, 38912


class GsMlp(ctypes.Structure):
    """gs_mlp: the learned-start network (device parameter and packed-weight pointers, normalisation)."""

    _fields_ = [
        ("params", ctypes.c_void_p),
        ("packed", ctypes.c_void_p),
        ("in_shift", ctypes.c_float * 8),
        ("in_scale", ctypes.c_float * 8),
        ("out_shift", ctypes.c_float * 4),
        ("out_scale", ctypes.c_float * 4),
    ]


# Every symbol that include/gsolve.h declares.
EXPORTED_SYMBOLS = (
    "gs_solve_batch_f64",
    "gs_solve_batch_f64_noalloc",
    "gs_tally_codes",
    "gs_batch_summary_f64",
    "gs_last_error",
    "gs_version",
    "gs_abi_sizes",
    "gs_set_kernel_timing",
    "gs_last_kernel_ms",
    "gs_solve_expsum_lsq_f64",
    "gs_shutdown",
    "gs_solve_batch_f32",
    "gs_solve_batch_deflated_f64",
    "gs_solve_batch_stationarity_f64",
    "gs_solve_batch_reparametrized_f64",
    "gs_solve_batch_stationarity_f32",
    "gs_solve_batch_reparametrized_f32",
    "gs_lsq_set_shape",
    "gs_lsq_get_shape",
    "gs_lsq_set_graph",
    "gs_dense_set_shape",
    "gs_dense_get_shape",
    "gs_mlp_start_henon_heiles",
    "gs_solve_learned_henon_heiles_f64",
    "gs_mlp_grad",
    "gs_mlp_adam",
    "gs_mlp_pack",
)

_LIB = None


class LibraryMissing(RuntimeError):
    """libgsolve_hip is not built: there is no fallback path."""


def load_library(path: os.PathLike | None = None):
    """Load libgsolve_hip, declare the signatures and check the ABI version and struct sizes."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise LibraryMissing(
            f"{p} is not built; run __graft_entry__.build() (hipcc --offload-arch=gfx950). "
            "The batched solve has no CPU or PyTorch fallback."
        )
    lib = ctypes.CDLL(str(p))
    vp = ctypes.c_void_p
    lib.gs_solve_batch_f64.restype = ctypes.c_int
    lib.gs_solve_batch_f64.argtypes = [
        ctypes.POINTER(GsMethod), ctypes.POINTER(GsFamily), ctypes.c_int64, vp, vp, ctypes.c_int64,
        vp, vp, vp, ctypes.POINTER(GsOptionsF64), ctypes.c_int32, vp,
    ]
    lib.gs_solve_batch_f64_noalloc.restype = ctypes.c_int
    lib.gs_solve_batch_f64_noalloc.argtypes = lib.gs_solve_batch_f64.argtypes
    lib.gs_tally_codes.restype = ctypes.c_int
    lib.gs_tally_codes.argtypes = [vp, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64), vp]
    lib.gs_batch_summary_f64.restype = ctypes.c_int
    lib.gs_batch_summary_f64.argtypes = [
        vp, ctypes.c_int64, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64), vp,
    ]
    lib.gs_last_error.restype = ctypes.c_char_p
    lib.gs_last_error.argtypes = []
    lib.gs_version.restype = ctypes.c_int32
    lib.gs_version.argtypes = []
    lib.gs_abi_sizes.restype = None
    lib.gs_abi_sizes.argtypes = [ctypes.POINTER(ctypes.c_int64)]
    lib.gs_set_kernel_timing.restype = None
    lib.gs_set_kernel_timing.argtypes = [ctypes.c_int]
    lib.gs_last_kernel_ms.restype = ctypes.c_double
    lib.gs_last_kernel_ms.argtypes = []
    lib.gs_solve_expsum_lsq_f64.restype = ctypes.c_int
    lib.gs_solve_expsum_lsq_f64.argtypes = [
        ctypes.POINTER(GsMethod), ctypes.c_int64, vp, vp, vp, vp, vp, ctypes.c_int64, vp,
        ctypes.POINTER(GsOptionsF64), vp,
    ]
    lib.gs_mlp_start_henon_heiles.restype = ctypes.c_int
    lib.gs_mlp_start_henon_heiles.argtypes = [ctypes.POINTER(GsMlp), ctypes.c_int64, vp, ctypes.c_double, vp, vp]
    lib.gs_solve_learned_henon_heiles_f64.restype = ctypes.c_int
    lib.gs_solve_learned_henon_heiles_f64.argtypes = [
        ctypes.POINTER(GsMlp), ctypes.POINTER(GsMethod), ctypes.c_int64, vp, ctypes.c_double, vp, vp,
        ctypes.POINTER(GsOptionsF64), ctypes.POINTER(ctypes.c_int64), vp,
    ]
    lib.gs_mlp_grad.restype = ctypes.c_int
    lib.gs_mlp_grad.argtypes = [ctypes.POINTER(GsMlp), vp, vp, ctypes.c_double, vp, ctypes.c_int32, vp, vp, vp]
    lib.gs_mlp_adam.restype = ctypes.c_int
    lib.gs_mlp_adam.argtypes = [vp, vp, vp, vp, vp, ctypes.c_int64, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                ctypes.c_float, vp]
    lib.gs_mlp_pack.restype = ctypes.c_int
    lib.gs_mlp_pack.argtypes = [vp, vp, vp]
    lib.gs_lsq_set_shape.restype = ctypes.c_int
    lib.gs_lsq_set_shape.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.gs_lsq_get_shape.restype = ctypes.c_int
    lib.gs_lsq_get_shape.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_void_p]
    lib.gs_dense_set_shape.restype = ctypes.c_int
    lib.gs_dense_set_shape.argtypes = [ctypes.c_int, ctypes.c_int]
    lib.gs_dense_get_shape.restype = ctypes.c_int
    lib.gs_dense_get_shape.argtypes = [ctypes.c_int]
    lib.gs_lsq_set_graph.restype = ctypes.c_int
    lib.gs_lsq_set_graph.argtypes = [ctypes.c_int]
    lib.gs_shutdown.restype = ctypes.c_int
    lib.gs_shutdown.argtypes = []
    lib.gs_solve_batch_f32.restype = ctypes.c_int
    lib.gs_solve_batch_f32.argtypes = lib.gs_solve_batch_f64.argtypes
    lib.gs_solve_batch_deflated_f64.restype = ctypes.c_int
    lib.gs_solve_batch_deflated_f64.argtypes = [
        ctypes.POINTER(GsMethod), ctypes.POINTER(GsFamily), ctypes.POINTER(GsDeflation), ctypes.c_int64, vp, vp,
        ctypes.c_int64, vp, vp, vp, ctypes.POINTER(GsOptionsF64), ctypes.c_int32, vp,
    ]
    lib.gs_solve_batch_stationarity_f64.restype = ctypes.c_int
    lib.gs_solve_batch_stationarity_f64.argtypes = [
        ctypes.POINTER(GsMethod), ctypes.POINTER(GsFamily), ctypes.POINTER(GsStationarity), ctypes.c_int64, vp, vp,
        ctypes.c_int64, vp, vp, vp, ctypes.POINTER(GsOptionsF64), ctypes.c_int32, vp,
    ]
    lib.gs_solve_batch_reparametrized_f64.restype = ctypes.c_int
    lib.gs_solve_batch_reparametrized_f64.argtypes = [
        ctypes.POINTER(GsMethod), ctypes.POINTER(GsFamily), ctypes.POINTER(GsChart), ctypes.c_int64, vp, vp,
        ctypes.c_int64, vp, vp, vp, ctypes.POINTER(GsOptionsF64), ctypes.c_int32, vp,
    ]
    lib.gs_solve_batch_stationarity_f32.restype = ctypes.c_int
    lib.gs_solve_batch_stationarity_f32.argtypes = lib.gs_solve_batch_stationarity_f64.argtypes
    lib.gs_solve_batch_reparametrized_f32.restype = ctypes.c_int
    lib.gs_solve_batch_reparametrized_f32.argtypes = lib.gs_solve_batch_reparametrized_f64.argtypes
    if lib.gs_version() != ABI_VERSION:
        raise RuntimeError(f"libgsolve_hip ABI version {lib.gs_version()} != {ABI_VERSION}")
    sizes = (ctypes.c_int64 * 4)()
    lib.gs_abi_sizes(sizes)
    want = (ctypes.sizeof(GsStatusF64), ctypes.sizeof(GsOptionsF64), ctypes.sizeof(GsMethod),
            ctypes.sizeof(GsFamily))
    if tuple(sizes) != want or STATUS_DTYPE.itemsize != 72 or want[0] != 72:
        raise RuntimeError(f"ABI struct sizes differ: library {tuple(sizes)}, binding {want}")
    if path is None:
        _LIB = lib
    return lib


def last_error(lib) -> str:
    msg = lib.gs_last_error()
    return msg.decode() if msg else ""
