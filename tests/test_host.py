"""Host logic of the binding: the C-ABI library loads and exports what include/gsolve.h declares,
the struct layouts match Julia's SolverStatus{Float64}, and solve_batch validates its arguments with
the reference's errors (src/nonlinear/batched.jl:113-150, 205-210, 257-307; options.jl:207-245)
before any device work."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest
import torch

ROOT = Path(__file__).resolve().parents[1]


def test_library_exports_every_declared_symbol(gs):
    header = (ROOT / "include" / "gsolve.h").read_text()
    declared = set(re.findall(r"^\S.*?\b(gs_\w+)\(", header, flags=re.M))
    from importlib import import_module

    abi = import_module(gs.__name__ + "._abi")
    lib = ctypes.CDLL(str(abi.LIB_PATH))
    for name in declared:
        assert hasattr(lib, name), name
    assert declared == set(abi.EXPORTED_SYMBOLS)
    loaded = gs.load_library()
    assert loaded.gs_version() == 2


def test_status_layout_matches_julia(gs):
    dt = gs.STATUS_DTYPE
    assert dt.itemsize == 72
    offs = {n: dt.fields[n][1] for n in dt.names}
    assert offs == {"code": 0, "f_calls": 2, "iterations": 4, "fnorm": 8, "stepnorm": 16,
                    "step_failures": 24, "promoted": 28, "stationary": 29, "jvp_calls": 30, "rank": 32,
                    "nullnorm": 40, "gradratio": 48, "rate": 56, "contracted": 64, "reused": 65}


def test_return_codes(gs):
    assert [int(c) for c in gs.ReturnCode] == list(range(6))


def test_options_validation(gs):
    with pytest.raises(gs.ArgumentError, match="f_abstol must be non-negative"):
        gs.Options(f_abstol=-1.0)
    with pytest.raises(gs.ArgumentError, match="x_reltol must be non-negative"):
        gs.Options(x_reltol=-1)
    with pytest.raises(gs.ArgumentError):
        gs.Options(f_abstol=float("nan"))
    with pytest.raises(gs.ArgumentError, match="max_iterations = 1 is below min_iterations = 2"):
        gs.Options(min_iterations=2, max_iterations=1)
    with pytest.raises(gs.ArgumentError, match="max_stalls must be at least 1"):
        gs.Options(max_stalls=0)
    with pytest.raises(gs.ArgumentError, match="kappa"):
        gs.Options(kappa=1.0)
    o = gs.Options()
    assert o.f_reltol == np.sqrt(np.finfo(float).eps) and o.x_reltol == 2 * np.finfo(float).eps
    assert o.max_iterations == 100 and o.max_stalls == 2 and o.kappa == 0.1 and o.rate_test


def test_method_validation(gs):
    with pytest.raises(gs.ArgumentError, match="Backtracking needs 0 < c₁ < 1"):
        gs.Backtracking(c1=1.5)
    with pytest.raises(gs.ArgumentError, match="the step of Static must be positive and finite"):
        gs.Static(0.0)
    with pytest.raises(gs.ArgumentError, match="Refinement maxiter must be ≥ 1"):
        gs.Refinement(maxiter=0)
    with pytest.raises(gs.ArgumentError, match="regularization"):
        gs.Newton(regularization=-1.0)
    with pytest.raises(gs.ArgumentError, match="keep_factor"):
        gs.Newton(keep_factor=True)
    assert isinstance(gs.Newton(refinement=True).refinement, gs.Refinement)


def _cpu_batch(gs, n=3):
    X = torch.zeros((n, 2), dtype=torch.float64)
    return X, gs.NonlinearProblem(gs.Coupled2())


def test_validation_order_and_messages(gs):
    """The host checks raise before any launch, so they run without a GPU on CPU tensors."""
    X, prob = _cpu_batch(gs)
    P = torch.zeros((2, 2), dtype=torch.float64)
    with pytest.raises(gs.DimensionMismatch, match="PerSystem parameter length 2 does not match batch length 3"):
        gs.solve_batch(X, prob, gs.Newton(), gs.PerSystem(P))
    with pytest.raises(gs.DimensionMismatch, match="order length 1 does not match batch length 3"):
        gs.solve_batch(X, prob, gs.Newton(), (2.125, 1.5), order=[1])
    with pytest.raises(gs.ArgumentError, match=r"order is not a permutation of 1:3"):
        gs.solve_batch(X, prob, gs.Newton(), (2.125, 1.5), order=[1, 1, 2])
    with pytest.raises(gs.ArgumentError, match=r"order is not a permutation of 1:3"):
        gs.solve_batch(X, prob, gs.Newton(), (2.125, 1.5), order=[1, 4, 2])
    S = torch.empty((2, 72), dtype=torch.uint8)
    with pytest.raises(gs.DimensionMismatch, match="status vector length 2 does not match batch length 3"):
        gs.solve_batch(X, S, prob, gs.Newton(), (2.125, 1.5))
    with pytest.raises(gs.ArgumentError, match="workgroupsize must be a positive integer"):
        gs.solve_batch(X, prob, gs.Newton(), (2.125, 1.5), workgroupsize=0)
    with pytest.raises(gs.ArgumentError, match="pass either `options` or keyword arguments, not both"):
        gs.solve_batch(X, prob, gs.Newton(), (2.125, 1.5), options=gs.Options(), max_iterations=3)
    with pytest.raises(gs.ArgumentError, match="SolverStatus real type Float64"):
        gs.solve_batch(X.half(), prob, gs.Newton(), (2.125, 1.5))
    # a Float32 batch takes SolverStatus{Float32} and Options{Float32} (batched.jl:268-272, options.jl:241-245)
    with pytest.raises(gs.ArgumentError, match="SolverStatus real type Float64 does not match batch real element type Float32"):
        gs.solve_batch(X.float(), gs.status_buffer(3, "cpu"), prob, gs.Newton(), (2.125, 1.5))
    with pytest.raises(gs.ArgumentError, match=r"options are Options\{Float64\}, but the iterate has element type Float32"):
        gs.solve_batch(X.float(), prob, gs.Newton(), (2.125, 1.5), options=gs.Options())
    with pytest.raises(gs.ArgumentError, match="device array"):
        gs.solve_batch(X.float(), prob, gs.Newton(), (2.125, 1.5), options=gs.Options(eltype="float32"))


def test_host_tensor_is_refused_not_solved(gs):
    """There is no CPU fallback: a host X raises."""
    X, prob = _cpu_batch(gs)
    with pytest.raises(gs.ArgumentError, match="device array"):
        gs.solve_batch(X, prob, gs.Newton(), (2.125, 1.5))


def test_empty_batch(gs):
    X = torch.zeros((0, 2), dtype=torch.float64, device="cuda" if torch.cuda.is_available() else "cpu")
    prob = gs.NonlinearProblem(gs.Coupled2())
    if X.device.type == "cpu":
        pytest.skip("empty-batch call needs a device tensor")
    counts = gs.solve_batch(X, prob, gs.Newton(), (2.125, 1.5))
    assert counts == {k: 0 for k in ("SUCCESS", "STALLED", "MAXITERS", "SINGULAR", "NONFINITE",
                                     "LINESEARCH_FAILED")}
    assert gs.isconverged(counts)


def test_isconverged_counts(gs):
    ok = dict(SUCCESS=10, STALLED=0, MAXITERS=0, SINGULAR=0, NONFINITE=0, LINESEARCH_FAILED=0)
    assert gs.isconverged(ok)
    for k in ("STALLED", "MAXITERS", "SINGULAR", "NONFINITE", "LINESEARCH_FAILED"):
        assert not gs.isconverged({**ok, k: 1})


def test_workloads_are_reproducible(workloads):
    a = workloads.henon_heiles_states(1000, 0)
    b = workloads.henon_heiles_states(1000, 0)
    assert np.array_equal(a, b)
    E = workloads.henon_heiles_energy(a)
    assert E.min() >= 0.01 - 1e-12 and E.max() <= 0.15 + 1e-12
    rec, x0 = workloads.tanh_linear3_batch(100, 0)
    assert rec.shape == (100, 12) and not x0.any()


def test_least_squares_batch_validation(gs, workloads):
    """The least-squares batch (ExpSumFits + LevenbergMarquardt) checks its arguments on the host
    before any device work."""
    K, m, off, Y, X0, _ = workloads.expsum_fits(8, seed=1)
    data = gs.ExpSumData(torch.as_tensor(K), torch.as_tensor(m), torch.as_tensor(off), torch.as_tensor(Y))
    prob = gs.LeastSquaresProblem(gs.ExpSumFits())
    lm = gs.Newton(globalization=gs.LevenbergMarquardt())
    X = torch.as_tensor(X0).clone()
    with pytest.raises(gs.ArgumentError, match="LevenbergMarquardt"):
        gs.solve_batch(X, prob, gs.Newton(), data)
    with pytest.raises(gs.ArgumentError, match="order"):
        gs.solve_batch(X, prob, lm, data, order=np.arange(1, 9))
    with pytest.raises(gs.ArgumentError, match="workgroupsize"):
        gs.solve_batch(X, prob, lm, data, workgroupsize=64)
    with pytest.raises(gs.ArgumentError, match="ExpSumData"):
        gs.solve_batch(X, prob, lm, gs.PerSystem(np.zeros(8)))
    with pytest.raises(gs.DimensionMismatch, match="ExpSumData length 8 does not match batch length 7"):
        gs.solve_batch(X[:7].clone(), prob, lm, data)
    with pytest.raises(gs.DimensionMismatch):
        gs.solve_batch(torch.zeros(8, 4, dtype=torch.float64), prob, lm, data)
    with pytest.raises(gs.ArgumentError, match="device array"):
        gs.solve_batch(X, prob, lm, data)
    with pytest.raises(gs.DimensionMismatch, match="lengths differ"):
        gs.ExpSumData(torch.as_tensor(K), torch.as_tensor(m[:4]), torch.as_tensor(off), torch.as_tensor(Y))
    with pytest.raises(gs.ArgumentError, match="int32"):
        gs.ExpSumData(torch.as_tensor(K).long(), torch.as_tensor(m), torch.as_tensor(off), torch.as_tensor(Y))
    with pytest.raises(gs.ArgumentError, match="least-squares"):
        gs.solve_batch(torch.zeros(4, 4, dtype=torch.float64), gs.NonlinearProblem(gs.HenonHeilesMidpoint()), lm,
                       gs.BatchParams(0.1, torch.zeros(4, 4, dtype=torch.float64)))
    with pytest.raises(gs.ArgumentError, match="least-squares family"):
        gs.LeastSquaresProblem(gs.HenonHeilesMidpoint())
    with pytest.raises(gs.ArgumentError, match="μ_increase > 1"):
        gs.LevenbergMarquardt(mu_increase=1.0)
    with pytest.raises(gs.ArgumentError, match=r"scaling ∈ \(:none, :diagonal\); got scaling = :bad"):
        gs.LevenbergMarquardt(scaling="bad")


def test_least_squares_entry_point_checks_method_without_gpu(gs):
    """gs_solve_expsum_lsq_f64 rejects a method outside the LevenbergMarquardt constructor's bounds
    (which also bound its damping loop) before touching any pointer."""
    lib = gs.load_library()
    from importlib import import_module

    abi = import_module(gs.__name__ + "._abi")
    m = gs.Newton(globalization=gs.LevenbergMarquardt()).to_c()
    bad = abi.GsMethod.from_buffer_copy(m)
    bad.lm_mu_increase = 1.0
    o = gs.Options().to_c()
    rc = lib.gs_solve_expsum_lsq_f64(ctypes.byref(bad), 1, None, None, None, None, None, 0, None, ctypes.byref(o), None)
    assert rc == abi.GS_EINVAL and "μ_increase" in abi.last_error(lib)
    assert m.lm_scaling == abi.GS_LM_SCALING_NONE
    assert gs.Newton(globalization=gs.LevenbergMarquardt(scaling="diagonal")).to_c().lm_scaling == abi.GS_LM_SCALING_DIAGONAL
    bad = abi.GsMethod.from_buffer_copy(m)
    bad.lm_scaling = 2
    rc = lib.gs_solve_expsum_lsq_f64(ctypes.byref(bad), 1, None, None, None, None, None, 0, None, ctypes.byref(o), None)
    assert rc == abi.GS_EINVAL and "scaling" in abi.last_error(lib)
    bad = abi.GsMethod.from_buffer_copy(m)
    bad.globalization = abi.GS_STATIC
    rc = lib.gs_solve_expsum_lsq_f64(ctypes.byref(bad), 1, None, None, None, None, None, 0, None, ctypes.byref(o), None)
    assert rc == abi.GS_EFAMILY
    rc = lib.gs_lsq_set_shape(5, 256, 1)
    assert rc == abi.GS_EINVAL
    assert lib.gs_lsq_set_shape(1, 64, 65) == abi.GS_EINVAL and lib.gs_lsq_set_shape(1, 96, 0) == abi.GS_EINVAL


def test_dense_shape_setting_without_gpu(gs):
    """gs_dense_set_shape takes, per size, the segment widths that hold its rows (0 restores the
    default) and refuses any other width or size; it touches no device."""
    lib = gs.load_library()
    from importlib import import_module

    abi = import_module(gs.__name__ + "._abi")
    allowed = {12: (16, 32), 16: (16, 32), 24: (32, 64), 32: (32, 64), 48: (64,), 64: (64,)}
    assert allowed == gs.autotune.DENSE_LANES
    try:
        for n, widths in allowed.items():
            assert lib.gs_dense_get_shape(n) == widths[0]
            for lanes in (8, 16, 32, 64, 128):
                rc = lib.gs_dense_set_shape(n, lanes)
                if lanes in widths:
                    assert rc == 0 and lib.gs_dense_get_shape(n) == lanes
                else:
                    assert rc == abi.GS_EINVAL and f"n = {n}" in abi.last_error(lib)
            assert lib.gs_dense_set_shape(n, 0) == 0 and lib.gs_dense_get_shape(n) == widths[0]
        assert lib.gs_dense_set_shape(40, 64) == abi.GS_EINVAL and lib.gs_dense_get_shape(40) == abi.GS_EINVAL
        # n ≤ 8 runs one lane per system: no segment width
        assert lib.gs_dense_set_shape(8, 16) == abi.GS_EINVAL and lib.gs_dense_get_shape(8) == abi.GS_EINVAL
    finally:
        for n in allowed:
            lib.gs_dense_set_shape(n, 0)
    with pytest.raises(gs.ArgumentError, match="order 2, 4, 6, 8, 12, 16, 24, 32, 48 or 64"):
        gs.DenseSin(np.eye(40))
    assert gs.DenseSin(np.eye(6)).n == 6
    assert gs.DenseSin(np.eye(64) * 2.0).n == 64


def test_reuse_method_and_entry_checks(gs):
    """Newton(; reuse = k) crosses the ABI as gs_method.reuse; the constructor's refusals
    (newton.jl:60-90) and the entry point's refusal of reuse where no held factor exists (the
    least-squares batch) happen before any device call."""
    assert gs.Newton(reuse=3).to_c().reuse == 3 and gs.Newton().to_c().reuse == 0
    assert gs.Newton(reuse=2, keep_factor=True).to_c().reuse == 2
    with pytest.raises(gs.ArgumentError, match="reuse must be ≥ 1"):
        gs.Newton(reuse=0)
    with pytest.raises(gs.ArgumentError, match="LevenbergMarquardt globalization takes no reuse"):
        gs.Newton(globalization=gs.LevenbergMarquardt(), reuse=2)
    lib = gs.load_library()
    from importlib import import_module

    abi = import_module(gs.__name__ + "._abi")
    m = abi.GsMethod.from_buffer_copy(gs.Newton(globalization=gs.LevenbergMarquardt()).to_c())
    m.reuse = 2
    o = gs.Options().to_c()
    rc = lib.gs_solve_expsum_lsq_f64(ctypes.byref(m), 1, None, None, None, None, None, 0, None, ctypes.byref(o), None)
    assert rc == abi.GS_EINVAL and "no reuse" in abi.last_error(lib)
    fam = abi.GsFamily()
    fam.id, fam.n, fam.m, fam.p_len = abi.GS_FAM_FPU_GL3, 48, 48, 32
    rc = lib.gs_solve_batch_f64(ctypes.byref(gs.Newton(reuse=2).to_c()), ctypes.byref(fam), 1, None, None, 0, None,
                                None, None, ctypes.byref(o), 0, None)
    assert rc == abi.GS_EINVAL  # neither S nor codes: refused first, before the family checks



def test_dogleg_entry_checks(gs):
    """gs_solve_batch_f64 refuses a DogLeg radius outside the constructor's clamp and DogLeg with
    reuse (dogleg.jl:44-60, newton.jl:66-70) before any device call; the dense solve takes DogLeg at
    every size (its batch then reaches the pointer checks)."""
    lib = gs.load_library()
    from importlib import import_module

    abi = import_module(gs.__name__ + "._abi")
    fam = abi.GsFamily()
    fam.id, fam.n, fam.m, fam.p_len = abi.GS_FAM_COUPLED2, 2, 2, 2
    o = gs.Options().to_c()
    codes = (ctypes.c_uint8 * 1)()
    m = abi.GsMethod.from_buffer_copy(gs.Newton(globalization=gs.DogLeg()).to_c())
    assert m.globalization == abi.GS_DOGLEG and m.dl_radius == 0.0
    m.dl_radius = 1e-9
    rc = lib.gs_solve_batch_f64(ctypes.byref(m), ctypes.byref(fam), 1, None, None, 0, None, codes, None,
                                ctypes.byref(o), 0, None)
    assert rc == abi.GS_EINVAL and "radius" in abi.last_error(lib)
    m.dl_radius, m.reuse = 0.0, 2
    rc = lib.gs_solve_batch_f64(ctypes.byref(m), ctypes.byref(fam), 1, None, None, 0, None, codes, None,
                                ctypes.byref(o), 0, None)
    assert rc == abi.GS_EINVAL and "no reuse" in abi.last_error(lib)
    A = (ctypes.c_double * 256)()
    fam.id, fam.n, fam.m, fam.p_len, fam.data = abi.GS_FAM_DENSE_SIN, 16, 16, 16, ctypes.addressof(A)
    m.reuse = 0
    rc = lib.gs_solve_batch_f64(ctypes.byref(m), ctypes.byref(fam), 1, None, None, 0, None, codes, None,
                                ctypes.byref(o), 0, None)
    assert rc == abi.GS_EINVAL and "X must not be NULL" in abi.last_error(lib)


def test_float32_options_and_status_layout(gs):
    """Options(Float32) takes √eps, 2 eps and 1/10 of Float32 and holds every tolerance as its Float32
    value (options.jl:207-235); SolverStatus{Float32} is 44 bytes at Julia's offsets."""
    o = gs.Options(eltype="float32")
    f32 = np.float32
    assert o.f_reltol == float(np.sqrt(f32(np.finfo(f32).eps))) and o.x_reltol == 2 * float(np.finfo(f32).eps)
    assert o.kappa == float(f32(0.1)) and o.g_tol == o.f_reltol
    assert gs.Options(eltype="float32", f_abstol=1e-12).f_abstol == float(f32(1e-12))
    d = gs.STATUS_DTYPE_F32
    assert d.itemsize == 44
    assert [d.fields[n][1] for n in ("code", "f_calls", "iterations", "fnorm", "stepnorm", "step_failures",
                                     "promoted", "stationary", "jvp_calls", "rank", "nullnorm", "gradratio",
                                     "rate", "contracted", "reused")] == [0, 2, 4, 8, 12, 16, 20, 21, 22, 24, 28,
                                                                           32, 36, 40, 41]
    assert gs.status_buffer(5, "cpu", "float32").shape == (5, 44)



def test_wrapper_forms_have_no_host_path(gs):
    """The Deflated, StationarityProblem and Reparametrized batches refuse host tensors as the plain
    batch does: the library solves on the GPU and nothing falls back to the CPU."""
    import torch

    X2 = torch.zeros((4, 2), dtype=torch.float64)
    p2 = gs.PerSystem(torch.zeros((4, 2), dtype=torch.float64))
    prob = gs.NonlinearProblem(gs.Coupled2())
    with pytest.raises(gs.ArgumentError, match="device array"):
        gs.solve_batch(X2, gs.Deflated(prob, ((1.0, 2.0),)), gs.Newton(), p2)
    with pytest.raises(gs.ArgumentError, match="device array"):
        gs.solve_batch(X2, gs.Reparametrized(prob, gs.LinearChart(np.eye(2))), gs.Newton(), p2)
    X6 = torch.zeros((4, 6), dtype=torch.float64)
    with pytest.raises(gs.ArgumentError, match="device array"):
        gs.solve_batch(X6, gs.StationarityProblem(gs.HenonHeilesAction(), E=(3, 4), U=(5, 6)), gs.Newton(), (0.1, 1.0))
