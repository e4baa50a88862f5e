"""Parity of the wave-per-system solve (the FPU-β Gauss–Legendre s = 3 stage system, n = 48) with the
CPU oracle, through the C ABI.

The device follows the oracle's operation order in both factor precisions (the unblocked LU in
FP64 or FP32, the left folds, the Jacobian products over the same entries in the same order), and
the results are compared bit for bit; the mixed method is also held to the floating-point bar
(codes, iterations and promotions identical, iterates within 1e-10 relative). Rows that are not
tame (a non-finite or −0 entry, or one beyond 2^(emax−48)) go to the second kernel, whose LU
selects every entry per lane; the last tests force that path.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

H = 0.05
TOL = dict(f_abstol=1e-12, f_reltol=0.0)


def _fields_equal(st, so, names):
    bad = np.zeros(len(st), dtype=bool)
    for n in names:
        a, b = np.ascontiguousarray(st[n]), np.ascontiguousarray(so[n])
        ua = a.view({1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[a.dtype.itemsize])
        ub = b.view({1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[b.dtype.itemsize])
        bad |= ua != ub
    return bad


def _run(gs, device, method, QP, Z0, opts):
    Z = torch.as_tensor(Z0, device=device).clone()
    S = gs.status_buffer(len(Z0), device)
    gs.solve_batch(Z, S, gs.NonlinearProblem(gs.FpuGl3()), method, gs.BatchParams(H, torch.as_tensor(QP, device=device)),
                   opts)
    torch.cuda.synchronize()
    return Z.cpu().numpy(), gs.statuses(S)


def _oracle(gs, orc, method, QP, Z0, opts):
    return orc.solve_batch(method, gs.FpuGl3(), np.array(Z0, dtype=float), gs.BatchParams(H, QP), opts)


def _assert_bitwise(z, st, zo, so):
    bad_x = np.nonzero((z.view(np.uint64) != zo.view(np.uint64)).any(axis=1))[0]
    bad_s = np.nonzero(_fields_equal(st, so, st.dtype.names))[0]
    first = ""
    if bad_s.size:
        first = f"status {bad_s[0]}: device {st[bad_s[0]]}, oracle {so[bad_s[0]]}"
    elif bad_x.size:
        i = bad_x[0]
        first = f"z {i}: max |Δ| {np.abs(z[i] - zo[i]).max():.3e}"
    assert bad_s.size == 0 and bad_x.size == 0, f"{bad_s.size} statuses, {bad_x.size} iterates differ; {first}"


@pytest.mark.parametrize("case", ["default", "min_iterations", "best_maxiter1", "regularized", "maxnorm", "refined"])
def test_fp64_bitwise(gs, orc, workloads, device, case):
    QP = workloads.fpu_states(512, seed=3)
    Z0 = np.zeros((512, 48))
    method = gs.Newton()
    opts = gs.Options(**TOL)
    if case == "min_iterations":
        opts = gs.Options(min_iterations=3, **TOL)
    elif case == "best_maxiter1":
        method = gs.Newton(best=True)
        opts = gs.Options(max_iterations=1, **TOL)
    elif case == "regularized":
        method = gs.Newton(regularization=1e-3)
    elif case == "maxnorm":
        opts = gs.Options(norm=gs.MaxNorm(), f_abstol=1e-13, f_reltol=0.0)
    elif case == "refined":
        method = gs.Newton(refinement=gs.Refinement())
    z, st = _run(gs, device, method, QP, Z0, opts)
    zo, so = _oracle(gs, orc, method, QP, Z0, opts)
    _assert_bitwise(z, st, zo, so)
    print(case, np.bincount(st["code"], minlength=6).tolist(), np.bincount(st["iterations"]).tolist())


def test_fp64_hostile_bitwise(gs, orc, workloads, device):
    """Large amplitudes (strongly nonlinear stages), NaN and Inf records, and nonzero starts."""
    QP = workloads.fpu_states(256, seed=4)
    Z0 = np.zeros((256, 48))
    QP[0:32] *= 10.0
    QP[32, 3] = np.nan
    QP[33, 20] = np.inf
    Z0[40:80] = np.random.default_rng(0).standard_normal((40, 48)) * 0.5
    opts = gs.Options(**TOL)
    z, st = _run(gs, device, gs.Newton(), QP, Z0, opts)
    zo, so = _oracle(gs, orc, gs.Newton(), QP, Z0, opts)
    _assert_bitwise(z, st, zo, so)
    assert st["code"][32] == 4 and st["code"][33] == 4


@pytest.mark.parametrize("refine", [True, False])
def test_mixed_against_oracle(gs, orc, workloads, device, refine):
    QP = workloads.fpu_states(512, seed=5)
    Z0 = np.zeros((512, 48))
    method = gs.Newton(linsolve=gs.LUFactorization(np.float32),
                       refinement=gs.Refinement() if refine else gs.NoRefinement())
    opts = gs.Options(**TOL)
    z, st = _run(gs, device, method, QP, Z0, opts)
    zo, so = _oracle(gs, orc, method, QP, Z0, opts)
    _assert_bitwise(z, st, zo, so)
    for f in ("code", "iterations", "promoted", "step_failures"):
        assert (st[f] == so[f]).all(), (f, np.nonzero(st[f] != so[f])[0][:8])
    assert np.abs(z - zo).max() <= 1e-10 * np.abs(zo).max()
    for f in ("fnorm", "gradratio"):
        a, b = st[f], so[f]
        floor = (np.abs(a) < 1e-12) & (np.abs(b) < 1e-12)
        ok = floor | np.isclose(a, b, rtol=1e-6, atol=0)
        assert ok.all(), (f, np.nonzero(~ok)[0][:8])
    print("mixed", refine, np.bincount(st["code"], minlength=6).tolist(), np.bincount(st["iterations"]).tolist(),
          "promoted", int(st["promoted"].sum()))


def test_full_batch_and_residual_recheck(gs, orc, workloads, device):
    """The 65536-system batch in both forms: every system converges, its final residual (a host
    recheck of the stage equations) is at the tolerance, the two forms agree to 1e-10, and a sample
    of the FP64 form matches the oracle bit for bit."""
    QP = workloads.fpu_states(65536, seed=0)
    Z0 = np.zeros((65536, 48))
    opts = gs.Options(**TOL)
    z64, s64 = _run(gs, device, gs.Newton(), QP, Z0, opts)
    mixed = gs.Newton(linsolve=gs.LUFactorization(np.float32), refinement=gs.Refinement())
    z32, s32 = _run(gs, device, mixed, QP, Z0, opts)
    assert (s64["code"] == 0).all() and (s32["code"] == 0).all()
    assert np.abs(z64 - z32).max() <= 1e-10 * np.abs(z64).max()
    res = np.abs(workloads.fpu_gl3_residual(z64, QP, H)).max(axis=1)
    assert res.max() <= 1e-11
    sample = np.random.default_rng(3).choice(65536, 1024, replace=False)
    zo, so = _oracle(gs, orc, gs.Newton(), QP[sample], Z0[sample], opts)
    _assert_bitwise(z64[sample], s64[sample], zo, so)


@pytest.mark.parametrize("case", ["huge_regularization", "f32_negative_zero"])
def test_exact_kernel_bitwise(gs, orc, workloads, device, case):
    """Systems whose rows are not tame are solved by the second kernel: a regularisation of 1e300
    puts entries beyond 2^975 on the diagonal (FP64 rows), and h = 1e-25 rounds the off-diagonal
    entries of the FP32 rows to −0. Both are compared with the oracle bit for bit."""
    QP = workloads.fpu_states(128, seed=6)
    Z0 = np.random.default_rng(1).standard_normal((128, 48)) * 0.1
    opts = gs.Options(max_iterations=5, **TOL)
    h = H
    if case == "huge_regularization":
        method = gs.Newton(regularization=1e300)
    else:
        method = gs.Newton(linsolve=gs.LUFactorization(np.float32), refinement=gs.Refinement())
        h = 1e-25
    Z = torch.as_tensor(Z0, device=device).clone()
    S = gs.status_buffer(len(Z0), device)
    prob = gs.NonlinearProblem(gs.FpuGl3())
    gs.solve_batch(Z, S, prob, method, gs.BatchParams(h, torch.as_tensor(QP, device=device)), opts)
    torch.cuda.synchronize()
    z, st = Z.cpu().numpy(), gs.statuses(S)
    zo, so = orc.solve_batch(method, gs.FpuGl3(), np.array(Z0), gs.BatchParams(h, QP), opts)
    _assert_bitwise(z, st, zo, so)
    assert (st["iterations"] > 0).all()
    print(case, np.bincount(st["code"], minlength=6).tolist(), np.bincount(st["iterations"]).tolist())


@pytest.mark.parametrize("case", ["exact_slope", "measured_slope", "refined", "mixed", "hostile", "short_search"])
def test_backtracking_bitwise(gs, orc, workloads, device, case):
    """Backtracking (backtracking.jl:354-390) on the wave-per-system solve: the plain method takes
    ExactStep's slope −2φ₀, a regularised or refined one MeasuredSlope's 2⟨F, J d⟩, the mixed one
    switches at its promotion (newton.jl:1244-1254). Half of the records have 40 times the amplitude,
    where the search shortens some steps (their iterates and iteration counts then differ from
    Static's); `hostile` adds non-finite records and nonzero starts, `short_search` caps the search at
    2 trips with a steep sufficient-decrease constant, so searches end at LINESEARCH_FAILED. Every
    case is compared with the oracle bit for bit."""
    rng = np.random.default_rng(7)
    QP = workloads.fpu_states(256, seed=8)
    QP[0:128] *= 40.0
    Z0 = np.zeros((256, 48))
    bt = gs.Backtracking()
    opts = gs.Options(max_iterations=40, **TOL)
    if case == "exact_slope":
        method = gs.Newton(globalization=bt)
    elif case == "measured_slope":
        method = gs.Newton(globalization=bt, regularization=1e-3)
    elif case == "refined":
        method = gs.Newton(globalization=bt, refinement=gs.Refinement())
    elif case == "mixed":
        method = gs.Newton(globalization=bt, linsolve=gs.LUFactorization(np.float32), refinement=gs.Refinement())
    elif case == "hostile":
        method = gs.Newton(globalization=bt)
        QP[200, 3] = np.nan
        QP[201, 20] = np.inf
        Z0[64:192] = rng.standard_normal((128, 48)) * 2.0
    else:
        method = gs.Newton(globalization=gs.Backtracking(c1=0.9, maxiter=2))
        Z0[128:] = rng.standard_normal((128, 48)) * 0.8
    z, st = _run(gs, device, method, QP, Z0, opts)
    zo, so = _oracle(gs, orc, method, QP, Z0, opts)
    _assert_bitwise(z, st, zo, so)
    if case == "exact_slope":
        zs, ss = _oracle(gs, orc, gs.Newton(), QP, Z0, opts)
        assert (ss["iterations"] != so["iterations"]).any() and (zs != zo).any(axis=1).any()
    print(case, np.bincount(st["code"], minlength=8).tolist(), np.bincount(st["iterations"]).tolist(),
          "failures", int(st["step_failures"].sum()), "promoted", int(st["promoted"].sum()))


@pytest.mark.parametrize("glob", ["static", "backtracking"])
@pytest.mark.parametrize("k", [2, 3])
@pytest.mark.parametrize("form", ["plain", "refined", "mixed", "regularized"])
def test_reuse_bitwise(gs, orc, workloads, device, glob, k, form):
    """reuse = k (newton.jl:1781-1874) on the wave-per-system solve: the factor's rows stay in
    registers across steps; a step along them refines with the JVP at x, a failed refinement or a
    search that fails without a step refactorises, a step of rate ≥ 1 refreshes the factor at the
    next step. Half of the records have 40 times the amplitude and half of the starts are nonzero,
    so steps along the held factor fail, stall and refresh. Compared with the oracle bit for bit,
    the `reused` flag included."""
    QP = workloads.fpu_states(256, seed=9)
    QP[0:128] *= 40.0
    Z0 = np.zeros((256, 48))
    Z0[64:192] = np.random.default_rng(2).standard_normal((128, 48)) * 0.5
    kw = dict(globalization=gs.Backtracking() if glob == "backtracking" else gs.Static(), reuse=k)
    if form == "refined":
        kw["refinement"] = gs.Refinement()
    elif form == "mixed":
        kw.update(linsolve=gs.LUFactorization(np.float32), refinement=gs.Refinement())
    elif form == "regularized":
        kw["regularization"] = 1e-3
    method = gs.Newton(**kw)
    opts = gs.Options(max_iterations=60, **TOL)
    z, st = _run(gs, device, method, QP, Z0, opts)
    zo, so = _oracle(gs, orc, method, QP, Z0, opts)
    _assert_bitwise(z, st, zo, so)
    assert st["reused"].any()
    print(glob, k, form, np.bincount(st["code"], minlength=8).tolist(), "failures", int(st["step_failures"].sum()),
          "reused", int(st["reused"].sum()))


def test_reuse_exact_kernel_bitwise(gs, orc, workloads, device):
    """reuse = 2 on systems the second kernel solves (rows beyond 2^975 from a regularisation of
    1e300; FP32 rows with −0 entries at h = 1e-25), against the oracle bit for bit."""
    QP = workloads.fpu_states(128, seed=6)
    Z0 = np.random.default_rng(1).standard_normal((128, 48)) * 0.1
    opts = gs.Options(max_iterations=5, **TOL)
    for h, method in ((H, gs.Newton(regularization=1e300, reuse=2)),
                      (1e-25, gs.Newton(linsolve=gs.LUFactorization(np.float32), refinement=gs.Refinement(),
                                        reuse=2))):
        Z = torch.as_tensor(Z0, device=device).clone()
        S = gs.status_buffer(len(Z0), device)
        gs.solve_batch(Z, S, gs.NonlinearProblem(gs.FpuGl3()), method,
                       gs.BatchParams(h, torch.as_tensor(QP, device=device)), opts)
        torch.cuda.synchronize()
        zo, so = orc.solve_batch(method, gs.FpuGl3(), np.array(Z0), gs.BatchParams(h, QP), opts)
        _assert_bitwise(Z.cpu().numpy(), gs.statuses(S), zo, so)


@pytest.mark.parametrize("case", ["default", "radius", "wide_radius", "refined", "mixed", "regularized", "hostile"])
def test_dogleg_bitwise(gs, orc, workloads, device, case):
    """DogLeg (dogleg.jl:116-206, newton.jl:1725-1780) on the wave-per-system solve: the Cauchy
    direction from Jregᵀ F, the Newton step inside the radius, the Cauchy point or the dogleg segment
    point, the radius updates, and the limited step that passes only the residual test. A small
    initial radius (1e-3) forces Cauchy and segment steps at every start; a wide one (1e10) forces
    rejected trials first. Compared with the oracle bit for bit."""
    QP = workloads.fpu_states(256, seed=10)
    QP[0:128] *= 40.0
    Z0 = np.zeros((256, 48))
    Z0[64:192] = np.random.default_rng(3).standard_normal((128, 48)) * 0.5
    kw = dict(globalization=gs.DogLeg())
    if case == "radius":
        kw["globalization"] = gs.DogLeg(radius=1e-3)
    elif case == "wide_radius":
        kw["globalization"] = gs.DogLeg(radius=1e10)
        Z0[0:64] = np.random.default_rng(4).standard_normal((64, 48)) * 3.0
    elif case == "refined":
        kw["refinement"] = gs.Refinement()
    elif case == "mixed":
        kw.update(linsolve=gs.LUFactorization(np.float32), refinement=gs.Refinement())
    elif case == "regularized":
        kw["regularization"] = 1e-3
    elif case == "hostile":
        QP[200, 3] = np.nan
        Z0[0:64] = np.random.default_rng(4).standard_normal((64, 48)) * 3.0
    method = gs.Newton(**kw)
    opts = gs.Options(max_iterations=60, **TOL)
    z, st = _run(gs, device, method, QP, Z0, opts)
    zo, so = _oracle(gs, orc, method, QP, Z0, opts)
    _assert_bitwise(z, st, zo, so)
    print(case, np.bincount(st["code"], minlength=8).tolist(), "iterations", float(st["iterations"].mean()),
          "failures", int(st["step_failures"].sum()))
