"""Parity of the dense family (GS_FAM_DENSE_SIN: the §10.E problem A x + sin.(x)/10 − b of
benchmark/device_problems.jl:133-175; one lane per system at n = 2, 4, 6, 8, a wave segment per system
at n = 12, 16, 24, 32, 48, 64) with the CPU oracle,
through the C ABI. The device evaluates the family's fixed sine and cosine and folds every sum in the
oracle's order, so results are compared bit for bit: iterates and every status byte.

The batch mixes the benchmark's start x₀ = 0 with far starts (the line search and the iteration cap
work), starts that make F overflow and right-hand sides with NaN or Inf (NONFINITE), so each exit of
the loop runs on the device.
"""
import numpy as np
import pytest
import torch

from test_gpu_parity import assert_parity

pytestmark = pytest.mark.gpu

METHODS = {
    "static": lambda gs: gs.Newton(),
    "backtracking": lambda gs: gs.Newton(globalization=gs.Backtracking()),
    "mixed-refine": lambda gs: gs.Newton(linsolve=gs.LUFactorization("float32"), refinement=gs.Refinement()),
    "mixed-norefine": lambda gs: gs.Newton(linsolve=gs.LUFactorization("float32")),
    "mixed-backtracking-refine": lambda gs: gs.Newton(linsolve=gs.LUFactorization("float32"),
                                                      globalization=gs.Backtracking(), refinement=gs.Refinement()),
    "plain-refine-regularized": lambda gs: gs.Newton(refinement=gs.Refinement(), regularization=1e-3),
    "best-backtracking": lambda gs: gs.Newton(globalization=gs.Backtracking(), best=True),
    # reuse = k (newton.jl:1781-1874): the device grid's gs-reuse4 column, and held factors under
    # Backtracking and in the mixed form
    "reuse4": lambda gs: gs.Newton(reuse=4),
    "backtracking-reuse3": lambda gs: gs.Newton(globalization=gs.Backtracking(), reuse=3),
    "mixed-refine-reuse2": lambda gs: gs.Newton(linsolve=gs.LUFactorization("float32"), refinement=gs.Refinement(),
                                                reuse=2),
}


def _batch(workloads, n, nsys, seed):
    A, B, _ = workloads.tene_batch(n, nsys, seed=seed)
    rng = np.random.default_rng(seed + 100)
    X0 = np.zeros((nsys, n))
    far = rng.random(nsys) < 0.2
    X0[far] = rng.uniform(-30.0, 30.0, size=(int(far.sum()), n))
    huge = np.nonzero(rng.random(nsys) < 0.01)[0]
    X0[huge, 0] = 1e300
    bad = np.nonzero(rng.random(nsys) < 0.01)[0]
    B[bad, bad % n] = np.where(bad % 2 == 0, np.nan, np.inf)
    return A, B, X0


def _gpu(gs, device, fam, method, X0, B, opts, counts=False, order=None):
    X = torch.as_tensor(X0, device=device).clone()
    p = gs.PerSystem(torch.as_tensor(B, device=device))
    prob = gs.NonlinearProblem(fam)
    if counts:
        c = gs.solve_batch(X, prob, method, p, opts, order=order)
        return X.cpu().numpy(), c
    S = gs.status_buffer(X.shape[0], device)
    gs.solve_batch(X, S, prob, method, p, opts, order=order)
    return X.cpu().numpy(), gs.statuses(S)


@pytest.mark.parametrize("n", [2, 4, 6, 8, 12, 16, 24, 32, 48, 64])
@pytest.mark.parametrize("name", list(METHODS))
def test_dense_every_method_bitwise(gs, orc, workloads, device, n, name):
    nsys = 2048 if n < 32 else (1024 if n == 32 else 512)
    A, B, X0 = _batch(workloads, n, nsys, seed=n)
    fam = gs.DenseSin(A)
    m = METHODS[name](gs)
    for opts in (gs.Options(f_abstol=1e-12, f_reltol=0.0, max_iterations=40),
                 gs.Options(norm=gs.MaxNorm(), max_iterations=25)):
        print(f"dense n={n} {name} norm={type(opts.norm).__name__}", flush=True)
        x, st = _gpu(gs, device, fam, m, X0, B, opts)
        xo, so = orc.solve_batch(m, fam, X0.copy(), gs.PerSystem(B), opts)
        assert_parity(st, x, so, xo, bitwise=True)
    codes = np.bincount(st["code"], minlength=6)
    assert codes[0] > 0.7 * nsys and codes[4] > 0  # mostly SUCCESS, and the NONFINITE exits ran


@pytest.mark.parametrize("name", ["static", "backtracking", "mixed-refine", "mixed-backtracking-refine"])
def test_dense_nonmonotone_bitwise(gs, orc, workloads, device, name):
    """A weak diagonal (A = 0.02 I + R/20) leaves sin.(x)/10 in charge: many roots, far starts that
    wander, line searches that shorten, fail or stall, the iteration cap."""
    n, nsys = 16, 2048
    rng = np.random.default_rng(11)
    A = 0.02 * np.eye(n) + 0.05 * (workloads.tene_matrix(n, seed=12) - 2.0 * np.eye(n))
    B = rng.uniform(-0.1, 0.1, size=(nsys, n))
    X0 = rng.uniform(-20.0, 20.0, size=(nsys, n))
    fam = gs.DenseSin(A)
    m = METHODS[name](gs)
    opts = gs.Options(f_abstol=1e-12, f_reltol=0.0, max_iterations=60)
    x, st = _gpu(gs, device, fam, m, X0, B, opts)
    xo, so = orc.solve_batch(m, fam, X0.copy(), gs.PerSystem(B), opts)
    assert_parity(st, x, so, xo, bitwise=True)
    print(name, np.bincount(st["code"], minlength=6).tolist())


def test_dense_counts_form_and_order(gs, orc, workloads, device):
    n, nsys = 24, 3000
    A, B, X0 = _batch(workloads, n, nsys, seed=7)
    fam = gs.DenseSin(A)
    opts = gs.Options(f_abstol=1e-12, f_reltol=0.0)
    order = np.random.default_rng(8).permutation(nsys) + 1
    x, counts = _gpu(gs, device, fam, gs.Newton(), X0, B, opts, counts=True, order=order)
    xo, so = orc.solve_batch(gs.Newton(), fam, X0.copy(), gs.PerSystem(B), opts)
    assert np.array_equal(x.view(np.uint64), xo.view(np.uint64))
    want = np.bincount(so["code"], minlength=6)
    assert [counts[k] for k in ("SUCCESS", "STALLED", "MAXITERS", "SINGULAR", "NONFINITE", "LINESEARCH_FAILED")] \
        == want.tolist()


def test_dense_full_size_converges_to_the_built_roots(gs, workloads, device):
    """The benchmark's own batch from x₀ = 0 at 2^16 systems: every system converges to its root."""
    for n in (16, 32, 64):
        A, B, Xs = workloads.tene_batch(n, 1 << 16)
        x, st = _gpu(gs, device, gs.DenseSin(A), gs.Newton(), np.zeros((1 << 16, n)), B,
                     gs.Options(f_abstol=1e-12, f_reltol=0.0))
        assert (st["code"] == 0).all()
        assert np.abs(x - Xs).max() < 1e-12
        assert np.abs(workloads.tene_residual(x, A, B)).max() < 1e-12


@pytest.mark.parametrize("n", [12, 16, 24, 32])
def test_dense_lanes_per_system_agree_bitwise(gs, orc, workloads, device, n):
    """Every segment width a size takes (four, two or one systems per wave) writes the same bytes,
    those of the oracle; the width is the launch's choice alone."""
    A, B, X0 = _batch(workloads, n, 1000, seed=n + 1)
    fam = gs.DenseSin(A)
    m = METHODS["mixed-backtracking-refine"](gs)
    opts = gs.Options(f_abstol=1e-12, f_reltol=0.0, max_iterations=40)
    xo, so = orc.solve_batch(m, fam, X0.copy(), gs.PerSystem(B), opts)
    try:
        for lanes in gs.autotune.DENSE_LANES[n]:
            gs.dense_set_shape(n, lanes)
            assert gs.dense_get_shape(n) == lanes
            print(f"dense n={n} lanes={lanes}", flush=True)
            x, st = _gpu(gs, device, fam, m, X0, B, opts)
            assert_parity(st, x, so, xo, bitwise=True)
    finally:
        gs.dense_set_shape(n, 0)
    assert gs.dense_get_shape(n) == gs.autotune.DENSE_LANES[n][0]


def test_dense_tuner_sets_a_shape_that_changes_no_result(gs, workloads, device):
    n, nsys = 16, 1 << 14
    A, B, _ = workloads.tene_batch(n, nsys)
    fam = gs.DenseSin(A)
    prob = gs.NonlinearProblem(fam)
    p = gs.PerSystem(torch.as_tensor(B, device=device))
    opts = gs.Options(f_abstol=1e-12, f_reltol=0.0)
    X0 = torch.zeros((nsys, n), dtype=torch.float64, device=device)
    ref = _gpu(gs, device, fam, gs.Newton(), X0.cpu().numpy(), B, opts)
    try:
        res = gs.tune_dense(X0, prob, gs.Newton(), p, opts, repeats=1)
        assert res["lanes"] in gs.autotune.DENSE_LANES[n] and gs.dense_get_shape(n) == res["lanes"]
        assert set(res["seconds"]) == set(gs.autotune.DENSE_LANES[n])
        got = _gpu(gs, device, fam, gs.Newton(), X0.cpu().numpy(), B, opts)
        assert np.array_equal(got[0].view(np.uint64), ref[0].view(np.uint64))
        assert got[1].tobytes() == ref[1].tobytes()
    finally:
        gs.dense_set_shape(n, 0)
    assert (X0 == 0).all()  # the tuner solved copies


def test_workgroup_tuner_changes_no_result(gs, workloads, device):
    xk = workloads.henon_heiles_states(1 << 14, seed=9)
    Xk = torch.as_tensor(xk, device=device)
    prob = gs.NonlinearProblem(gs.HenonHeilesMidpoint())
    p = gs.BatchParams(0.1, Xk)
    opts = gs.Options(f_abstol=1e-12, f_reltol=0.0)
    res = gs.tune_workgroup(Xk, prob, gs.Newton(), p, opts, repeats=1)
    assert res["workgroupsize"] in (64, 128, 256) and set(res["seconds"]) == {64, 128, 256}
    X1, X2 = Xk.clone(), Xk.clone()
    S1, S2 = gs.status_buffer(Xk.shape[0], device), gs.status_buffer(Xk.shape[0], device)
    gs.solve_batch(X1, S1, prob, gs.Newton(), p, opts)
    gs.solve_batch(X2, S2, prob, gs.Newton(), p, opts, workgroupsize=res["workgroupsize"])
    assert torch.equal(X1, X2) and torch.equal(S1, S2)


F32_METHODS = {
    "static": lambda gs: gs.Newton(),
    "backtracking": lambda gs: gs.Newton(globalization=gs.Backtracking()),
    "reuse4": lambda gs: gs.Newton(reuse=4),
    "backtracking-reuse3": lambda gs: gs.Newton(globalization=gs.Backtracking(), reuse=3),
    "refine-regularized": lambda gs: gs.Newton(refinement=gs.Refinement(), regularization=1e-3),
    "best-backtracking": lambda gs: gs.Newton(globalization=gs.Backtracking(), best=True),
    # LUFactorization(Float32) on a Float32 batch is the working precision: the same solve
    "f32-factor": lambda gs: gs.Newton(linsolve=gs.LUFactorization("float32")),
}


@pytest.mark.parametrize("n", [2, 4, 6, 8, 12, 16, 24, 32, 48, 64])
@pytest.mark.parametrize("name", list(F32_METHODS))
def test_dense_float32_bitwise(gs, orc, workloads, device, n, name):
    """A Float32 batch of the dense family (the element type of the reference's device grid,
    benchmark/devices.jl:47): A, b and the start in Float32, the family's FP32 sine and cosine, the
    LU and every fold in Float32, SolverStatus{Float32}. Compared with the oracle's Float32 solve byte
    for byte (iterates and the 42 meaningful status bytes), under the grid's tolerance √eps(Float32)
    and under MaxNorm."""
    nsys = 2048 if n < 32 else (1024 if n == 32 else 512)
    A, B, X0 = _batch(workloads, n, nsys, seed=n + 1)
    # a start of 1e300 becomes 1e30: far beyond the others and finite in Float32
    A32, B32 = A.astype(np.float32), B.astype(np.float32)
    X32 = np.where(np.abs(X0) > 1e30, 1e30, X0).astype(np.float32)
    fam = gs.DenseSin(A32.astype(np.float64))
    m = F32_METHODS[name](gs)
    tol = float(np.sqrt(np.finfo(np.float32).eps))
    for opts in (gs.Options(eltype="float32", f_abstol=tol, f_reltol=0.0, max_iterations=40),
                 gs.Options(eltype="float32", norm=gs.MaxNorm(), max_iterations=25)):
        print(f"dense f32 n={n} {name} norm={type(opts.norm).__name__}", flush=True)
        X = torch.as_tensor(X32, device=device).clone()
        S = gs.status_buffer(nsys, device, "float32")
        gs.solve_batch(X, S, gs.NonlinearProblem(fam), m, gs.PerSystem(torch.as_tensor(B32, device=device)), opts)
        x, st = X.cpu().numpy(), gs.statuses(S)
        xo, so = orc.solve_batch(m, fam, X32.copy(), gs.PerSystem(B32), opts)
        assert x.dtype == np.float32 and st.dtype.itemsize == 44
        bad_x = np.nonzero((x.view(np.uint32) != xo.view(np.uint32)).any(axis=1))[0]
        sb = st.view(np.uint8).reshape(nsys, 44)[:, :42]
        ob = so.view(np.uint8).reshape(nsys, 44)[:, :42]
        bad_s = np.nonzero((sb != ob).any(axis=1))[0]
        msg = f"status {bad_s[0]}: {st[bad_s[0]]} vs {so[bad_s[0]]}" if bad_s.size else (
            f"x {bad_x[0]}" if bad_x.size else "")
        assert bad_s.size == 0 and bad_x.size == 0, f"{bad_s.size} statuses, {bad_x.size} iterates differ; {msg}"
    codes = np.bincount(st["code"], minlength=6)
    assert codes[0] > 0.6 * nsys and codes[4] > 0
    if "reuse" in name:
        assert st["reused"].any()


@pytest.mark.parametrize("n", [2, 4, 6, 8, 12, 16, 24, 32, 48, 64])
@pytest.mark.parametrize("eltype", ["float64", "float32"])
def test_dense_dogleg_bitwise(gs, orc, workloads, device, n, eltype):
    """DogLeg on the dense family at every size (the lane-per-system form at n ≤ 8, and 12 in Float32;
    the segment kernel above, whose Cauchy direction takes Jregᵀ F from the columns of A), in both
    precisions, with the radius unset and set small, and in FP64 the mixed form with refinement and a
    regularised form; against the oracle bit for bit."""
    nsys = 2048 if n < 32 else (1024 if n == 32 else 512)
    A, B, X0 = _batch(workloads, n, nsys, seed=n + 7)
    X0 = np.where(np.abs(X0) > 1e30, 1e30, X0)
    dt = np.float32 if eltype == "float32" else np.float64
    fam = gs.DenseSin(A.astype(dt).astype(np.float64))
    Bt, Xt = B.astype(dt), X0.astype(dt)
    tol = float(np.sqrt(np.finfo(np.float32).eps)) if eltype == "float32" else 1e-12
    methods = [gs.Newton(globalization=gs.DogLeg()), gs.Newton(globalization=gs.DogLeg(radius=1e-2))]
    if eltype == "float64":
        methods += [gs.Newton(globalization=gs.DogLeg(radius=2.0), linsolve=gs.LUFactorization("float32"),
                              refinement=gs.Refinement()),
                    gs.Newton(globalization=gs.DogLeg(), refinement=gs.Refinement(), regularization=1e-3)]
    for m in methods:
        opts = gs.Options(eltype=eltype, f_abstol=tol, f_reltol=0.0, max_iterations=60)
        X = torch.as_tensor(Xt, device=device).clone()
        S = gs.status_buffer(len(Xt), device, eltype)
        gs.solve_batch(X, S, gs.NonlinearProblem(fam), m, gs.PerSystem(torch.as_tensor(Bt, device=device)), opts)
        x, st = X.cpu().numpy(), gs.statuses(S)
        xo, so = orc.solve_batch(m, fam, Xt.copy(), gs.PerSystem(Bt), opts)
        w = 4 if eltype == "float32" else 8
        u = np.uint32 if eltype == "float32" else np.uint64
        size = st.dtype.itemsize
        meaningful = 42 if eltype == "float32" else 66
        bad_x = np.nonzero((x.view(u) != xo.view(u)).any(axis=1))[0]
        sb = st.view(np.uint8).reshape(len(st), size)[:, :meaningful]
        ob = so.view(np.uint8).reshape(len(so), size)[:, :meaningful]
        bad_s = np.nonzero((sb != ob).any(axis=1))[0]
        assert bad_s.size == 0 and bad_x.size == 0, f"{bad_s.size} statuses, {bad_x.size} iterates differ (w={w})"
        print(f"dense n={n} {eltype} dogleg", np.bincount(st["code"], minlength=6).tolist(), flush=True)
