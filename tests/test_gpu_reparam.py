"""Parity of the Reparametrized batch (src/nonlinear/reparametrized.jl:981-1031) with the CPU oracle,
through the C ABI (gs_solve_batch_reparametrized_f64).

The device composes each family's F and J with the chart (x = g(y), J̃ = J(g(y)) ∂g(y), each entry
a left fold) in the oracle's order, and writes x* = g(y*) to the batch: the linear chart on the
polynomial families is compared bit for bit; the polar chart and the tanh family go through
cos/sin/tanh, whose last bits differ between ROCm and glibc, and are held to the floating-point bar
(codes and iterations identical, solutions within 1e-10).
"""
import math

import numpy as np
import pytest
import torch

from test_gpu_parity import assert_parity
from test_oracle_reparam import R1

pytestmark = pytest.mark.gpu

TOL = dict(f_abstol=1e-12, f_reltol=0.0)


def _gpu(gs, device, rep, Y0, p, method, opts, counts=False, order=None):
    X = torch.as_tensor(Y0, device=device).clone()
    if isinstance(p, gs.PerSystem):
        pd = gs.PerSystem(torch.as_tensor(np.asarray(p.data), device=device))
    else:
        pd = gs.BatchParams(p.shared, torch.as_tensor(np.asarray(p.per_system), device=device))
    if counts:
        tally = gs.solve_batch(X, rep, method, pd, opts, order=order)
        return X.cpu().numpy(), tally
    S = gs.status_buffer(X.shape[0], device)
    gs.solve_batch(X, S, rep, method, pd, opts, order=order)
    return X.cpu().numpy(), gs.statuses(S)


METHODS = {
    "static": lambda gs: gs.Newton(),
    "backtracking": lambda gs: gs.Newton(globalization=gs.Backtracking()),
    "refine-regularized": lambda gs: gs.Newton(refinement=gs.Refinement(), regularization=1e-6),
    "mixed-refine": lambda gs: gs.Newton(linsolve=gs.LUFactorization("float32"), refinement=gs.Refinement()),
    "best": lambda gs: gs.Newton(best=True),
    "dogleg": lambda gs: gs.Newton(globalization=gs.DogLeg()),
    "mixed-dogleg-refine": lambda gs: gs.Newton(linsolve=gs.LUFactorization("float32"),
                                                globalization=gs.DogLeg(radius=2.0), refinement=gs.Refinement()),
    "reuse3": lambda gs: gs.Newton(reuse=3),
    "reuse2-backtracking-refine": lambda gs: gs.Newton(reuse=2, globalization=gs.Backtracking(),
                                                       refinement=gs.Refinement()),
    "mixed-reuse3-refine": lambda gs: gs.Newton(linsolve=gs.LUFactorization("float32"), reuse=3,
                                                refinement=gs.Refinement()),
}


def _linear_cases(gs, workloads, n):
    rng = np.random.default_rng(31)
    cases = []
    Y0 = rng.uniform(-12, 12, size=(n, 2))
    cases.append((gs.Coupled2(), Y0, gs.PerSystem(np.tile([10.0, 10.0], (n, 1))),
                  np.array([[1.5, 0.7], [-0.4, 0.9]]), np.array([0.3, -1.1])))
    cases.append((gs.Hetero2(), rng.uniform(-3, 3, size=(n, 2)), gs.PerSystem(rng.integers(0, 6, size=(n, 1))),
                  np.array([[2.0, 0.0], [0.5, 1.0]]), np.zeros(2)))
    cases.append((gs.ScalarSquare(), rng.uniform(0.2, 3, size=(n, 1)),
                  gs.PerSystem(rng.uniform(0.5, 4, size=(n, 1))), np.array([[0.5]]), np.array([0.25])))
    xk = workloads.henon_heiles_states(n, seed=32)
    B4 = np.eye(4) + 0.1 * rng.standard_normal((4, 4))
    cases.append((gs.HenonHeilesMidpoint(), np.ascontiguousarray(np.linalg.solve(B4, xk.T).T), gs.BatchParams(0.1, xk), B4, np.zeros(4)))
    return cases


@pytest.mark.parametrize("name", list(METHODS))
def test_linear_chart_bitwise(gs, orc, workloads, device, name):
    method = METHODS[name](gs)
    for fam, Y0, p, B, c in _linear_cases(gs, workloads, 4096):
        rep = gs.Reparametrized(gs.NonlinearProblem(fam), gs.LinearChart(B, c))
        for opts in (gs.Options(max_iterations=40, **TOL), gs.Options(norm=gs.MaxNorm(), max_iterations=40, **TOL)):
            x, st = _gpu(gs, device, rep, Y0, p, method, opts)
            xo, so = orc.solve_batch(method, fam, Y0.copy(), p, opts, reparametrized=rep)
            assert_parity(st, x, so, xo, bitwise=True)
        print(name, type(fam).__name__, np.bincount(st["code"], minlength=6).tolist())


def _tolerance(gs, st, x, so, xo):
    """Codes identical; a converged system may take one iteration more or fewer where its residual
    crosses the tolerance at the last bit (at most 1% of them); solutions within 1e-10."""
    assert (st["code"] == so["code"]).all()
    ok = so["code"] == gs.SUCCESS
    di = np.abs(st["iterations"][ok].astype(int) - so["iterations"][ok].astype(int))
    assert di.max() <= 1 and (di != 0).mean() <= 0.01, (di.max(), (di != 0).sum())
    scale = np.maximum(np.abs(xo[ok]).max(axis=1, keepdims=True), 1.0)
    assert (np.abs(x[ok] - xo[ok]) <= 1e-10 * scale).all()
    return ok


def test_polar_chart_and_tanh_to_tolerance(gs, orc, workloads, device):
    """The polar chart of the reference's tests on the coupled family (starts in the basin of the
    root (5.798, 5.798)), and the tanh family under a linear chart."""
    n = 4096
    rng = np.random.default_rng(33)
    r0 = math.hypot(*R1)
    Y0 = np.column_stack([r0 + rng.uniform(-1.5, 1.5, n), math.pi / 4 + rng.uniform(-0.3, 0.3, n)])
    p = gs.PerSystem(np.tile([10.0, 10.0], (n, 1)))
    rep = gs.Reparametrized(gs.NonlinearProblem(gs.Coupled2()), gs.PolarChart())
    for m in (gs.Newton(), gs.Newton(globalization=gs.Backtracking()), gs.Newton(globalization=gs.DogLeg()),
              gs.Newton(reuse=3, refinement=gs.Refinement())):
        opts = gs.Options(max_iterations=40, **TOL)
        x, st = _gpu(gs, device, rep, Y0, p, m, opts)
        xo, so = orc.solve_batch(m, gs.Coupled2(), Y0.copy(), p, opts, reparametrized=rep)
        ok = _tolerance(gs, st, x, so, xo)
        assert ok.mean() > 0.9
    rec, X0 = workloads.tanh_linear3_batch(2048, 34)
    B = np.eye(3) + 0.2 * rng.standard_normal((3, 3))
    rep = gs.Reparametrized(gs.NonlinearProblem(gs.TanhLinear3()), gs.LinearChart(B))
    Y0 = np.ascontiguousarray(np.linalg.solve(B, X0.T).T)
    x, st = _gpu(gs, device, rep, Y0, gs.PerSystem(rec), gs.Newton(), gs.Options(**TOL))
    xo, so = orc.solve_batch(gs.Newton(), gs.TanhLinear3(), Y0.copy(), gs.PerSystem(rec), gs.Options(**TOL),
                             reparametrized=rep)
    assert _tolerance(gs, st, x, so, xo).all()


def test_counts_form_and_order(gs, orc, device):
    n = 2048
    rng = np.random.default_rng(35)
    Y0 = rng.uniform(-12, 12, size=(n, 2))
    p = gs.PerSystem(np.tile([10.0, 10.0], (n, 1)))
    rep = gs.Reparametrized(gs.NonlinearProblem(gs.Coupled2()), gs.LinearChart(2 * np.eye(2), [1.0, -1.0]))
    order = rng.permutation(n) + 1
    x, counts = _gpu(gs, device, rep, Y0, p, gs.Newton(), gs.Options(**TOL), counts=True, order=order)
    xo, so = orc.solve_batch(gs.Newton(), gs.Coupled2(), Y0.copy(), p, gs.Options(**TOL), reparametrized=rep)
    assert (x.view(np.uint64) == xo.view(np.uint64)).all()
    assert [counts[k] for k in ("SUCCESS", "STALLED", "MAXITERS", "SINGULAR", "NONFINITE", "LINESEARCH_FAILED")] \
        == np.bincount(so["code"], minlength=6).tolist()


def test_refusals(gs, device):
    X = torch.zeros((4, 3), dtype=torch.float64, device=device)
    p = gs.PerSystem(torch.zeros((4, 12), dtype=torch.float64, device=device))
    rep = gs.Reparametrized(gs.NonlinearProblem(gs.TanhLinear3()), gs.PolarChart())
    with pytest.raises(gs.ArgumentError, match="polar chart maps 2"):
        gs.solve_batch(X, rep, gs.Newton(), p)
    with pytest.raises(gs.ArgumentError, match="takes no reuse"):
        gs.solve_batch(X, gs.Reparametrized(gs.NonlinearProblem(gs.TanhLinear3()), gs.LinearChart(np.eye(3))),
                       gs.Newton(globalization=gs.DogLeg(), reuse=2), p)


F32_METHODS = {
    "static": lambda gs: gs.Newton(),
    "backtracking": lambda gs: gs.Newton(globalization=gs.Backtracking()),
    "dogleg": lambda gs: gs.Newton(globalization=gs.DogLeg()),
    "refine": lambda gs: gs.Newton(refinement=gs.Refinement()),
    "reuse3": lambda gs: gs.Newton(reuse=3),
    "best": lambda gs: gs.Newton(best=True),
}


@pytest.mark.parametrize("name", list(F32_METHODS))
def test_float32_linear_chart_bitwise(gs, orc, workloads, device, name):
    """gs_solve_batch_reparametrized_f32: the linear-chart cases of the Float64 test with Float32
    starts and records (the chart's B and c rounded to Float32 on both sides), bit for bit against the
    oracle's Float32 solve; the polar chart is refused in Float32."""
    method = F32_METHODS[name](gs)
    for fam, Y0, p, B, c in _linear_cases(gs, workloads, 4096):
        Y0 = Y0.astype(np.float32)
        if isinstance(p, gs.PerSystem):
            data = np.asarray(p.data)
            p = gs.PerSystem(data if data.dtype.kind == "i" else data.astype(np.float32))
        else:
            p = gs.BatchParams(p.shared, np.asarray(p.per_system).astype(np.float32))
        rep = gs.Reparametrized(gs.NonlinearProblem(fam), gs.LinearChart(B, c))
        for opts in (gs.Options(eltype="float32", f_abstol=1e-5, max_iterations=40),
                     gs.Options(eltype="float32", norm=gs.MaxNorm(), f_abstol=1e-5, max_iterations=40)):
            X = torch.as_tensor(Y0, device=device).clone()
            if isinstance(p, gs.PerSystem):
                pd = gs.PerSystem(torch.as_tensor(np.asarray(p.data), device=device))
            else:
                pd = gs.BatchParams(p.shared, torch.as_tensor(np.asarray(p.per_system), device=device))
            S = gs.status_buffer(X.shape[0], device, "float32")
            gs.solve_batch(X, S, rep, method, pd, opts)
            x, st = X.cpu().numpy(), gs.statuses(S)
            xo, so = orc.solve_batch(method, fam, Y0.copy(), p, opts, reparametrized=rep)
            assert np.array_equal(x.view(np.uint32), xo.view(np.uint32)), type(fam).__name__
            sb = st.view(np.uint8).reshape(len(st), 44)[:, :42]
            ob = so.view(np.uint8).reshape(len(so), 44)[:, :42]
            assert (sb == ob).all(), type(fam).__name__
        print(name, type(fam).__name__, np.bincount(st["code"], minlength=6).tolist())
    X = torch.zeros((4, 2), dtype=torch.float32, device=device)
    polar = gs.Reparametrized(gs.NonlinearProblem(gs.Coupled2()), gs.PolarChart())
    with pytest.raises(gs.ArgumentError, match="linear chart"):
        gs.solve_batch(X, polar, method, gs.PerSystem(torch.ones((4, 2), dtype=torch.float32, device=device)),
                       gs.Options(eltype="float32"))
