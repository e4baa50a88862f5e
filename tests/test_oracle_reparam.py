"""The oracle's Reparametrized batch (src/nonlinear/reparametrized.jl, covariant = false) over the
registered charts, pinned by what the mathematics fixes:

- the reference's batch KAT (test/nonlinear/reparametrized.jl:383-412): F(x) = x − c under the chart
  x = 2y reaches x = c to 8 eps in every method form;
- affine invariance of Newton's method: under x = B y + c with B invertible the Newton iterates in y
  map onto those in x, so the reparametrized solve from y₀ = B⁻¹(x₀ − c) takes the plain solve's
  iterations and reaches its solution (a transposed or dropped chart Jacobian breaks both);
- the polar chart of the reference's tests reaches the closed-form roots of the coupled family, and
  at r = 0 (a rank-one chart Jacobian) the solve reports SINGULAR.
"""
import math

import numpy as np
import pytest

TOL = dict(f_abstol=1e-12, f_reltol=0.0)
R1 = (-4 + math.sqrt(96), -4 + math.sqrt(96))  # a root of the coupled family at p = (10, 10)


def test_reference_batch_kat(gs, orc):
    """x − c with c = (2, 3) (the Hetero2 fixture of code 1) under x = 2y, from y = (1, 1) and (0, 0)."""
    fam = gs.Hetero2()
    rep = gs.Reparametrized(gs.NonlinearProblem(fam), gs.LinearChart(2 * np.eye(2)))
    for m in (gs.Newton(), gs.Newton(globalization=gs.Backtracking()), gs.Newton(refinement=gs.Refinement()),
              gs.Newton(linsolve=gs.LUFactorization("float32"), refinement=gs.Refinement())):
        x, s = orc.solve_batch(m, fam, np.array([[1.0, 1.0], [0.0, 0.0]]), gs.PerSystem(np.array([[1], [1]])),
                               gs.Options(**TOL), reparametrized=rep)
        assert (s["code"] == gs.SUCCESS).all()
        assert np.allclose(x, [[2.0, 3.0], [2.0, 3.0]], rtol=8 * np.finfo(float).eps, atol=0)


def test_affine_invariance(gs, orc):
    rng = np.random.default_rng(5)
    n = 512
    fam = gs.Coupled2()
    p = gs.PerSystem(np.tile([10.0, 10.0], (n, 1)))
    X0 = np.array(R1) + rng.uniform(-1.5, 1.5, size=(n, 2))
    B = np.array([[1.5, 0.7], [-0.4, 0.9]])
    c = np.array([0.3, -1.1])
    Y0 = np.linalg.solve(B, (X0 - c).T).T
    for m in (gs.Newton(), gs.Newton(globalization=gs.Backtracking())):
        xp, sp = orc.solve_batch(m, fam, X0.copy(), p, gs.Options(**TOL))
        rep = gs.Reparametrized(gs.NonlinearProblem(fam), gs.LinearChart(B, c))
        x, s = orc.solve_batch(m, fam, Y0.copy(), p, gs.Options(**TOL), reparametrized=rep)
        assert (s["code"] == gs.SUCCESS).all() and (sp["code"] == gs.SUCCESS).all()
        assert (s["iterations"] == sp["iterations"]).all()
        assert np.abs(x - xp).max() <= 1e-12 * np.abs(xp).max()


def test_polar_chart_roots_and_singular_origin(gs, orc):
    fam = gs.Coupled2()
    rep = gs.Reparametrized(gs.NonlinearProblem(fam), gs.PolarChart())
    rng = np.random.default_rng(6)
    r0 = math.hypot(*R1)
    Y0 = np.column_stack([r0 + rng.uniform(-0.8, 0.8, 256), math.pi / 4 + rng.uniform(-0.2, 0.2, 256)])
    p = gs.PerSystem(np.tile([10.0, 10.0], (256, 1)))
    x, s = orc.solve_batch(gs.Newton(), fam, Y0.copy(), p, gs.Options(**TOL), reparametrized=rep)
    assert (s["code"] == gs.SUCCESS).all() and (s["iterations"] <= 6).all()
    assert np.abs(x - np.array(R1)).max() <= 1e-12 * R1[0]
    x0, s0 = orc.solve_batch(gs.Newton(), fam, np.array([[0.0, 0.3]]), gs.PerSystem(np.array([[10.0, 10.0]])),
                             gs.Options(**TOL), reparametrized=rep)
    assert s0["code"][0] == gs.SINGULAR


def test_reparametrized_checks(gs):
    prob = gs.NonlinearProblem(gs.Coupled2())
    rep = gs.Reparametrized(prob, gs.PolarChart())
    with pytest.raises(gs.ArgumentError, match="cannot wrap a Reparametrized"):
        gs.Reparametrized(rep, gs.PolarChart())
    with pytest.raises(gs.ArgumentError, match="covariant = false"):
        gs.Reparametrized(prob, gs.PolarChart(), covariant=True)
    with pytest.raises(gs.ArgumentError, match="registered chart"):
        gs.Reparametrized(prob, lambda y, p: y)
    with pytest.raises(gs.ArgumentError, match="square B"):
        gs.LinearChart(np.zeros((2, 3)))
    with pytest.raises(gs.ArgumentError, match="polar chart maps 2"):
        gs.Reparametrized(gs.NonlinearProblem(gs.TanhLinear3()), gs.PolarChart()).to_c(3)
    with pytest.raises(gs.DimensionMismatch):
        gs.Reparametrized(prob, gs.LinearChart(np.eye(3))).to_c(2)


def test_float32_linear_chart(gs, orc):
    """The Float32 batch with the linear chart: the reference's batch KAT in Float32 (x − c under
    x = 2y reaches c to 8 eps(Float32)), and affine invariance against the plain Float32 solve (the
    same iteration counts, the same solution to Float32 rounding); the polar chart has no Float32 form."""
    fam = gs.Hetero2()
    rep = gs.Reparametrized(gs.NonlinearProblem(fam), gs.LinearChart(2 * np.eye(2)))
    o32 = gs.Options(eltype="float32", f_abstol=1e-6)
    for m in (gs.Newton(), gs.Newton(globalization=gs.Backtracking()), gs.Newton(refinement=gs.Refinement())):
        x, s = orc.solve_batch(m, fam, np.array([[1.0, 1.0], [0.0, 0.0]], dtype=np.float32),
                               gs.PerSystem(np.array([[1], [1]])), o32, reparametrized=rep)
        assert x.dtype == np.float32 and (s["code"] == gs.SUCCESS).all()
        assert np.allclose(x, [[2.0, 3.0], [2.0, 3.0]], rtol=8 * np.finfo(np.float32).eps, atol=0)
    rng = np.random.default_rng(6)
    n = 512
    fam = gs.Coupled2()
    p = gs.PerSystem(np.tile([10.0, 10.0], (n, 1)).astype(np.float32))
    X0 = (np.array(R1) + rng.uniform(-1.5, 1.5, size=(n, 2))).astype(np.float32)
    B = np.array([[1.0, 0.5], [0.0, 2.0]])  # exact in Float32, so the chart is the same map
    c = np.array([0.25, -1.0])
    Y0 = np.linalg.solve(B, (X0.astype(np.float64) - c).T).T.astype(np.float32)
    for m in (gs.Newton(), gs.Newton(globalization=gs.Backtracking())):
        xp, sp = orc.solve_batch(m, fam, X0.copy(), p, o32)
        rep = gs.Reparametrized(gs.NonlinearProblem(fam), gs.LinearChart(B, c))
        x, s = orc.solve_batch(m, fam, Y0.copy(), p, o32, reparametrized=rep)
        assert (s["code"] == gs.SUCCESS).all() and (sp["code"] == gs.SUCCESS).all()
        assert np.abs(s["iterations"].astype(int) - sp["iterations"].astype(int)).max() <= 1
        assert np.abs(x - xp).max() <= 1e-5 * np.abs(xp).max()
    polar = gs.Reparametrized(gs.NonlinearProblem(fam), gs.PolarChart())
    with pytest.raises(RuntimeError):
        orc.solve_batch(gs.Newton(), fam, X0[:4].copy(), gs.PerSystem(p.data[:4]), o32, reparametrized=polar)
