"""The oracle's StationarityProblem batch (src/nonlinear/stationarity.jl) over the Hénon–Heiles
two-step discrete action, pinned without the oracle's own formulas:

- ∇S by complex-step differentiation of an independent numpy S (exact to rounding for a
  polynomial), ∇²S by central differences of that gradient (exact to rounding: the gradient of a
  cubic is quadratic);
- λ = 0 (a quadratic action): every index-set and datum form against numpy.linalg.solve of the
  linear stationarity equations;
- λ = 1: the converged discrete Euler–Lagrange residual re-evaluated by the complex-step gradient;
- the index-set semantics of the reference's tests (test/nonlinear/stationarity.jl:519-583): the
  entries outside U keep their bits, a permutation of E and U leaves the solution, the datum forms
  (none, a number, a vector, PerSystem, the keyword over the problem's own) and the refusals.
"""
import numpy as np
import pytest

H = 0.1


def action_np(x, h, lam):
    """S(x) on numpy (real or complex) arrays, written from the definition of L_d and V."""
    def V(q):
        return 0.5 * (q[0] ** 2 + q[1] ** 2) + lam * (q[0] ** 2 * q[1] - q[1] ** 3 / 3)

    def Ld(a, b):
        d = b - a
        return (d[0] ** 2 + d[1] ** 2) / (2 * h) - h * V((a + b) / 2)

    return Ld(x[0:2], x[2:4]) + Ld(x[2:4], x[4:6])


def grad_cs(x, h, lam):
    g = np.zeros(6)
    for i in range(6):
        z = x.astype(complex)
        z[i] += 1e-30j
        g[i] = action_np(z, h, lam).imag / 1e-30
    return g


def hess_cd(x, h, lam, d=1e-3):
    Hm = np.zeros((6, 6))
    for j in range(6):
        e = np.zeros(6)
        e[j] = d
        Hm[:, j] = (grad_cs(x + e, h, lam) - grad_cs(x - e, h, lam)) / (2 * d)
    return Hm


def states(n, seed):
    """Full states (q₋, q, q₊) with q₊ the linear extrapolation 2q − q₋ as the start."""
    rng = np.random.default_rng(seed)
    q0 = rng.uniform(-0.4, 0.4, size=(n, 2))
    q1 = q0 + H * rng.uniform(-0.3, 0.3, size=(n, 2))
    return np.hstack([q0, q1, 2 * q1 - q0])


def test_action_gradient_and_hessian(orc):
    rng = np.random.default_rng(0)
    for lam in (0.0, 1.0, -0.7):
        for _ in range(20):
            x = rng.uniform(-1.0, 1.0, size=6)
            g, Hm = orc.hh_action_eval(H, lam, x)
            gr = grad_cs(x, H, lam)
            assert np.allclose(g, gr, rtol=1e-13, atol=1e-13 * np.abs(gr).max())
            Hr = hess_cd(x, H, lam)
            assert np.allclose(Hm, Hr, rtol=1e-9, atol=1e-9)
            assert (Hm == Hm.T).all() and (Hm[0:2, 4:6] == 0).all()


def _linear_solution(x0, E, U, b, h):
    """λ = 0: ∇S(x) = H x, so (H x)_E + b = 0 is linear in x[U]."""
    Hm = hess_cd(np.zeros(6), h, 0.0)
    e = np.arange(6) if E is None else np.array(E) - 1
    u = np.arange(6) if U is None else np.array(U) - 1
    fixed = np.setdiff1d(np.arange(6), u)
    rhs = -(b if b is not None else 0.0) - Hm[np.ix_(e, fixed)] @ x0[fixed]
    x = x0.copy()
    x[u] = np.linalg.solve(Hm[np.ix_(e, u)], rhs)
    return x


@pytest.mark.parametrize("E,U,b", [
    ((3, 4), (5, 6), None),
    ((3, 4), (5, 6), (0.2, -0.1)),
    ((1, 2), (1, 2), 0.05),
    (None, None, (0.1, -0.2, 0.3, 0.0, -0.05, 0.02)),
    ((3, 4), (1, 2), None),
])
def test_linear_action_closed_form(gs, orc, E, U, b):
    X0 = states(64, 1)
    prob = gs.StationarityProblem(gs.HenonHeilesAction(), E=E, U=U, b=b)
    x, s = orc.solve_batch(gs.Newton(), gs.HenonHeilesAction(), X0.copy(), (H, 0.0),
                           gs.Options(f_abstol=1e-12, f_reltol=0.0), stationarity=prob)
    assert (s["code"] == gs.SUCCESS).all()
    bv = None if b is None else np.broadcast_to(np.asarray(b, dtype=float), (2 if E else 6,))
    u = np.arange(6) if U is None else np.array(U) - 1
    fixed = np.setdiff1d(np.arange(6), u)
    for i in range(len(X0)):
        ref = _linear_solution(X0[i], E, U, bv, H)
        assert np.allclose(x[i], ref, rtol=1e-11, atol=1e-12)
        assert (x[i, fixed].view(np.uint64) == X0[i, fixed].view(np.uint64)).all()


def test_discrete_euler_lagrange_residual(gs, orc):
    """λ = 1, E = (3, 4), U = (5, 6): q₊ solves D₂L_d(q₋, q) + D₁L_d(q, q₊) = 0, checked with the
    complex-step gradient of the numpy action at the result; Backtracking and the mixed factor
    reach the same solutions."""
    X0 = states(256, 2)
    act = gs.HenonHeilesAction()
    prob = gs.StationarityProblem(act, E=(3, 4), U=(5, 6))
    opts = gs.Options(f_abstol=1e-13, f_reltol=0.0)
    x, s = orc.solve_batch(gs.Newton(), act, X0.copy(), (H, 1.0), opts, stationarity=prob)
    assert (s["code"] == gs.SUCCESS).all() and (s["iterations"] <= 4).all()
    for i in range(0, 256, 8):
        assert np.abs(grad_cs(x[i], H, 1.0)[2:4]).max() <= 2e-13
    for m in (gs.Newton(globalization=gs.Backtracking()),
              gs.Newton(linsolve=gs.LUFactorization("float32"), refinement=gs.Refinement())):
        x2, s2 = orc.solve_batch(m, act, X0.copy(), (H, 1.0), opts, stationarity=prob)
        assert (s2["code"] == gs.SUCCESS).all()
        assert np.allclose(x2, x, rtol=0, atol=1e-12)


def test_index_permutation_and_datum_forms(gs, orc):
    act = gs.HenonHeilesAction()
    X0 = states(32, 3)
    opts = gs.Options(f_abstol=1e-12, f_reltol=0.0)
    solve = lambda prob, **kw: orc.solve_batch(gs.Newton(), act, X0.copy(), (H, 1.0), opts, stationarity=prob, **kw)
    x1, _ = solve(gs.StationarityProblem(act, E=(3, 4), U=(5, 6)))
    x2, _ = solve(gs.StationarityProblem(act, E=(4, 3), U=(6, 5)))
    assert np.allclose(x1, x2, rtol=0, atol=1e-13)
    # a number datum is that number in every equation
    xa, sa = solve(gs.StationarityProblem(act, E=(3, 4), U=(5, 6), b=0.125))
    xb, sb = solve(gs.StationarityProblem(act, E=(3, 4), U=(5, 6), b=(0.125, 0.125)))
    assert (xa.view(np.uint64) == xb.view(np.uint64)).all()
    # the keyword datum overrides the problem's own; PerSystem gives system i its own datum
    bs = np.random.default_rng(4).uniform(-0.1, 0.1, size=(32, 2))
    prob_ps = gs.StationarityProblem(act, E=(3, 4), U=(5, 6), b=gs.PerSystem(bs))
    xp, _ = solve(prob_ps)
    xk, _ = solve(gs.StationarityProblem(act, E=(3, 4), U=(5, 6), b=0.5), b=gs.PerSystem(bs))
    assert (xp.view(np.uint64) == xk.view(np.uint64)).all()
    for i in (0, 7, 31):
        xi, _ = orc.solve_batch(gs.Newton(), act, X0[i:i + 1].copy(), (H, 1.0), opts,
                                stationarity=gs.StationarityProblem(act, E=(3, 4), U=(5, 6), b=tuple(bs[i])))
        assert (xi[0].view(np.uint64) == xp[i].view(np.uint64)).all()
        assert np.abs(grad_cs(xp[i], H, 1.0)[2:4] + bs[i]).max() <= 2e-13


def test_stationarity_checks(gs):
    """The constructor's checks (stationarity.jl:60-118) and the solve's (:133-157)."""
    act = gs.HenonHeilesAction()
    S = gs.StationarityProblem
    with pytest.raises(gs.ArgumentError, match="both E and U to be :"):
        S(act, E=(1, 2))
    with pytest.raises(gs.ArgumentError, match="length\\(E\\) == length\\(U\\)"):
        S(act, E=(1, 2), U=(1,))
    with pytest.raises(gs.ArgumentError, match="must be non-empty"):
        S(act, E=(), U=())
    with pytest.raises(gs.ArgumentError, match="indices in E must be >= 1"):
        S(act, E=(0, 1), U=(1, 2))
    with pytest.raises(gs.ArgumentError, match="E has duplicate"):
        S(act, E=(1, 1), U=(1, 2))
    with pytest.raises(gs.ArgumentError, match="U has duplicate"):
        S(act, E=(1, 2), U=(2, 2))
    with pytest.raises(gs.ArgumentError, match="callable datum"):
        S(act, b=lambda p: p)
    with pytest.raises(gs.ArgumentError, match="index in E exceeds 6"):
        S(act, E=(1, 7), U=(1, 2)).to_c(6, None, 4)
    with pytest.raises(gs.ArgumentError, match="datum b must have length 2, got 3"):
        S(act, E=(3, 4), U=(5, 6), b=(1.0, 2.0, 3.0)).to_c(6, None, 4)
    with pytest.raises(gs.ArgumentError, match="datum b must have length 6, got 2"):
        S(act).to_c(6, gs.PerSystem(np.zeros((4, 2))), 4)
    with pytest.raises(gs.ArgumentError, match="registered action"):
        S(gs.Coupled2())


def test_float32_batch_against_float64(gs, orc):
    """The Float32 batch (SolverStatus{Float32}, records and datum in Float32): λ = 0 against the
    closed-form solution of the linear equations, and λ = 1 against the Float64 solve, to Float32
    rounding; the entries outside U keep their bits, every system reaches SUCCESS."""
    n = 2048
    X0 = states(n, 41)
    act = gs.HenonHeilesAction()
    opts32 = gs.Options(eltype="float32", f_abstol=1e-5, max_iterations=30)
    prob = gs.StationarityProblem(act, E=(3, 4), U=(5, 6))
    for m in (gs.Newton(), gs.Newton(globalization=gs.Backtracking()), gs.Newton(globalization=gs.DogLeg()),
              gs.Newton(reuse=2)):
        x64, s64 = orc.solve_batch(m, act, X0.copy(), (H, 1.0), gs.Options(f_abstol=1e-13, max_iterations=30),
                                   stationarity=prob)
        X32 = X0.astype(np.float32)
        x32, s32 = orc.solve_batch(m, act, X32.copy(), (H, 1.0), opts32, stationarity=prob)
        assert x32.dtype == np.float32 and s32.dtype.itemsize == 44
        assert (s32["code"] == gs.SUCCESS).all() and (s64["code"] == gs.SUCCESS).all()
        assert (x32[:, 0:4].view(np.uint32) == X32[:, 0:4].view(np.uint32)).all()
        assert np.abs(x32 - x64).max() <= 1e-5
    # λ = 0: q₊ solves a linear system; its closed form from the Float64 linear solve
    x32, s32 = orc.solve_batch(gs.Newton(), act, X0.astype(np.float32), (H, 0.0), opts32, stationarity=prob)
    Hm = hess_cd(np.zeros(6), H, 0.0)
    for i in range(0, n, 97):
        x = X0[i].copy()
        g0 = grad_cs(np.concatenate([x[0:4], [0.0, 0.0]]), H, 0.0)
        sol = np.linalg.solve(Hm[np.ix_([2, 3], [4, 5])], -g0[[2, 3]])
        assert np.abs(x32[i, 4:6] - sol).max() <= 1e-5 * max(1.0, np.abs(sol).max())
