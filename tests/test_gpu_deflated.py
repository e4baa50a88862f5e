"""Parity of the Deflated batch (src/nonlinear/deflation.jl:714-760) with the CPU oracle, through the C
ABI (gs_solve_batch_deflated_f64).

The device wraps each lane-per-system family in the deflation M(x) F(x) with the product-rule
Jacobian and the status hook that reports the inner residual; it follows the oracle's operation
order (the integer power, the division-free (M, ∇M) recurrence, the rank-1 Jacobian term), so the
polynomial families are compared bit for bit, the tanh family to the floating-point bar.
"""
import math

import numpy as np
import pytest
import torch

from test_gpu_parity import assert_parity

pytestmark = pytest.mark.gpu

TOL = dict(f_abstol=1e-12, f_reltol=0.0)


def _gpu(gs, device, defl, X0, p_host, method, opts, counts=False):
    X = torch.as_tensor(X0, device=device).clone()
    if isinstance(p_host, gs.PerSystem):
        p = gs.PerSystem(torch.as_tensor(np.asarray(p_host.data), device=device))
    else:
        p = gs.BatchParams(p_host.shared, torch.as_tensor(np.asarray(p_host.per_system), device=device))
    if counts:
        tally = gs.solve_batch(X, defl, method, p, opts)
        return X.cpu().numpy(), tally
    S = gs.status_buffer(X.shape[0], device)
    gs.solve_batch(X, S, defl, method, p, opts)
    return X.cpu().numpy(), gs.statuses(S)


def _both(gs, orc, device, fam, roots, X0, p, method, opts, **kw):
    defl = gs.Deflated(gs.NonlinearProblem(fam), roots, **kw)
    x, st = _gpu(gs, device, defl, X0, p, method, opts)
    xo, so = orc.solve_batch(method, fam, X0.copy(), p, opts, deflated=defl)
    return x, st, xo, so


METHODS = {
    "static": lambda gs: gs.Newton(),
    "backtracking": lambda gs: gs.Newton(globalization=gs.Backtracking()),
    "refine-regularized": lambda gs: gs.Newton(refinement=gs.Refinement(), regularization=1e-6),
    "mixed-refine": lambda gs: gs.Newton(linsolve=gs.LUFactorization("float32"), refinement=gs.Refinement()),
    "mixed-backtracking": lambda gs: gs.Newton(linsolve=gs.LUFactorization("float32"),
                                               globalization=gs.Backtracking(), refinement=gs.Refinement()),
    "dogleg": lambda gs: gs.Newton(globalization=gs.DogLeg()),
    "mixed-dogleg": lambda gs: gs.Newton(linsolve=gs.LUFactorization("float32"), globalization=gs.DogLeg(radius=2.0),
                                         refinement=gs.Refinement()),
    # reuse = k: the refinement along a held factor takes the deflated JVP M (J v) + ⟨∇M, v⟩ F
    "reuse3": lambda gs: gs.Newton(reuse=3),
    "reuse2-backtracking-refine": lambda gs: gs.Newton(reuse=2, globalization=gs.Backtracking(),
                                                       refinement=gs.Refinement()),
    "mixed-reuse3-refine": lambda gs: gs.Newton(linsolve=gs.LUFactorization("float32"), reuse=3,
                                                refinement=gs.Refinement()),
}


@pytest.mark.parametrize("name", list(METHODS))
def test_coupled_family_bitwise(gs, orc, device, name):
    """The coupled family at p = (10, 10) (four real roots) deflated at two of them, from 4096 seeded
    starts, in every method form (DogLeg included); also MaxNorm and a power of 3 with a shift of 0.5."""
    rng = np.random.default_rng(21)
    X0 = rng.uniform(-12.0, 12.0, size=(4096, 2))
    p = gs.PerSystem(np.tile([10.0, 10.0], (4096, 1)))
    roots = ((-4 + math.sqrt(96), -4 + math.sqrt(96)), (4 + math.sqrt(32), 4 - math.sqrt(32)))
    method = METHODS[name](gs)
    for opts, kw in ((gs.Options(max_iterations=60, **TOL), {}),
                     (gs.Options(norm=gs.MaxNorm(), max_iterations=60, **TOL), dict(power=3, shift=0.5))):
        x, st, xo, so = _both(gs, orc, device, gs.Coupled2(), roots, X0, p, method, opts, **kw)
        assert_parity(st, x, so, xo, bitwise=True)
    print(name, np.bincount(st["code"], minlength=6).tolist())


def test_scalar_square_and_midpoint_bitwise(gs, orc, workloads, device):
    """x² − p deflated at +√p (starts on and around the root: a start on it is NONFINITE at 0
    iterations), and the Hénon–Heiles midpoint step deflated at its undeflated solution."""
    p = np.linspace(0.5, 4.0, 1024)
    X0 = (np.sqrt(p) * np.random.default_rng(2).uniform(0.5, 1.5, size=1024)).reshape(-1, 1)
    X0[:8, 0] = np.sqrt(p[:8])
    fam = gs.ScalarSquare()
    # one root shared by the batch: the root of the first system
    x, st, xo, so = _both(gs, orc, device, fam, (math.sqrt(p[0]),), X0, gs.PerSystem(p), gs.Newton(),
                          gs.Options(**TOL))
    assert_parity(st, x, so, xo, bitwise=True)
    assert st["code"][0] == gs.NONFINITE and st["iterations"][0] == 0
    xk = workloads.henon_heiles_states(2048, seed=4)
    hh = gs.HenonHeilesMidpoint()
    bp = gs.BatchParams(0.1, xk)
    x1, s1 = orc.solve_batch(gs.Newton(), hh, xk.copy(), bp, gs.Options(**TOL))
    root = tuple(x1[0])
    for m in (gs.Newton(), gs.Newton(globalization=gs.Backtracking())):
        x, st, xo, so = _both(gs, orc, device, hh, (root,), xk.copy(), bp, m, gs.Options(max_iterations=30, **TOL))
        assert_parity(st, x, so, xo, bitwise=True)


def test_tanh_family_to_tolerance(gs, orc, workloads, device):
    """The tanh family (a transcendental residual: ROCm's and glibc's tanh may differ in the last
    bit) deflated at the root of its first system: every code identical; the converged systems to
    the floating-point bar (iterations equal, iterates within 1e-10, both residuals at the
    tolerance: their last steps are near the round-off of x, so stepnorm and rate are not
    compared). The systems the deflation leaves without a root near their path wander for 50
    iterations (MAXITERS), where last-bit differences grow, so their norms and iterates are not
    compared."""
    rec, X0 = workloads.tanh_linear3_batch(2048, 9)
    fam = gs.TanhLinear3()
    p = gs.PerSystem(rec)
    x1, _ = orc.solve_batch(gs.Newton(), fam, X0.copy(), p, gs.Options(**TOL))
    x, st, xo, so = _both(gs, orc, device, fam, (tuple(x1[0]),), X0, p, gs.Newton(), gs.Options(max_iterations=50, **TOL))
    assert (st["code"] == so["code"]).all()
    ok = so["code"] == gs.SUCCESS
    assert ok.sum() > len(ok) * 0.9
    assert (st["iterations"][ok] == so["iterations"][ok]).all()
    scale = np.abs(xo[ok]).max(axis=1, keepdims=True)
    assert (np.abs(x[ok] - xo[ok]) <= 1e-10 * scale).all()
    assert (st["fnorm"][ok] <= 1e-12).all() and (so["fnorm"][ok] <= 1e-12).all()


def test_one_deflation_for_a_batch(gs, orc, device):
    """test/nonlinear/deflation.jl:304-320: every system of a batch starts at the same iterate and sees
    the same roots, so all four end at the one root the deflation leaves them, away from r₁; the
    counts form counts the same four successes."""
    fam = gs.Coupled2()
    prob = gs.NonlinearProblem(fam)
    p = gs.PerSystem(np.tile([10.0, 10.0], (4, 1)))
    x0 = np.tile([0.5, -0.3], (4, 1))
    opts = gs.Options(**TOL)
    r1, _ = orc.solve_batch(gs.Newton(), fam, x0[:1].copy(), gs.PerSystem(np.array([[10.0, 10.0]])), opts)
    defl = gs.Deflated(prob, (tuple(r1[0]),))
    x, st = _gpu(gs, device, defl, x0, p, gs.Newton(), opts)
    assert (st["code"] == gs.SUCCESS).all()
    assert (np.linalg.norm(x - r1[0], axis=1) > 0.5).all()
    assert (x == x[0]).all()
    x2, counts = _gpu(gs, device, defl, x0, p, gs.Newton(), opts, counts=True)
    assert counts["SUCCESS"] == 4 and (x2 == x).all()


def test_refusals(gs, device):
    """best = true is refused as the reference refuses it; a Float32 batch has no deflated device
    form."""
    fam = gs.Coupled2()
    defl = gs.Deflated(gs.NonlinearProblem(fam), ((1.0, 2.0),))
    X = torch.zeros((4, 2), dtype=torch.float64, device=device)
    p = gs.PerSystem(torch.zeros((4, 2), dtype=torch.float64, device=device))
    with pytest.raises(gs.ArgumentError, match="best = true"):
        gs.solve_batch(X, defl, gs.Newton(best=True), p)
    with pytest.raises(gs.ArgumentError, match="element type of its roots"):
        gs.solve_batch(X.float(), defl, gs.Newton(), gs.PerSystem(torch.zeros((4, 2), device=device)))
