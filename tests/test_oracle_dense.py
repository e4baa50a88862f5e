"""The oracle's dense family (family 10, the §10.E problem of benchmark/device_problems.jl:133-175,
F(x, b) = A x + sin.(x)/10 − b) pinned by what the mathematics fixes: its sine and cosine against
libm to 1 ulp, its Jacobian against central differences of its residual, the roots the batch was
built around, and Newton's local quadratic convergence. The loop, the LU and the globalizations it
runs are the oracle's generic ones, pinned by tests/test_oracle.py at every size they are templated on.
"""
import math

import numpy as np
import pytest


def test_fsincos_within_one_ulp(orc):
    x = np.concatenate([np.linspace(-100.0, 100.0, 400001), np.pi * np.arange(-64, 65) / 4,
                        [1e-300, -1e-300, 5e-324, 1e6 + 0.5, -3e7]])
    s, c = orc.fsincos(x)
    for got, want in ((s, np.sin(x)), (c, np.cos(x))):
        ulp = np.spacing(np.abs(want))
        assert (np.abs(got - want) <= 1.0 * ulp).all(), np.abs(got - want).max()
    # signed zero, and NaN where the reduction has no exact quadrant
    s0, c0 = orc.fsincos([0.0, -0.0])
    assert math.copysign(1, s0[0]) > 0 and math.copysign(1, s0[1]) < 0 and (c0 == 1).all()
    sn, cn = orc.fsincos([np.nan, np.inf, -np.inf, 2.0 ** 50, -(2.0 ** 51)])
    assert np.isnan(sn).all() and np.isnan(cn).all()
    # Float32: within 1 ulp of the correctly rounded float
    x32 = np.linspace(-50.0, 50.0, 100001).astype(np.float32)
    s32, c32 = orc.fsincos(x32, f32=True)
    for got, want in ((s32, np.sin(x32.astype(np.float64))), (c32, np.cos(x32.astype(np.float64)))):
        w32 = want.astype(np.float32)
        assert (np.abs(got.astype(np.float32) - w32) <= np.spacing(np.abs(w32))).all()


@pytest.mark.parametrize("n", [2, 4, 6, 8, 12, 16, 24, 32, 48, 64])
def test_dense_jacobian_is_the_derivative(orc, workloads, n):
    A, B, _ = workloads.tene_batch(n, 4, seed=3)
    rng = np.random.default_rng(n)
    x = rng.uniform(-2, 2, n)
    F, J = orc.dense_eval(A, B[0], x)
    assert np.allclose(F, x @ A.T + np.sin(x) / 10 - B[0], rtol=0, atol=1e-14)
    h = 1e-6
    Jfd = np.empty((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = h
        Jfd[:, j] = (orc.dense_eval(A, B[0], x + e)[0] - orc.dense_eval(A, B[0], x - e)[0]) / (2 * h)
    assert np.abs(J - Jfd).max() < 1e-8
    # off the diagonal J is A itself, bit for bit (the product's partials are A's entries)
    off = ~np.eye(n, dtype=bool)
    assert np.array_equal(J[off], A[off])


@pytest.mark.parametrize("n", [16, 32, 64])
def test_dense_solve_finds_the_built_roots_quadratically(gs, orc, workloads, n):
    A, B, Xs = workloads.tene_batch(n, 512)
    fam = gs.DenseSin(A)
    x, s = orc.solve_batch(gs.Newton(), fam, np.zeros((512, n)), gs.PerSystem(B), gs.Options(f_abstol=1e-12, f_reltol=0.0))
    assert (s["code"] == 0).all()
    assert np.abs(x - Xs).max() < 1e-12  # A = 2I + R is well conditioned: ‖A⁻¹‖ ≤ 1
    # the residual norms of the first three iterates fall quadratically
    e = []
    for k in (1, 2, 3):
        xk, sk = orc.solve_batch(gs.Newton(), fam, np.zeros((512, n)), gs.PerSystem(B),
                                 gs.Options(f_abstol=0.0, f_reltol=0.0, x_reltol=0.0, max_iterations=k))
        e.append(np.abs(xk - Xs).max(axis=1))
    assert (e[1] <= 4 * e[0] ** 2 + 1e-15).all() and (e[2] <= 4 * e[1] ** 2 + 1e-15).all()
    # the status's gradient ratio is the last factorisation's ‖Jᴴ F‖/(‖J‖_F ‖F‖) at the previous iterate
    xp, sp = orc.solve_batch(gs.Newton(), fam, np.zeros((512, n)), gs.PerSystem(B),
                             gs.Options(f_abstol=0.0, f_reltol=0.0, x_reltol=0.0, max_iterations=2))
    for i in range(0, 512, 97):
        F, J = orc.dense_eval(A, B[i], xp[i])
        want = np.linalg.norm(J.T @ F) / (np.linalg.norm(J) * np.linalg.norm(F))
        assert math.isclose(s["gradratio"][i], want, rel_tol=1e-9)
