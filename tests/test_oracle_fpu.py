"""Pins the oracle's FPU-β Gauss–Legendre s = 3 stage system (oracle/families.hpp FpuGl3): its
Jacobian against central differences of its residual, its tableau (c and every entry of A², read
back through F and J) against the collocation construction of test/helpers/stage_problems.jl:148-160,
and the solve of the linear chain (β = 0), where Newton is exact after one step, against
numpy.linalg.solve of the same linear stage system."""
import math

import numpy as np

H = 0.05


def _tableau():
    # Gauss nodes by Golub–Welsch and A from A V = Ĉ, as the reference's helper builds them
    s = 3
    beta = [k / math.sqrt(4 * k * k - 1) for k in range(1, s)]
    T = np.diag(beta, 1) + np.diag(beta, -1)
    c = (np.sort(np.linalg.eigvalsh(T)) + 1) / 2
    V = np.array([[c[j] ** k for k in range(s)] for j in range(s)])
    Ch = np.array([[c[i] ** (k + 1) / (k + 1) for k in range(s)] for i in range(s)])
    A = Ch @ np.linalg.inv(V)
    return c, A


def _laplacian():
    L = -2.0 * np.eye(16) + np.eye(16, k=1) + np.eye(16, k=-1)
    return L


def test_jacobian_against_central_differences(orc, workloads):
    qp = workloads.fpu_states(3, seed=7)
    rng = np.random.default_rng(2)
    for s in range(3):
        z = rng.standard_normal(48) * 0.02
        _, J = orc.fpu_gl3_eval(H, 1.0, qp[s], z)
        eps = 1e-6
        for j in range(48):
            zp, zm = z.copy(), z.copy()
            zp[j] += eps
            zm[j] -= eps
            fd = (orc.fpu_gl3_eval(H, 1.0, qp[s], zp)[0] - orc.fpu_gl3_eval(H, 1.0, qp[s], zm)[0]) / (2 * eps)
            assert np.abs(fd - J[:, j]).max() <= 1e-9, j
        # the structure: 3 × 3 blocks of tridiagonal 16 × 16
        mask = np.abs(np.subtract.outer(np.arange(48) % 16, np.arange(48) % 16)) <= 1
        assert np.all(J[~mask] == 0.0)


def test_tableau_read_back_through_F_and_J(orc):
    c, A = _tableau()
    A2 = A @ A
    # c: q = 0, p = e_m, z = 0, β = 0 → F_{a,m} = −h c_a
    for m in (0, 7, 15):
        qp = np.zeros(32)
        qp[16 + m] = 1.0
        F, _ = orc.fpu_gl3_eval(H, 0.0, qp, np.zeros(48))
        for a in range(3):
            assert abs(-F[a * 16 + m] / H - c[a]) <= 4e-16
    # A²: for β = 0, J_{(a,m),(b,m)} = δ_ab − h² A²_ab (−2), so A²_ab = (J − δ)/(2h²)
    _, J = orc.fpu_gl3_eval(H, 0.0, np.zeros(32), np.zeros(48))
    for a in range(3):
        for b in range(3):
            v = (J[a * 16 + 5, b * 16 + 5] - (1.0 if a == b else 0.0)) / (2 * H * H)
            assert abs(v - A2[a, b]) <= 1e-13 * max(1.0, abs(A2[a, b])), (a, b, v, A2[a, b])
    # row sums of A² are c²/2 (the Nyström relation Σ_j (A²)_ij = Σ_j a_ij c_j = c_i²/2)
    assert np.allclose(A2.sum(axis=1), c ** 2 / 2, rtol=1e-14, atol=0)


def test_linear_chain_solve_is_the_linear_system(gs, orc, workloads):
    c, A = _tableau()
    A2 = A @ A
    L = _laplacian()
    qp = workloads.fpu_states(16, seed=9)
    opts = gs.Options(f_abstol=1e-14, f_reltol=0.0)
    fam = gs.FpuGl3(beta=0.0)
    z, S = orc.solve_batch(gs.Newton(), fam, np.zeros((16, 48)), gs.BatchParams(H, qp), opts)
    M = np.eye(48) - H * H * np.kron(A2, L)
    for s in range(16):
        q, p = qp[s, :16], qp[s, 16:]
        rhs = H * np.kron(c, p) + H * H * np.kron(A2, L) @ np.kron(np.ones(3), q)
        zs = np.linalg.solve(M, rhs)
        assert np.abs(z[s] - zs).max() <= 1e-14 * max(1.0, np.abs(zs).max())
    assert (S["code"] == 0).all() and (S["iterations"] <= 2).all()
