"""Synthetic batches of the benchmark configurations, from a documented generator.

Julia's Xoshiro/randn streams cannot be reproduced without Julia, so every batch here comes from
numpy's PCG64 (`numpy.random.default_rng(seed)`), drawn in the order written below. The oracle and
the device solve always receive the same arrays.
"""
from __future__ import annotations

import math

import numpy as np


def henon_heiles_potential(q1, q2):
    """V(q) = ½(q₁² + q₂²) + q₁²q₂ − q₂³/3."""
    return 0.5 * (q1 * q1 + q2 * q2) + q1 * q1 * q2 - q2 ** 3 / 3.0


def henon_heiles_states(nsys: int, seed: int = 0, e_lo: float = 0.01, e_hi: float = 0.15) -> np.ndarray:
    """States x_k = (q₁, q₂, p₁, p₂) on energy shells E ~ U[e_lo, e_hi] (below the escape energy 1/6).

    Per system: E, then q uniform in the bounded component of {V(q) ≤ E} — the inside of the
    triangle q₂ > −½, q₂ < 1 ± √3 q₁ whose sides are the level set V = 1/6 — by rejection from the
    box [−1, 1] × [−½, 1], then |p| = √(2(E − V(q))) at a uniform angle θ ∈ [0, 2π).
    """
    rng = np.random.default_rng(seed)
    E = rng.uniform(e_lo, e_hi, size=nsys)
    q = np.empty((nsys, 2))
    todo = np.arange(nsys)
    s3 = math.sqrt(3.0)
    while todo.size:
        c1 = rng.uniform(-1.0, 1.0, size=todo.size)
        c2 = rng.uniform(-0.5, 1.0, size=todo.size)
        inside = (c2 > -0.5) & (c2 < 1.0 - s3 * c1) & (c2 < 1.0 + s3 * c1)
        ok = inside & (henon_heiles_potential(c1, c2) <= E[todo])
        q[todo[ok], 0] = c1[ok]
        q[todo[ok], 1] = c2[ok]
        todo = todo[~ok]
    V = henon_heiles_potential(q[:, 0], q[:, 1])
    pm = np.sqrt(np.maximum(2.0 * (E - V), 0.0))
    theta = rng.uniform(0.0, 2.0 * math.pi, size=nsys)
    x = np.empty((nsys, 4))
    x[:, 0:2] = q
    x[:, 2] = pm * np.cos(theta)
    x[:, 3] = pm * np.sin(theta)
    return x


def henon_heiles_energy(x: np.ndarray) -> np.ndarray:
    return 0.5 * (x[:, 2] ** 2 + x[:, 3] ** 2) + henon_heiles_potential(x[:, 0], x[:, 1])


def henon_heiles_midpoint_residual(x: np.ndarray, xk: np.ndarray, h: float) -> np.ndarray:
    """F(x) = x − x_k − h f((x_k + x)/2), vectorised over systems (the convergence recheck)."""
    m = (xk + x) / 2.0
    f = np.stack([m[:, 2], m[:, 3], -m[:, 0] - 2.0 * m[:, 0] * m[:, 1],
                  -m[:, 1] - m[:, 0] ** 2 + m[:, 1] ** 2], axis=1)
    return (x - xk) - h * f


def tanh_linear3_batch(nsys: int, seed: int = 0):
    """A_p = 3I + 0.5·N(0,1) per entry (9 draws, column-major), then b_p ~ N(0,1) (3 draws), per
    system in that order; x₀ = 0. Returns (records (nsys, 12): vec(A_p) then b_p, x₀ (nsys, 3))."""
    rng = np.random.default_rng(seed)
    z = rng.standard_normal(size=(nsys, 12))
    rec = np.empty((nsys, 12))
    rec[:, :9] = 0.5 * z[:, :9]
    for i in range(3):
        rec[:, i + 3 * i] += 3.0
    rec[:, 9:] = z[:, 9:]
    return rec, np.zeros((nsys, 3))


def tanh_linear3_residual(x: np.ndarray, rec: np.ndarray) -> np.ndarray:
    A = rec[:, :9].reshape(-1, 3, 3).transpose(0, 2, 1)  # column-major records
    return np.einsum("sij,sj->si", A, x) + 0.1 * np.tanh(x) - rec[:, 9:]


def fpu_states(nsys: int, seed: int = 0, sigma: float = 0.3) -> np.ndarray:
    """FPU-β chain states: (q, p) ~ N(0, σ²) entrywise, 32 draws per system (q₁..q₁₆ then p₁..p₁₆).
    Returns (nsys, 32)."""
    rng = np.random.default_rng(seed)
    return rng.standard_normal(size=(nsys, 32)) * sigma


def fpu_gl3_residual(z: np.ndarray, qp: np.ndarray, h: float, beta: float = 1.0) -> np.ndarray:
    """The stage residual of FpuGl3, vectorised over systems (a host recheck; not the oracle)."""
    import math

    s15 = math.sqrt(15.0)
    c = np.array([0.5 - s15 / 10, 0.5, 0.5 + s15 / 10])
    A = np.array([[5 / 36, 2 / 9 - s15 / 15, 5 / 36 - s15 / 30],
                  [5 / 36 + s15 / 24, 2 / 9, 5 / 36 - s15 / 24],
                  [5 / 36 + s15 / 30, 2 / 9 + s15 / 15, 5 / 36]])
    A2 = A @ A
    q, p = qp[:, :16], qp[:, 16:]
    n = len(z)
    Q = q[:, None, :] + z.reshape(n, 3, 16)
    pad = np.zeros((n, 3, 1))
    Qe = np.concatenate([pad, Q, pad], axis=2)
    r = np.diff(Qe, axis=2)                      # 17 bonds
    dphi = r + beta * r ** 3
    g = dphi[:, :, 1:] - dphi[:, :, :-1]
    R = z.reshape(n, 3, 16) - h * c[None, :, None] * p[:, None, :] - h * h * np.einsum("ab,nbm->nam", A2, g)
    return R.reshape(n, 48)


def expsum_fits(nfit: int, seed: int = 0, m_lo: int = 16, m_hi: int = 128, sigma: float = 1e-3):
    """Exponential-sum least-squares fits y(t) = Σ_{k≤K} a_k exp(−λ_k t) + noise.

    Per fit, drawn in this order: K ~ U{1..4}, m ~ U{m_lo..m_hi}, a_k ~ U[0.5, 2] (K draws),
    log10 λ_k ~ U[−1, 1] (K draws), then m noise samples N(0, σ²). The points t_i = 5(i−1)/(m−1)
    are equispaced on [0, 5], so only y travels. The start is a_k = 1, λ_k = 10^(−1 + (2k−1)/K).
    Returns (K (int32), m (int32), y offsets (int64), Y (concatenated), X0 (nfit, 8), truth (nfit, 8)),
    with x = (a_1..a_K, λ_1..λ_K) in the first 2K slots of a row.
    """
    rng = np.random.default_rng(seed)
    K = rng.integers(1, 5, size=nfit).astype(np.int32)
    m = rng.integers(m_lo, m_hi + 1, size=nfit).astype(np.int32)
    a = rng.uniform(0.5, 2.0, size=(nfit, 4))
    lam = 10.0 ** rng.uniform(-1.0, 1.0, size=(nfit, 4))
    off = np.zeros(nfit, dtype=np.int64)
    off[1:] = np.cumsum(m[:-1], dtype=np.int64)
    Y = np.empty(int(m.sum()))
    noise = rng.standard_normal(size=len(Y)) * sigma
    X0 = np.zeros((nfit, 8))
    truth = np.zeros((nfit, 8))
    for k in range(1, 5):
        sel = K == k
        X0[sel, :k] = 1.0
        X0[sel, k:2 * k] = 10.0 ** (-1.0 + (2.0 * np.arange(1, k + 1) - 1.0) / k)
        truth[sel, :k] = a[sel, :k]
        truth[sel, k:2 * k] = lam[sel, :k]
    for mm in np.unique(m):
        sel = np.nonzero(m == mm)[0]
        t = 5.0 * np.arange(mm) / (mm - 1)
        yy = np.zeros((len(sel), mm))
        for k in range(4):
            use = (K[sel] > k)[:, None]
            yy += np.where(use, a[sel, k:k + 1] * np.exp(-lam[sel, k:k + 1] * t[None, :]), 0.0)
        idx = off[sel][:, None] + np.arange(mm)[None, :]
        Y[idx] = yy
    Y += noise
    return K, m, off, Y, X0, truth


def expsum_residual_norms(X8, K, m, off, Y):
    """‖F(x)‖₂ of each fit (a host recheck)."""
    out = np.empty(len(K))
    for i in range(len(K)):
        k, mm = int(K[i]), int(m[i])
        t = 5.0 * np.arange(mm) / (mm - 1)
        f = (X8[i, :k, None] * np.exp(-X8[i, k:2 * k, None] * t[None, :])).sum(axis=0) - Y[off[i]:off[i] + mm]
        out[i] = np.sqrt((f * f).sum())
    return out


def tene_matrix(n: int, seed: int | None = None) -> np.ndarray:
    """A = 2I + R of the §10.E problem (benchmark/device_problems.jl:133-145): R with a zero diagonal
    and entries uniform on [−1, 1] scaled by 1/(2(n − 1)), so each off-diagonal row sum is at most 1/2.
    One matrix per n (numpy PCG64 from the seed n: the reference's Xoshiro stream is Julia's)."""
    rng = np.random.default_rng(n if seed is None else seed)
    R = 2.0 * rng.random((n, n)) - 1.0
    np.fill_diagonal(R, 0.0)
    R *= 0.5 / (n - 1)
    return 2.0 * np.eye(n) + R


def tene_batch(n: int, nsys: int, seed: int = 1):
    """(A, B, X*) of the §10.E batch (device_problems.jl:159-175): each system's own root x* uniform on
    [−1, 1]ⁿ and b = A x* + sin.(x*) / 10; the solve starts from x₀ = 0."""
    A = tene_matrix(n)
    rng = np.random.default_rng(seed)
    Xs = 2.0 * rng.random((nsys, n)) - 1.0
    B = Xs @ A.T + np.sin(Xs) / 10.0
    return A, B, Xs


def tene_residual(x: np.ndarray, A: np.ndarray, B: np.ndarray) -> np.ndarray:
    """F(x, b) = A x + sin.(x) / 10 − b, row-wise in numpy (the host FP64 recheck)."""
    return x @ A.T + np.sin(x) / 10.0 - B
