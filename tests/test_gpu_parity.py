"""Parity of the HIP batched solve with the CPU oracle, through the C ABI (libgsolve_hip).

The bar: return codes and iteration counts identical (except, for a family with transcendental
functions, systems whose final residual is within 1% of the tolerance), iterates within 1e-10
relative, and the floating-point status fields within 8 eps relative — the tolerance of the
reference's own R2-vs-R1 comparison `status_agrees` (test/nonlinear/batched.jl:42-52). For the
polynomial families, whose device arithmetic follows the oracle's operation order exactly, the
results are compared bit for bit as well.
"""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

EPS = np.finfo(float).eps
FLOAT_FIELDS = ("fnorm", "stepnorm", "gradratio", "rate", "nullnorm")
INT_FIELDS = ("code", "f_calls", "iterations", "step_failures", "promoted", "stationary", "jvp_calls",
              "rank", "contracted", "reused")


def _close(a, b, rtol):
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    both_nan = np.isnan(a) & np.isnan(b)
    same = a == b
    with np.errstate(invalid="ignore"):
        rel = np.abs(a - b) <= rtol * np.maximum(np.abs(a), np.abs(b))
    return both_nan | same | rel


def assert_parity(st, x, so, xo, *, bitwise=False, near_tol=None, tol=None):
    """Compare device statuses/iterates (st, x) with the oracle's (so, xo)."""
    mask = np.ones(len(st), dtype=bool)
    if near_tol is not None:
        # systems whose final residual lies within 1% of the tolerance may differ in code
        mask = ~((np.abs(so["fnorm"] - tol) <= 0.01 * tol) | (np.abs(st["fnorm"] - tol) <= 0.01 * tol))
    for f in INT_FIELDS:
        bad = np.nonzero((st[f] != so[f]) & mask)[0]
        assert bad.size == 0, f"{f} differs at {bad[:10]}: device {st[f][bad[:5]]}, oracle {so[f][bad[:5]]}"
    for f in FLOAT_FIELDS:
        if near_tol is None:
            ok = _close(st[f], so[f], 8 * EPS) | ~mask
        else:
            # a transcendental residual differs in its last bit between ROCm's and glibc's libm, so
            # the norms at the round-off floor (below 1e-10) agree in magnitude only
            a, b = st[f], so[f]
            floor = (np.abs(a) < 1e-10) & (np.abs(b) < 1e-10)
            if f in ("rate", "gradratio"):
                # the rate, and the gradient ratio at the last factorised iterate, of a last step
                # that started at the floor are ratios of floor-level norms
                floor |= (st["stepnorm"] < 1e-10) | (so["stepnorm"] < 1e-10)
            ok = _close(a, b, 1e-6) | floor | ~mask
        assert ok.all(), f"{f} differs at {np.nonzero(~ok)[0][:10]}"
    fin = np.isfinite(xo)
    scale = np.maximum(np.where(fin, np.abs(xo), 0.0).max(axis=1, keepdims=True), 1e-300)
    with np.errstate(invalid="ignore"):
        rel = np.where(fin, np.abs(np.where(fin, x, 0.0) - np.where(fin, xo, 0.0)) / scale, 0.0)
    # a non-finite oracle entry must be the same non-finite value on the device
    rel = np.where(~fin & ~((np.isnan(xo) & np.isnan(x)) | (xo == x)), np.inf, rel)
    rel = np.where(fin & ~np.isfinite(x), np.inf, rel)
    assert (rel[mask] <= 1e-10).all(), f"iterates differ by {rel[mask].max():.3e}"
    if bitwise:
        assert np.array_equal(x.view(np.uint64), xo.view(np.uint64)), "iterates differ in bits"
        a = st.view(np.uint8).reshape(len(st), 72)[:, :66]
        b = so.view(np.uint8).reshape(len(so), 72)[:, :66]
        if not np.array_equal(a, b):
            cols = np.nonzero((a != b).any(axis=0))[0]
            names = sorted({n for n, (dt, off) in st.dtype.fields.items() for c in cols if off <= c < off + dt.itemsize})
            rows = np.nonzero((a != b).any(axis=1))[0]
            raise AssertionError(f"statuses differ in bits: bytes {cols.tolist()} (fields {names}) at {rows[:10]} "
                                 f"({rows.size} systems); device {a[rows[0]].tolist()} oracle {b[rows[0]].tolist()}")


def same_bytes(a, b):
    """Equality of two status arrays byte for byte (NaN fields compare equal to themselves)."""
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def run_gpu(gs, device, method, family, X0, p_host, opts, order=None, workgroupsize=None, counts=False):
    X = torch.as_tensor(X0, device=device).clone()
    if isinstance(p_host, gs.PerSystem):
        p = gs.PerSystem(torch.as_tensor(np.asarray(p_host.data), device=device))
    elif isinstance(p_host, gs.BatchParams):
        p = gs.BatchParams(p_host.shared, torch.as_tensor(np.asarray(p_host.per_system), device=device))
    else:
        p = p_host
    prob = gs.NonlinearProblem(family)
    if counts:
        c = gs.solve_batch(X, prob, method, p, opts, order=order, workgroupsize=workgroupsize)
        return X.cpu().numpy(), c
    S = gs.status_buffer(X.shape[0], device, "float32" if X.dtype == torch.float32 else "float64")
    gs.solve_batch(X, S, prob, method, p, opts, order=order, workgroupsize=workgroupsize)
    return X.cpu().numpy(), gs.statuses(S)


def run_both(gs, orc, device, method, family, X0, p, opts, **kw):
    x, st = run_gpu(gs, device, method, family, X0, p, opts, **kw)
    xo, so = orc.solve_batch(method, family, X0.copy(), p, opts, order=kw.get("order"))
    return x, st, xo, so


HH_METHODS = {
    "newton-static": lambda gs: gs.Newton(),
    "newton-backtracking": lambda gs: gs.Newton(globalization=gs.Backtracking()),
    "mixed-static-refine": lambda gs: gs.Newton(linsolve=gs.LUFactorization("float32"), refinement=gs.Refinement()),
    "mixed-backtracking-refine": lambda gs: gs.Newton(linsolve=gs.LUFactorization("float32"),
                                                      globalization=gs.Backtracking(), refinement=gs.Refinement()),
    "mixed-static-norefine": lambda gs: gs.Newton(linsolve=gs.LUFactorization("float32")),
    "plain-refine": lambda gs: gs.Newton(refinement=gs.Refinement()),
    "best-regularized": lambda gs: gs.Newton(best=True, regularization=1e-3),
    # reuse = k: the held factor (newton.jl:1781-1874)
    "reuse1-static": lambda gs: gs.Newton(reuse=1),
    "reuse2-static": lambda gs: gs.Newton(reuse=2),
    "reuse3-backtracking": lambda gs: gs.Newton(reuse=3, globalization=gs.Backtracking()),
    "reuse3-refine-regularized": lambda gs: gs.Newton(reuse=3, refinement=gs.Refinement(), regularization=1e-3),
    "reuse3-mixed-static-refine": lambda gs: gs.Newton(reuse=3, linsolve=gs.LUFactorization("float32"),
                                                       refinement=gs.Refinement()),
    "reuse2-mixed-backtracking-refine": lambda gs: gs.Newton(reuse=2, linsolve=gs.LUFactorization("float32"),
                                                             globalization=gs.Backtracking(),
                                                             refinement=gs.Refinement()),
    "reuse4-mixed-static-norefine": lambda gs: gs.Newton(reuse=4, linsolve=gs.LUFactorization("float32")),
    # DogLeg: the trust-region trial loop on the factorised Jacobian (dogleg.jl:116-206)
    "dogleg": lambda gs: gs.Newton(globalization=gs.DogLeg()),
    "dogleg-radius": lambda gs: gs.Newton(globalization=gs.DogLeg(radius=0.05)),
    "dogleg-mixed-refine": lambda gs: gs.Newton(globalization=gs.DogLeg(), linsolve=gs.LUFactorization("float32"),
                                                refinement=gs.Refinement()),
    "dogleg-refine-regularized": lambda gs: gs.Newton(globalization=gs.DogLeg(radius=1e-3), refinement=gs.Refinement(),
                                                      regularization=1e-3),
}


@pytest.mark.parametrize("name", list(HH_METHODS))
def test_henon_heiles_midpoint_parity(gs, orc, workloads, device, name):
    """The n = 4 implicit-midpoint family at h = 0.1 on 65536 seeded states, every method form."""
    n = 65536
    xk = workloads.henon_heiles_states(n, seed=1)
    m = HH_METHODS[name](gs)
    opts = gs.Options(f_abstol=1e-12, f_reltol=0.0)
    x, st, xo, so = run_both(gs, orc, device, m, gs.HenonHeilesMidpoint(), xk.copy(), gs.BatchParams(0.1, xk), opts)
    assert_parity(st, x, so, xo, bitwise=True)
    assert (st["code"] == gs.SUCCESS).mean() > 0.99


def test_henon_heiles_hard_steps(gs, orc, workloads, device):
    """Large steps (h = 2) from far-out starts: a mix of codes, with Backtracking and MaxNorm."""
    n = 16384
    xk = workloads.henon_heiles_states(n, seed=2, e_lo=0.1, e_hi=0.166)
    for m in (gs.Newton(), gs.Newton(globalization=gs.Backtracking())):
        for opts in (gs.Options(f_abstol=1e-12, f_reltol=0.0, max_iterations=8),
                     gs.Options(norm=gs.MaxNorm(), min_iterations=1, x_reltol=1e-6, roundoff=True)):
            x, st, xo, so = run_both(gs, orc, device, m, gs.HenonHeilesMidpoint(), xk * 3.0, gs.BatchParams(2.0, xk * 3.0), opts)
            assert_parity(st, x, so, xo, bitwise=True)
            assert len(np.unique(st["code"])) >= 2


def test_reuse_hard_steps(gs, orc, workloads, device):
    """Newton(; reuse = k) where the held factor matters: large steps (h = 2) from far-out starts, so
    that steps along a stale factor stall, refresh on θ ≥ 1, or fail their search and refactorise."""
    n = 16384
    xk = workloads.henon_heiles_states(n, seed=7, e_lo=0.1, e_hi=0.166)
    f32 = gs.LUFactorization("float32")
    methods = (gs.Newton(reuse=3), gs.Newton(reuse=3, globalization=gs.Backtracking()),
               gs.Newton(reuse=2, linsolve=f32, refinement=gs.Refinement()),
               gs.Newton(reuse=3, linsolve=f32, globalization=gs.Backtracking(), refinement=gs.Refinement()),
               gs.Newton(reuse=5, refinement=gs.Refinement(), regularization=1e-2, globalization=gs.Backtracking()))
    for m in methods:
        for opts in (gs.Options(f_abstol=1e-12, f_reltol=0.0, max_iterations=30),
                     gs.Options(norm=gs.MaxNorm(), min_iterations=1, x_reltol=1e-6, roundoff=True)):
            x, st, xo, so = run_both(gs, orc, device, m, gs.HenonHeilesMidpoint(), xk * 3.0, gs.BatchParams(2.0, xk * 3.0),
                                     opts)
            assert_parity(st, x, so, xo, bitwise=True)
            assert st["reused"].any() and len(np.unique(st["code"])) >= 2


def test_dogleg_hard_steps(gs, orc, workloads, device):
    """DogLeg where the radius matters: large steps (h = 2) from far-out starts give Cauchy and
    segment steps, radius floors (LINESEARCH_FAILED), stalls and mixed-precision promotions."""
    n = 16384
    xk = workloads.henon_heiles_states(n, seed=7, e_lo=0.1, e_hi=0.166)
    f32 = gs.LUFactorization("float32")
    methods = (gs.Newton(globalization=gs.DogLeg()),
               gs.Newton(globalization=gs.DogLeg(), linsolve=f32, refinement=gs.Refinement()),
               gs.Newton(globalization=gs.DogLeg(radius=0.5), regularization=1e-2))
    for m in methods:
        for opts in (gs.Options(f_abstol=1e-12, f_reltol=0.0, max_iterations=30),
                     gs.Options(norm=gs.MaxNorm(), min_iterations=1, x_reltol=1e-6, roundoff=True)):
            x, st, xo, so = run_both(gs, orc, device, m, gs.HenonHeilesMidpoint(), xk * 3.0, gs.BatchParams(2.0, xk * 3.0),
                                     opts)
            assert_parity(st, x, so, xo, bitwise=True)
            assert (st["code"] == gs.LINESEARCH_FAILED).any() and (st["code"] == gs.SUCCESS).any()


def test_dogleg_coupled_batch_and_tanh(gs, orc, workloads, device):
    """test/nonlinear/batched.jl:263-352 with Newton(; globalization = DogLeg()): bit-identical to
    the oracle, the exact root in 0 steps, the counts form; and the n = 3 tanh family."""
    P = np.array([[2.125, 1.5], [1.25, 2.0], [1.5, 2.5], [1.75, 3.0]])
    X0 = np.array([[2.0, 1.0], [2.5, 1.5], [3.5, 2.5], [4.0, 3.0]])
    opts = gs.Options(max_iterations=150)
    m = gs.Newton(globalization=gs.DogLeg())
    x, st, xo, so = run_both(gs, orc, device, m, gs.Coupled2(), X0.copy(), gs.PerSystem(P), opts)
    assert_parity(st, x, so, xo, bitwise=True)
    assert (st["code"] == gs.SUCCESS).all() and st["iterations"][0] == 0 and np.array_equal(x[0], X0[0])
    _, counts = run_gpu(gs, device, m, gs.Coupled2(), X0.copy(), gs.PerSystem(P), opts, counts=True)
    assert counts == dict(SUCCESS=4, STALLED=0, MAXITERS=0, SINGULAR=0, NONFINITE=0, LINESEARCH_FAILED=0)
    rec, x0 = workloads.tanh_linear3_batch(4096, 0)
    opts = gs.Options(f_abstol=1e-12, f_reltol=0.0, max_iterations=50)
    x, st, xo, so = run_both(gs, orc, device, m, gs.TanhLinear3(), x0.copy(), gs.PerSystem(rec), opts)
    assert_parity(st, x, so, xo, near_tol=True, tol=1e-12)


def test_reuse_coupled_batch(gs, orc, device):
    """test/nonlinear/batched.jl:263-352 with Newton(; reuse = 2): bit-identical to the oracle, the
    exact root in 0 steps, and more steps than a fresh factor per step on systems 3 and 4."""
    P = np.array([[2.125, 1.5], [1.25, 2.0], [1.5, 2.5], [1.75, 3.0]])
    X0 = np.array([[2.0, 1.0], [2.5, 1.5], [3.5, 2.5], [4.0, 3.0]])
    opts = gs.Options(max_iterations=150)
    _, sn = run_gpu(gs, device, gs.Newton(), gs.Coupled2(), X0.copy(), gs.PerSystem(P), opts)
    x, st, xo, so = run_both(gs, orc, device, gs.Newton(reuse=2), gs.Coupled2(), X0.copy(), gs.PerSystem(P), opts)
    assert_parity(st, x, so, xo, bitwise=True)
    assert (st["code"] == gs.SUCCESS).all() and st["iterations"][0] == 0 and np.array_equal(x[0], X0[0])
    assert st["iterations"][2] > sn["iterations"][2] and st["iterations"][3] > sn["iterations"][3]
    _, counts = run_gpu(gs, device, gs.Newton(reuse=2), gs.Coupled2(), X0.copy(), gs.PerSystem(P), opts, counts=True)
    assert counts == dict(SUCCESS=4, STALLED=0, MAXITERS=0, SINGULAR=0, NONFINITE=0, LINESEARCH_FAILED=0)


def test_full_size_henon_heiles(gs, orc, workloads, device):
    """The benchmark size, 2^20 systems: every code and iterate against the oracle, and the FP64
    residual recheck ‖F(x)‖₂ ≤ 1e-12 of every SUCCESS."""
    n = 1 << 20
    xk = workloads.henon_heiles_states(n, seed=0)
    opts = gs.Options(f_abstol=1e-12, f_reltol=0.0)
    x, st, xo, so = run_both(gs, orc, device, gs.Newton(), gs.HenonHeilesMidpoint(), xk.copy(), gs.BatchParams(0.1, xk), opts)
    assert_parity(st, x, so, xo, bitwise=True)
    ok = st["code"] == gs.SUCCESS
    r = np.sqrt((workloads.henon_heiles_midpoint_residual(x, xk, 0.1) ** 2).sum(axis=1))
    assert ok.all() and (r[ok] <= 1e-12).all()
    # the implicit midpoint rule conserves the quadratic part only; the energy drift stays O(h²)
    assert np.abs(workloads.henon_heiles_energy(x) - workloads.henon_heiles_energy(xk)).max() < 1e-3


def test_tanh_linear3_parity(gs, orc, workloads, device):
    """The n = 3 family of the CPU configuration (batch 4096, max_iterations 50, seed 0) — tanh is
    ROCm's on the device and glibc's on the host, so codes may differ within 1% of the tolerance."""
    rec, x0 = workloads.tanh_linear3_batch(4096, 0)
    for m in (gs.Newton(), gs.Newton(globalization=gs.Backtracking())):
        opts = gs.Options(f_abstol=1e-12, f_reltol=0.0, max_iterations=50)
        x, st, xo, so = run_both(gs, orc, device, m, gs.TanhLinear3(), x0.copy(), gs.PerSystem(rec), opts)
        assert_parity(st, x, so, xo, near_tol=True, tol=1e-12)
        r = np.sqrt((workloads.tanh_linear3_residual(x, rec) ** 2).sum(axis=1))
        assert (r[st["code"] == gs.SUCCESS] <= 1e-12 * (1 + 1e-9)).all()


def test_coupled_exact_root_singular_and_counts(gs, orc, device):
    """test/nonlinear/batched.jl:263-375: bit-identical exact root with 0 iterations, a singular
    start, and the counts form."""
    P = np.array([[2.125, 1.5], [1.25, 2.0], [1.5, 2.5], [1.75, 3.0]])
    X0 = np.array([[2.0, 1.0], [2.5, 1.5], [3.5, 2.5], [4.0, 3.0]])
    opts = gs.Options(max_iterations=150)
    for m in (gs.Newton(), gs.Newton(globalization=gs.Backtracking())):
        x, st, xo, so = run_both(gs, orc, device, m, gs.Coupled2(), X0.copy(), gs.PerSystem(P), opts)
        assert_parity(st, x, so, xo, bitwise=True)
        assert st["iterations"][0] == 0 and np.array_equal(x[0], X0[0])
        assert (st["code"] == gs.SUCCESS).all()
        _, counts = run_gpu(gs, device, m, gs.Coupled2(), X0.copy(), gs.PerSystem(P), opts, counts=True)
        assert counts == dict(SUCCESS=4, STALLED=0, MAXITERS=0, SINGULAR=0, NONFINITE=0, LINESEARCH_FAILED=0)
    Xs = X0.copy()
    Xs[1] = [4.0, 4.0]
    x, st, xo, so = run_both(gs, orc, device, gs.Newton(), gs.Coupled2(), Xs, gs.PerSystem(P), gs.Options(max_iterations=60))
    assert list(st["code"]) == [gs.SUCCESS, gs.SINGULAR, gs.SUCCESS, gs.SUCCESS]
    assert_parity(st, x, so, xo, bitwise=True)
    # a shared parameter (not PerSystem) reaches every system
    x, st, xo, so = run_both(gs, orc, device, gs.Newton(), gs.Coupled2(), X0[1:].copy(), (2.125, 1.5), opts)
    assert_parity(st, x, so, xo, bitwise=True)


def test_six_codes_and_tally(gs, orc, device):
    """batched.jl:461-498: the six codes from one Backtracking batch with max_iterations = 3."""
    codes = np.array([1, 2, 3, 4, 5, 6, 1, 3], dtype=np.int64)
    X0 = np.tile([2.5, 3.5], (8, 1))
    m = gs.Newton(globalization=gs.Backtracking())
    opts = gs.Options(max_iterations=3)
    x, st, xo, so = run_both(gs, orc, device, m, gs.Hetero2(), X0.copy(), gs.PerSystem(codes), opts)
    assert list(st["code"]) == [0, 1, 2, 3, 4, 5, 0, 2]
    assert_parity(st, x, so, xo, bitwise=True)
    x2, counts = run_gpu(gs, device, m, gs.Hetero2(), X0.copy(), gs.PerSystem(codes), opts, counts=True)
    assert counts == dict(SUCCESS=2, STALLED=1, MAXITERS=2, SINGULAR=1, NONFINITE=1, LINESEARCH_FAILED=1)
    assert not gs.isconverged(counts)
    assert np.array_equal(x2[0], x[0]) and np.array_equal(x2[6], x[6])


def test_order_permutation(gs, orc, device):
    """batched.jl:501-583: order = perm solves the same systems into the same slots."""
    N = 7
    P = np.array([[1.0 + 0.2 * i, 1.5 + 0.2 * i] for i in range(1, N + 1)])
    X0 = np.array([[1.5 + 0.2 * i, 0.8 + 0.1 * i] for i in range(1, N + 1)])
    perm = [5, 2, 7, 1, 4, 6, 3]
    opts = gs.Options()
    x_u, st_u = run_gpu(gs, device, gs.Newton(), gs.Coupled2(), X0.copy(), gs.PerSystem(P), opts)
    x_o, st_o = run_gpu(gs, device, gs.Newton(), gs.Coupled2(), X0.copy(), gs.PerSystem(P), opts, order=perm)
    assert np.array_equal(x_u, x_o) and same_bytes(st_u, st_o)
    x_o2, st_o2 = run_gpu(gs, device, gs.Newton(), gs.Coupled2(), X0.copy(), gs.PerSystem(P), opts,
                          order=torch.tensor(perm))
    assert np.array_equal(x_u, x_o2)
    inv = np.argsort(np.array(perm) - 1)
    x_p, st_p = run_gpu(gs, device, gs.Newton(), gs.Coupled2(), X0[np.array(perm) - 1].copy(),
                        gs.PerSystem(P[np.array(perm) - 1]), opts)
    assert np.array_equal(x_p[inv], x_u)


def test_hostile_starts(gs, orc, device):
    """batched.jl:647-667: NaN/Inf starts → NONFINITE; a subnormal start converges."""
    X0 = np.array([[np.nan, 1.0], [np.inf, 1.0], [np.finfo(float).tiny / 16, 1.0], [2.0, 1.0]])
    P = np.tile([2.125, 1.5], (4, 1))
    x, st, xo, so = run_both(gs, orc, device, gs.Newton(), gs.Coupled2(), X0.copy(), gs.PerSystem(P), gs.Options())
    assert list(st["code"]) == [gs.NONFINITE, gs.NONFINITE, gs.SUCCESS, gs.SUCCESS]
    assert_parity(st, x, so, xo, bitwise=True)


def test_scalar_trajectory_and_best(gs, orc, device):
    """newton.jl:727-733: ‖F(x_k)‖ of Newton on x² + 1 from 0.5 for k = 0..8, and best = true."""
    best_nr_f = (1.25, 1.5625, 1.0850694444444444, 3.4600428713151907, 1.2166349630462578,
                 1.7081737551644687, 1.0300641461766078, 8.823069005566088, 2.487724019230605)
    for k, want in enumerate(best_nr_f):
        x, st = run_gpu(gs, device, gs.Newton(), gs.ScalarSquare(), np.array([[0.5]]), -1.0, gs.Options(max_iterations=k))
        assert st[0]["fnorm"] == pytest.approx(want, rel=8 * EPS)
    x, st = run_gpu(gs, device, gs.Newton(best=True), gs.ScalarSquare(), np.array([[0.5]]), -1.0, gs.Options(max_iterations=8))
    assert st[0]["code"] == gs.MAXITERS and st[0]["fnorm"] == pytest.approx(best_nr_f[6], rel=8 * EPS)
    # the step test of solver.jl:199-224
    x, st = run_gpu(gs, device, gs.Newton(), gs.ScalarSquare(), np.array([[10.0]]), 2.0,
                    gs.Options(x_reltol=0.5, f_reltol=1e-12, kappa=1e-6))
    assert st[0]["code"] == gs.STALLED and st[0]["iterations"] == 4


def test_workgroup_sizes_agree(gs, workloads, device):
    """The result does not depend on the launch's workgroup size (batched.jl:183-195)."""
    n = 5000
    xk = workloads.henon_heiles_states(n, seed=3)
    opts = gs.Options(f_abstol=1e-12, f_reltol=0.0)
    ref = run_gpu(gs, device, gs.Newton(), gs.HenonHeilesMidpoint(), xk.copy(), gs.BatchParams(0.1, xk), opts)
    # sizes above the kernel's bound of 256 run as workgroups of 256, as any size up to the device's
    # maximum of 1024 reaches the reference's launch (batched.jl:200-210)
    for ws in (1, 7, 64, 256, 257, 1000, 1024):
        got = run_gpu(gs, device, gs.Newton(), gs.HenonHeilesMidpoint(), xk.copy(), gs.BatchParams(0.1, xk), opts,
                      workgroupsize=ws)
        assert np.array_equal(got[0], ref[0]) and same_bytes(got[1], ref[1])
    # above the device's maximum the binding refuses it before any launch, as an ArgumentError
    with pytest.raises(gs.ArgumentError, match="exceeds the device maximum of 1024"):
        run_gpu(gs, device, gs.Newton(), gs.HenonHeilesMidpoint(), xk.copy(), gs.BatchParams(0.1, xk), opts,
                workgroupsize=1025)


def test_batch_summary(gs, workloads, device):
    """batch_summary (batched.jl:80-89) against the host reduction of the same statuses."""
    n = 20000
    xk = workloads.henon_heiles_states(n, seed=4, e_lo=0.1, e_hi=0.166)
    X = torch.as_tensor(xk * 3.0, device=device)
    S = gs.status_buffer(n, device)
    gs.solve_batch(X, S, gs.NonlinearProblem(gs.HenonHeilesMidpoint()), gs.Newton(),
                   gs.BatchParams(2.0, torch.as_tensor(xk * 3.0, device=device)), max_iterations=6)
    st = gs.statuses(S)
    bs = gs.batch_summary(S)
    fin = st["fnorm"]
    want = math.sqrt(float(np.sum(fin.astype(np.longdouble) ** 2))) if np.isfinite(fin).all() else float("nan")
    assert bs.failure_count == int((st["code"] != 0).sum())
    assert bs.max_iterations == int(st["iterations"].max())
    assert bs.total_iterations == int(st["iterations"].sum())
    assert bs.all_converged == (bs.failure_count == 0)
    if math.isfinite(want):
        assert bs.fnorm == pytest.approx(want, rel=1e-12)
    empty = gs.status_buffer(0, device)
    assert gs.batch_summary(empty) == gs.BatchSummary(0.0, True, 0, 0, 0)


def test_status_store_paths_agree(gs, workloads, device):
    """The statuses of whole waves in natural order are stored through LDS, those of a partial wave,
    of a batch with `order` or of a status buffer that is not 16-byte aligned lane by lane: every
    path writes the same bytes (the batch of 5000 ends in a partial wave)."""
    n = 5000
    xk = workloads.henon_heiles_states(n, seed=5)
    opts = gs.Options(f_abstol=1e-12, f_reltol=0.0)
    ref = run_gpu(gs, device, gs.Newton(), gs.HenonHeilesMidpoint(), xk.copy(), gs.BatchParams(0.1, xk), opts)
    ident = list(range(1, n + 1))
    got = run_gpu(gs, device, gs.Newton(), gs.HenonHeilesMidpoint(), xk.copy(), gs.BatchParams(0.1, xk), opts,
                  order=ident)
    assert np.array_equal(got[0], ref[0]) and same_bytes(got[1], ref[1])
    # a status buffer one record (72 bytes) into an allocation: 8-byte but not 16-byte aligned
    big = gs.status_buffer(n + 1, device)
    S = big[1:]
    assert S.data_ptr() % 16 == 8
    X = torch.as_tensor(xk.copy(), device=device)
    gs.solve_batch(X, S, gs.NonlinearProblem(gs.HenonHeilesMidpoint()), gs.Newton(),
                   gs.BatchParams(0.1, torch.as_tensor(xk, device=device)), opts)
    assert np.array_equal(X.cpu().numpy(), ref[0]) and same_bytes(gs.statuses(S), ref[1])


def test_row_store_paths_agree(gs, workloads, device):
    """Rows of X of whole waves go through LDS only where X is 16-byte aligned: a batch of 3-entry
    rows (24 bytes) one row into an allocation takes the lane-by-lane store, with the same bytes."""
    rec, X0 = workloads.tanh_linear3_batch(3000, 41)
    opts = gs.Options(f_abstol=1e-12, f_reltol=0.0)
    prob = gs.NonlinearProblem(gs.TanhLinear3())
    p = gs.PerSystem(torch.as_tensor(rec, device=device))
    Xa = torch.as_tensor(X0, device=device).clone()
    Sa = gs.status_buffer(3000, device)
    gs.solve_batch(Xa, Sa, prob, gs.Newton(), p, opts)
    big = torch.empty((3001, 3), dtype=torch.float64, device=device)
    Xu = big[1:]
    assert Xu.data_ptr() % 16 == 8
    Xu.copy_(torch.as_tensor(X0, device=device))
    Su = gs.status_buffer(3000, device)
    gs.solve_batch(Xu, Su, prob, gs.Newton(), p, opts)
    assert torch.equal(Xa, Xu) and same_bytes(gs.statuses(Sa), gs.statuses(Su))


def test_record_load_paths_agree(gs, workloads, device):
    """The parameter records of whole waves arrive through LDS only where they are 16-byte aligned
    with a 16-byte multiple stride: records one double into an allocation take the lane-by-lane
    loads, with the same bytes out."""
    rec, X0 = workloads.tanh_linear3_batch(3000, 43)
    opts = gs.Options(f_abstol=1e-12, f_reltol=0.0)
    prob = gs.NonlinearProblem(gs.TanhLinear3())
    out = []
    for shift in (0, 1):
        flat = torch.empty(3000 * 12 + shift, dtype=torch.float64, device=device)
        R = flat[shift:].view(3000, 12)
        R.copy_(torch.as_tensor(rec, device=device))
        assert (R.data_ptr() % 16 == 0) == (shift == 0)
        X = torch.as_tensor(X0, device=device).clone()
        S = gs.status_buffer(3000, device)
        gs.solve_batch(X, S, prob, gs.Newton(), gs.PerSystem(R), opts)
        out.append((X, gs.statuses(S)))
    assert torch.equal(out[0][0], out[1][0]) and same_bytes(out[0][1], out[1][1])
