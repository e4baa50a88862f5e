"""The oracle's mixed-precision step pinned by the reference's own known-answer tests
(test/nonlinear/refinement.jl): the refinement passes and their stopping rules, the slow-step
reading, mixed Newton on the moderately and the ill-conditioned system with and without promotion,
one failure per promotion site, the regularised direction, the floor of F, and Check 1.

The systems are the reference's: f_mod_r1 / f_ill_r1 (refinement.jl:40-96) as the quadratic-map
family 9, f_floor_r1 (:131-139) as fixture 104, the regularised x.^2 − p (:289-296) as fixture 101
and the cubic of Check 1 (:690-720) as fixture 105. A mistake that these checks catch: a refinement
that keeps a non-decreasing pass, an η floor taken as η, a promotion that refactorises in Float32,
a slow-step count that does not reset, a failure counted twice at one site.
"""
import math

import numpy as np
import pytest

EPS = np.finfo(np.float64).eps
SQRT_EPS = math.sqrt(EPS)
QUAD = 9


def _quad_record(A, b, c=(0.1, -0.2)):
    """vec(A) column-major, b, c — A given as (a11, a21, a12, a22) like the reference's SMatrix."""
    return np.array([*A, *b, *c], dtype=np.float64)


def _star(A, xs=(1.0, 2.0)):
    a11, a21, a12, a22 = A
    u1 = a11 * xs[0] + a12 * xs[1]
    u2 = a21 * xs[0] + a22 * xs[1]
    return (u1 + 0.1 * u1 ** 2, u2 - 0.2 * u2 ** 2)


# refinement.jl:40-45 and :75-80
A_MOD = (1.3, 1.1, 1.1, 1.1 * 1.1 / 1.3 + 2e-5)
A_ILL = (1.0, 1.0, 1.0, 1.0 + 1e-7)
P_MOD = _quad_record(A_MOD, _star(A_MOD))
P_ILL = _quad_record(A_ILL, _star(A_ILL))
X0 = np.array([2.0, 3.0])
X_STAR = np.array([1.0, 2.0])


def _opts(gs, **kw):
    return gs.Options(f_abstol=1e-12, f_reltol=0.0, **kw)


def _mixed(gs, refinement=True, **kw):
    return gs.Newton(linsolve=gs.LUFactorization(np.float32),
                     refinement=gs.Refinement() if refinement else gs.NoRefinement(), **kw)


def test_the_systems_are_the_reference_s(orc):
    """B is F(x*) + b: the residual vanishes at x* = (1, 2) up to the rounding of A x*, and the
    conditions at the start x0 are those the reference's comments state (cond(J) ≈ 1.8e5 and
    ≈ 5e7)."""
    for p, cond in ((P_MOD, 1.8e5), (P_ILL, 5e7)):
        F, _ = orc.fixture_eval(QUAD, X_STAR, p)
        assert np.abs(F).max() <= 8 * EPS * 10
        _, J = orc.fixture_eval(QUAD, X0, p)
        c = np.linalg.cond(J)
        assert 0.5 * cond < c < 2 * cond
        # J = Diagonal(1 .+ 2c .* u) · A
        u = np.array([[p[0], p[2]], [p[1], p[3]]]) @ X0
        assert np.allclose(J, np.diag(1 + 2 * p[6:8] * u) @ np.array([[p[0], p[2]], [p[1], p[3]]]), rtol=1e-15)


def test_refine_unit_checks_and_stopping_rules(orc):
    """refine!! unit checks (refinement.jl:328-404) on f_mod_r1 at x0 = (2, 3)."""
    F, J = orc.fixture_eval(QUAD, X0, P_MOD)
    fnorm = orc.rnorm(F)
    d0 = -orc.lu_solve(J, F, f32=True)

    # NoRefinement returns d0 itself and failed = false
    d, failed = orc.refine_fixture(QUAD, P_MOD, X0, d0, "f32", refinement=None)
    assert np.array_equal(d, d0) and not failed

    # Refinement() brings ‖F + J d‖ below max(η, √eps) ‖F‖ from above 1e-4
    r0_rel = np.linalg.norm(F + J @ d0) / fnorm
    assert r0_rel > 1e-4
    d_ref, failed = orc.refine_fixture(QUAD, P_MOD, X0, d0, "f32")
    r_rel = np.linalg.norm(F + J @ d_ref) / fnorm
    assert not failed
    assert r_rel <= SQRT_EPS
    assert r_rel < r0_rel / 100

    # an exact direction passes at once: the same bits, 0 passes
    d_exact = -orc.lu_solve(J, F)
    d, failed = orc.refine_fixture(QUAD, P_MOD, X0, d_exact, "f32")
    assert np.array_equal(d, d_exact) and not failed

    # the floor η_eff = max(η, √eps): a relative residual 1.2e-8 ≤ √eps passes even with η = 1e-12
    d_near = d_exact + np.linalg.solve(J, 1.2e-8 * F)
    r_near = np.linalg.norm(F + J @ d_near) / fnorm
    assert 1e-12 < r_near <= SQRT_EPS
    d, failed = orc.refine_fixture(QUAD, P_MOD, X0, d_near, "f32", refinement=(1e-12, 1))
    assert np.array_equal(d, d_near) and not failed

    # maxiter exhausted above η_eff ‖F‖: the improved direction, failed = true
    d_one, failed = orc.refine_fixture(QUAD, P_MOD, X0, d0, "f32", refinement=(1e-12, 1))
    assert failed
    assert np.linalg.norm(F + J @ d_one) < np.linalg.norm(F + J @ d0)

    # a pass that does not strictly reduce the residual stops at once and keeps the best direction
    d_bad, failed = orc.refine_fixture(QUAD, P_MOD, X0, d0, "bad", refinement=(SQRT_EPS, 3))
    assert failed and np.array_equal(d_bad, d0)

    # ‖F‖ = 0: d unchanged, not failed; a NaN ‖F‖: d unchanged, failed
    d, failed = orc.refine_fixture(QUAD, P_MOD, X0, d0, "f32", zero_fx=True, fnorm=0.0)
    assert np.array_equal(d, d0) and not failed
    d, failed = orc.refine_fixture(QUAD, P_MOD, X0, d0, "f32", fnorm=float("nan"))
    assert np.array_equal(d, d0) and failed

    # the norm formed from Fx when the caller holds none: the same direction and flag
    d_with, f_with = orc.refine_fixture(QUAD, P_MOD, X0, d0, "f32", fnorm=fnorm)
    d_without, f_without = orc.refine_fixture(QUAD, P_MOD, X0, d0, "f32")
    assert np.array_equal(d_with, d_without) and f_with == f_without and not f_without


def test_read_mixed_stall(orc):
    """The mixed step reads a slow TF step and nothing else (refinement.jl:406-464)."""
    eta_lin = max(SQRT_EPS, 1e-4)
    slow = dict(fnorm=0.9, stepnorm=0.1, fnorm_old=1.0)
    s1 = orc.read_mixed_stall(False, 0, lin_rel=10 * eta_lin, **slow)
    assert not s1["promoted"] and s1["slow_steps"] == 1 and s1["failures"] == 0 and not s1["info_promoted"]
    s2 = orc.read_mixed_stall(s1["promoted"], s1["slow_steps"], lin_rel=10 * eta_lin, **slow)
    assert s2["promoted"] and s2["slow_steps"] == 0
    assert s2["info_promoted"] and s2["failures"] == 1
    assert s2["fnorm"] == 0.9 and s2["stepnorm"] == 0.1
    # an accurate linear model resets the count: slow, accurate, slow does not promote
    s3 = orc.read_mixed_stall(s1["promoted"], s1["slow_steps"], lin_rel=eta_lin / 10, **slow)
    assert not s3["promoted"] and s3["slow_steps"] == 0 and s3["failures"] == 0
    s4 = orc.read_mixed_stall(s3["promoted"], s3["slow_steps"], lin_rel=10 * eta_lin, **slow)
    assert not s4["promoted"] and s4["slow_steps"] == 1 and s4["failures"] == 0
    # a step that halves ‖F‖ reads nothing, whatever its linear model
    s5 = orc.read_mixed_stall(s1["promoted"], s1["slow_steps"], fnorm=0.4, stepnorm=0.1, fnorm_old=1.0,
                              lin_rel=10 * eta_lin)
    assert not s5["promoted"] and s5["slow_steps"] == 0 and s5["failures"] == 0


def test_mixed_newton_on_the_moderate_system(gs, orc):
    """Mixed Newton and state invariants (refinement.jl:466-531): the FP64 control, then
    LUFactorization(Float32) + Refinement() in the control's iteration count, never promoted."""
    x64, s64 = orc.solve_fixture(QUAD, gs.Newton(), _opts(gs), X0, P_MOD)
    assert s64["code"] == 0
    assert np.linalg.norm(x64 - X_STAR) < 1e-10
    x32, s32 = orc.solve_fixture(QUAD, _mixed(gs), _opts(gs), X0, P_MOD)
    assert s32["code"] == 0
    assert s32["fnorm"] < 1e-11
    assert not s32["promoted"]
    assert s32["step_failures"] == 0
    assert s32["iterations"] == s64["iterations"]
    assert np.linalg.norm(x32 - X_STAR) < 1e-10
    assert math.isfinite(s32["gradratio"]) and s32["gradratio"] > 0
    assert s32["rank"] == -1 and math.isnan(s32["nullnorm"])


def test_promotion_on_the_ill_conditioned_system(gs, orc):
    """Promotion to the working precision (refinement.jl:533-588): with Refinement() the refinement
    fails on the Float32 factor and promotes within the step, one failure; with NoRefinement() two
    slow steps promote."""
    x, s = orc.solve_fixture(QUAD, _mixed(gs), _opts(gs), X0, P_ILL)
    assert s["code"] == 0 and s["promoted"] and s["step_failures"] == 1
    assert s["fnorm"] < 1e-12
    assert np.linalg.norm(x - X_STAR) < 1e-5
    assert math.isfinite(s["gradratio"])
    x, s = orc.solve_fixture(QUAD, _mixed(gs, refinement=False), _opts(gs), X0, P_ILL)
    assert s["code"] == 0 and s["promoted"] and s["step_failures"] >= 1
    assert s["fnorm"] < 1e-12
    assert np.linalg.norm(x - X_STAR) < 1e-5


@pytest.mark.parametrize("site", ["static", "dogleg", "reuse3"])
def test_one_failure_per_promotion_site(gs, orc, site):
    """A promotion counts one step-rule failure at each site that can make it: the line search, the
    DogLeg trial loop and a reused factor, once in the step and once in the solve
    (refinement.jl:590-608)."""
    kw = {"static": {}, "dogleg": {"globalization": gs.DogLeg()}, "reuse3": {"reuse": 3}}[site]
    m = _mixed(gs, **kw)
    st = orc.fixture_step(QUAD, m, X0, P_ILL)
    assert st["promoted"] and st["failures"] == 1
    if site == "static":
        # promote_factor!! refactorises the same Jacobian in Float64 (newton.jl:1332-1339), so the
        # promoting step is the plain Float64 step, bit for bit
        plain = orc.fixture_step(QUAD, gs.Newton(), X0, P_ILL)
        assert np.array_equal(st["x"], plain["x"]) and st["fnorm"] == plain["fnorm"]
    _, s = orc.solve_fixture(QUAD, m, _opts(gs), X0, P_ILL)
    assert s["promoted"] and s["step_failures"] == 1


@pytest.mark.parametrize("f32", [False, True])
def test_refinement_refines_the_regularised_direction(gs, orc, f32):
    """The factor of a regularised step is that of J + μI, and a Refinement refines that system's
    direction: at x = (1/2, 1), p = (0, 1), μ = 1 the step is (−1/8, 0), not (−1/4, 0), with no
    promotion and no failure (refinement.jl:289-318; x.^2 − p is fixture 101)."""
    lin = gs.LUFactorization(np.float32) if f32 else gs.LUFactorization()
    m = gs.Newton(linsolve=lin, regularization=1.0, refinement=gs.Refinement())
    x = np.array([0.5, 1.0])
    st = orc.fixture_step(101, m, x, (0.0, 1.0))
    assert np.allclose(st["x"] - x, [-0.125, 0.0], rtol=8 * EPS, atol=0)
    assert not st["info_promoted"] and not st["promoted"]
    assert st["failures"] == 0


def test_the_floor_of_f_under_a_mixed_step(gs, orc):
    """F rounds through Float32, so neither precision reaches 1e-14: the control and the mixed run
    both end MAXITERS at the cap, and the mixed run promotes nothing (refinement.jl:722-760)."""
    opts = gs.Options(f_abstol=1e-14, f_reltol=1e-14, x_abstol=0.0, x_reltol=0.0, max_iterations=10, max_stalls=2)
    x0 = np.array([1.5, 1.8])
    _, sc = orc.solve_fixture(104, gs.Newton(), opts, x0)
    assert sc["code"] == 2 and sc["iterations"] == 10
    _, sf = orc.solve_fixture(104, _mixed(gs, refinement=False), opts, x0)
    assert sf["code"] == sc["code"] and sf["iterations"] == sc["iterations"]
    assert not sf["promoted"] and sf["step_failures"] == 0
    assert sf["fnorm"] <= 4 * np.finfo(np.float32).eps * math.hypot(2.0, 3.0)


def _cubic(rng):
    A = rng.standard_normal((2, 2)) + 3 * np.eye(2)
    c = rng.standard_normal(2)
    b = 5 * rng.standard_normal(2)
    x0 = rng.standard_normal(2)
    return np.array([*A.reshape(-1, order="F"), *c, *b]), x0


def test_check_1_mixed_keeps_going_where_its_control_does(gs, orc):
    """Check 1 (refinement.jl:662-720): on well-conditioned A x + c .* x.^3 − b the Float32 factor
    converges with the control's code and iteration count and never promotes; over a wider draw of
    the same generator no mixed run ends STALLED and none promotes. The draws are this test's own
    (numpy PCG64, seed 7): the reference's Xoshiro stream is Julia's; the checks are properties of
    the generator, not of particular draws. Some draws of this generator wander for tens of
    iterations before Newton's local phase (a cubic's global phase), where a Float32 and a Float64
    factor take different paths; the equal-count check takes the first three draws whose control
    converges within 12 iterations, the local regime the reference's three draws are in."""
    rng = np.random.default_rng(7)
    opts = gs.Options()
    m32 = gs.Newton(linsolve=gs.LUFactorization(np.float32))
    local = 0
    while local < 3:
        p, x0 = _cubic(rng)
        _, s64 = orc.solve_fixture(105, gs.Newton(), opts, x0, p)
        _, s32 = orc.solve_fixture(105, m32, opts, x0, p)
        assert not s32["promoted"]
        if s64["code"] != 0 or s64["iterations"] > 12:
            continue
        local += 1
        assert s32["code"] == s64["code"] and s32["iterations"] == s64["iterations"]
    codes, promoted = [], []
    for _ in range(30):
        p, x0 = _cubic(rng)
        _, s32 = orc.solve_fixture(105, m32, opts, x0, p)
        codes.append(int(s32["code"]))
        promoted.append(bool(s32["promoted"]))
    assert codes.count(1) == 0
    assert sum(promoted) == 0
