"""The oracle's stopping rules pinned by the reference's own known-answer tests: the round-off stop,
the rate test's reading of the last two rates (test/nonlinear/stopping.jl:553-748), and rmaxnorm,
which the MaxNorm residual test reads (test/base/reductions.jl:250-275).

Every system here is exact in both precisions (steps and rates are dyadic), except FloorResidual,
whose iterates the test forms by hand from the oracle's own residual, as the reference's test does
from its own. A mistake that these checks catch: a round-off stop that reads a limited step or a NaN
rate, a rate test that reads only the last rate, `min_iterations` not bounding the stop, an rmaxnorm
that drops a NaN or returns −0.
"""
import math

import numpy as np
import pytest

ELTYPES = [False, True]  # f32
CYCLE, JUMP, LINEAR, JUMPJ, FLOOR = 109, 110, 106, 107, 108


def _opts(gs, f32, **kw):
    return gs.Options(eltype="float32" if f32 else "float64", **kw)


def _T(f32):
    return np.float32 if f32 else np.float64


@pytest.mark.parametrize("f32", ELTYPES)
def test_roundoff_stop_skips_a_limited_step(gs, orc, f32):
    """stopping.jl:565-580: rate 2 after rate 1/2 stops at step 3, unless a radius limits the steps."""
    opt = _opts(gs, f32, f_reltol=0.0, roundoff=True, max_iterations=5)
    x, st = orc.solve_fixture(CYCLE, gs.Newton(), opt, [1.0], [0.0], f32=f32)
    assert st["code"] == 1 and st["iterations"] == 3
    assert st["rate"] == 2
    assert x[0] == -0.25
    _, st_l = orc.solve_fixture(CYCLE, gs.Newton(), opt, [1.0], [1.0], f32=f32)
    assert st_l["code"] == 2 and st_l["iterations"] == 5
    assert not st_l["contracted"]


@pytest.mark.parametrize("f32", ELTYPES)
def test_rate_test_reads_the_larger_of_the_last_two_rates(gs, orc, f32):
    """stopping.jl:592-628."""
    T = _T(f32)
    # JumpStep: rate 1/1024 after rate 1 fails θ < 1 through the larger rate; the step test stalls
    x, st = orc.solve_fixture(JUMP, gs.Newton(), _opts(gs, f32, f_reltol=0.0, x_abstol=1e-2), [0.0], [0.0],
                              f32=f32)
    assert st["code"] == 1 and not st["contracted"]
    assert st["iterations"] == 3
    assert st["rate"] == 1 / 1024
    assert x[0] == 2 + 1 / 1024
    # rate 2 after 1/2 at the step norm where the smaller rate's bound holds with equality
    _, st_c = orc.solve_fixture(CYCLE, gs.Newton(),
                                _opts(gs, f32, f_reltol=0.0, x_abstol=1.0, kappa=0.5, min_iterations=3),
                                [1.0], [0.0], f32=f32)
    assert st_c["code"] == 1 and not st_c["contracted"]
    assert st_c["iterations"] == 3 and st_c["rate"] == 2
    # the first rate stands beside the NaN of the first step: no rate, and the step test stalls
    a = float(T(99) / T(100))
    _, st_lin = orc.solve_fixture(LINEAR, gs.Newton(), _opts(gs, f32, f_reltol=0.0, x_abstol=0.5), [1.0],
                                  [a, 1.0], f32=f32)
    assert st_lin["code"] == 1 and not st_lin["contracted"]
    assert st_lin["iterations"] == 2


def test_a_nan_rate_never_stops_a_solve(gs, orc):
    """stopping.jl:647-656: rate NaN, then 3/2 ≥ 1 with no rate below 1 before it."""
    _, st = orc.solve_fixture(LINEAR, gs.Newton(), gs.Options(roundoff=True, max_iterations=5), [1.0], [-0.5, 1.0])
    assert st["code"] == 2 and st["iterations"] == 5


@pytest.mark.parametrize("f32", ELTYPES)
def test_roundoff_stop_where_the_increment_stops_decreasing(gs, orc, f32):
    """stopping.jl:668-709: FloorResidual with δ = 64√eps(T), J = 1, from x₀ = 3."""
    T = _T(f32)
    eps = float(np.finfo(T).eps)
    delta = float(T(64) * np.sqrt(T(eps)))

    def F(x):
        return orc.fixture_eval(FLOOR, [x], [delta], f32=f32)[0][0]

    xs = [3.0]
    for _ in range(60):
        xs.append(float(T(xs[-1]) - T(F(xs[-1]))))
    steps = [abs(F(xs[k])) for k in range(60)]
    thetas = [math.nan] + [float(T(steps[k]) / T(steps[k - 1])) for k in range(1, 60)]
    # 1-based k as the reference: step k has θ_k = thetas[k - 1]
    k_stop = next(k for k in range(2, 61) if thetas[k - 1] >= 1 and thetas[k - 2] < 1)
    assert k_stop < 60
    assert all(s != 0 for s in steps)

    def opt(roundoff, **kw):
        return _opts(gs, f32, f_reltol=0.0, max_iterations=60, roundoff=roundoff, **kw)

    x, st = orc.solve_fixture(FLOOR, gs.Newton(), opt(True), [3.0], [delta], f32=f32)
    assert st["code"] == 1
    assert st["iterations"] == k_stop < 60
    assert x[0] == xs[k_stop]
    assert math.isclose(st["rate"], thetas[k_stop - 1], rel_tol=8 * eps)
    assert st["rate"] >= 1

    x_off, st_off = orc.solve_fixture(FLOOR, gs.Newton(), opt(False), [3.0], [delta], f32=f32)
    assert st_off["code"] == 2 and st_off["iterations"] == 60
    assert x_off[0] == xs[-1]

    k_min = next(k for k in range(k_stop + 1, 61) if thetas[k - 1] >= 1 and thetas[k - 2] < 1)
    x_min, st_min = orc.solve_fixture(FLOOR, gs.Newton(), opt(True, min_iterations=k_stop + 1), [3.0], [delta],
                                      f32=f32)
    assert st_min["code"] == 1 and st_min["iterations"] == k_min
    assert x_min[0] == xs[k_min]


@pytest.mark.parametrize("f32", ELTYPES)
def test_roundoff_stop_at_a_rate_of_exactly_one(gs, orc, f32):
    """stopping.jl:711-731: F = x from 4 with J = −1 at 1 and 2 elsewhere: steps −2, −1, +1."""
    x, st = orc.solve_fixture(JUMPJ, gs.Newton(), _opts(gs, f32, roundoff=True), [4.0], [0.0], f32=f32)
    assert st["code"] == 1 and st["iterations"] == 3
    assert st["rate"] == 1 and x[0] == 2
    _, st_off = orc.solve_fixture(JUMPJ, gs.Newton(), _opts(gs, f32), [4.0], [0.0], f32=f32)
    assert st_off["code"] == 2
    # J = 1/2 maps x to −x: every rate is 1, with no rate below 1 before it
    _, st_one = orc.solve_fixture(LINEAR, gs.Newton(), _opts(gs, f32, roundoff=True, max_iterations=6), [1.0],
                                  [1.0, 0.5], f32=f32)
    assert st_one["code"] == 2 and st_one["iterations"] == 6 and st_one["rate"] == 1


def _is_pos_zero(v):
    return v == 0 and math.copysign(1.0, v) > 0


@pytest.mark.parametrize("f32", ELTYPES)
def test_rmaxnorm_kats(orc, f32):
    """rmaxnorm is the largest modulus (reductions.jl:250-275)."""
    T = _T(f32)
    fi = np.finfo(T)
    fmax, fmin, eps, tiny = float(fi.max), float(fi.tiny), float(fi.eps), float(fi.smallest_subnormal)
    r = lambda v: orc.rmaxnorm(v, f32=f32)  # noqa: E731
    assert r([3, -4, 1]) == 4
    assert _is_pos_zero(r([]))
    assert _is_pos_zero(r([-0.0, 0.0]))
    assert r([tiny, fmin]) == fmin
    assert r([eps, -1]) == 1
    assert r([fmax] * 4) == fmax
    assert r([fmax / 2, -fmax]) == fmax
    for v in ([1, math.nan, 2], [math.nan, 1, 2], [1, 2, math.nan], [math.inf, math.nan], [math.nan, math.inf]):
        assert math.isnan(r(v))
    assert r([1, -math.inf]) == math.inf
    assert r([math.inf, 1]) == math.inf
    a = np.array([3.0, -4.0, 1.0])
    for k in range(-8, 9, 4):
        b = (T(10.0 ** k) * a.astype(T)).astype(T)
        assert r(b) == float(np.max(np.abs(b)))
    assert r([-3]) == 3
