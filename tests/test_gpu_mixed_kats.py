"""The device against the oracle on the systems of the reference's mixed-precision tests
(test/nonlinear/refinement.jl:40-96), whose oracle behaviour tests/test_oracle_mixed.py pins with the
reference's known answers: f_mod_r1 (a Float32 factor with Refinement recovers the Float64 step,
never promoted) and f_ill_r1 (the refinement fails and promotes once), as the quadratic-map family 9.

The batch holds the two reference systems from x0 = (2, 3) in its first rows, then perturbed copies of
both from spread starts, so every promotion path (within the step on a failed refinement, after two
slow steps without refinement, at each site: line search, DogLeg, reused factor) runs on the device.
The family is a polynomial, so the device and the oracle agree bit for bit.
"""
import numpy as np
import pytest

from test_gpu_parity import HH_METHODS, assert_parity, run_both, run_gpu
from test_oracle_mixed import P_ILL, P_MOD, X0

pytestmark = pytest.mark.gpu


def _batch(n, seed=3):
    rng = np.random.default_rng(seed)
    P = np.empty((n, 8))
    X = np.empty((n, 2))
    P[0], P[1] = P_MOD, P_ILL
    X[0] = X[1] = X0
    base = np.where(rng.random(n - 2)[:, None] < 0.5, P_MOD, P_ILL)
    # perturb A relative 1e-9 (keeps each system's conditioning), b so that x* moves a little
    P[2:] = base
    P[2:, :4] *= 1 + 1e-9 * rng.standard_normal((n - 2, 4))
    P[2:, 4:6] *= 1 + 1e-3 * rng.standard_normal((n - 2, 2))
    X[2:] = rng.uniform(0.5, 3.5, size=(n - 2, 2))
    return X, P


@pytest.mark.parametrize("name", list(HH_METHODS))
def test_quadmap_every_method_bitwise(gs, orc, device, name):
    X, P = _batch(4096)
    m = HH_METHODS[name](gs)
    opts = gs.Options(f_abstol=1e-12, f_reltol=0.0)
    x, st, xo, so = run_both(gs, orc, device, m, gs.QuadMap2(), X, gs.PerSystem(P), opts)
    assert_parity(st, x, so, xo, bitwise=True)


def test_quadmap_device_keeps_the_reference_answers(gs, orc, device):
    """The reference's assertions, read off the device's statuses (refinement.jl:498-545, 560-608)."""
    X, P = _batch(256)
    opts = gs.Options(f_abstol=1e-12, f_reltol=0.0)
    _, s64 = run_gpu(gs, device, gs.Newton(), gs.QuadMap2(), X, gs.PerSystem(P), opts)
    mixed = gs.Newton(linsolve=gs.LUFactorization("float32"), refinement=gs.Refinement())
    x, s = run_gpu(gs, device, mixed, gs.QuadMap2(), X, gs.PerSystem(P), opts)
    # f_mod_r1: the control's iteration count, no promotion, no failure
    assert s["code"][0] == 0 and not s["promoted"][0] and s["step_failures"][0] == 0
    assert s["iterations"][0] == s64["iterations"][0]
    assert np.linalg.norm(x[0] - [1.0, 2.0]) < 1e-10
    # f_ill_r1: promoted within the step, exactly one failure
    assert s["code"][1] == 0 and s["promoted"][1] and s["step_failures"][1] == 1
    assert s["fnorm"][1] < 1e-12 and np.linalg.norm(x[1] - [1.0, 2.0]) < 1e-5
    for kw in ({"globalization": gs.DogLeg()}, {"reuse": 3}):
        m = gs.Newton(linsolve=gs.LUFactorization("float32"), refinement=gs.Refinement(), **kw)
        _, s_site = run_gpu(gs, device, m, gs.QuadMap2(), X, gs.PerSystem(P), opts)
        assert s_site["promoted"][1] and s_site["step_failures"][1] == 1
    # without refinement, two slow steps promote
    nr = gs.Newton(linsolve=gs.LUFactorization("float32"))
    _, s_nr = run_gpu(gs, device, nr, gs.QuadMap2(), X, gs.PerSystem(P), opts)
    assert s_nr["code"][1] == 0 and s_nr["promoted"][1] and s_nr["step_failures"][1] >= 1


def test_quadmap_float32_batch_bitwise(gs, orc, device):
    X, P = _batch(4096)
    opts = gs.Options(eltype="float32", f_abstol=1e-6, f_reltol=0.0)
    for m in (gs.Newton(), gs.Newton(globalization=gs.Backtracking()), gs.Newton(globalization=gs.DogLeg()),
              gs.Newton(reuse=3)):
        X32, P32 = X.astype(np.float32), P.astype(np.float32)
        x, st, xo, so = run_both(gs, orc, device, m, gs.QuadMap2(), X32, gs.PerSystem(P32), opts)
        assert np.array_equal(x.view(np.uint32), xo.view(np.uint32))
        assert np.array_equal(st.view(np.uint8).reshape(-1, 44)[:, :42], so.view(np.uint8).reshape(-1, 44)[:, :42])
