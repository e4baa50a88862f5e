"""Parity of the Float32 batch (gs_solve_batch_f32, SolverStatus{Float32}) with the CPU oracle's
Float32 solve, through the C ABI. The device's Float32 arithmetic (correctly rounded division and
square root, explicit FMA only where StaticArrays' product has muladd) follows the oracle's operation
order, so the polynomial families agree bit for bit; the tanh family to the tolerance of its libm.
"""
import numpy as np
import pytest
import torch

from test_gpu_parity import run_gpu

pytestmark = pytest.mark.gpu

F32 = np.float32


def _opts(gs, **kw):
    return gs.Options(eltype="float32", **kw)


def _both(gs, orc, device, method, family, X0, p, opts, **kw):
    x, st = run_gpu(gs, device, method, family, X0, p, opts, **kw)
    xo, so = orc.solve_batch(method, family, X0.copy(), p, opts, order=kw.get("order"))
    return x, st, xo, so


def _assert_bitwise(x, st, xo, so):
    assert x.dtype == np.float32 and st.dtype.itemsize == 44
    bad_x = np.nonzero((x.view(np.uint32) != xo.view(np.uint32)).any(axis=1))[0]
    sb = st.view(np.uint8).reshape(len(st), 44)[:, :42]
    ob = so.view(np.uint8).reshape(len(so), 44)[:, :42]
    bad_s = np.nonzero((sb != ob).any(axis=1))[0]
    msg = f"status {bad_s[0]}: {st[bad_s[0]]} vs {so[bad_s[0]]}" if bad_s.size else (
        f"x {bad_x[0]}: {x[bad_x[0]]} vs {xo[bad_x[0]]}" if bad_x.size else "")
    assert bad_s.size == 0 and bad_x.size == 0, f"{bad_s.size} statuses, {bad_x.size} iterates differ; {msg}"


METHODS = {
    "static": lambda gs: gs.Newton(),
    "backtracking": lambda gs: gs.Newton(globalization=gs.Backtracking()),
    "dogleg": lambda gs: gs.Newton(globalization=gs.DogLeg()),
    "refine-regularized": lambda gs: gs.Newton(refinement=gs.Refinement(), regularization=1e-3),
    "reuse3-backtracking": lambda gs: gs.Newton(reuse=3, globalization=gs.Backtracking()),
    "best": lambda gs: gs.Newton(best=True),
    # LUFactorization(Float32) on a Float32 batch is the working precision: the same solve
    "f32-factor": lambda gs: gs.Newton(linsolve=gs.LUFactorization("float32")),
}


@pytest.mark.parametrize("name", list(METHODS))
def test_henon_heiles_f32_bitwise(gs, orc, workloads, device, name):
    xk = workloads.henon_heiles_states(65536, seed=11).astype(F32)
    m = METHODS[name](gs)
    x, st, xo, so = _both(gs, orc, device, m, gs.HenonHeilesMidpoint(), xk.copy(), gs.BatchParams(0.1, xk),
                          _opts(gs, f_abstol=1e-6))
    _assert_bitwise(x, st, xo, so)
    assert (st["code"] == gs.SUCCESS).mean() > 0.99
    if name == "f32-factor":
        x2, st2 = run_gpu(gs, device, gs.Newton(), gs.HenonHeilesMidpoint(), xk.copy(), gs.BatchParams(0.1, xk),
                          _opts(gs, f_abstol=1e-6))
        assert np.array_equal(x2.view(np.uint32), x.view(np.uint32)) and not st["promoted"].any()


def test_hard_steps_f32_bitwise(gs, orc, workloads, device):
    """Large steps from far-out states: mixed codes, Backtracking failures, DogLeg radius floors."""
    xk = (workloads.henon_heiles_states(16384, seed=2, e_lo=0.1, e_hi=0.166) * 3.0).astype(F32)
    for m in (gs.Newton(), gs.Newton(globalization=gs.Backtracking()), gs.Newton(globalization=gs.DogLeg()),
              gs.Newton(reuse=2, refinement=gs.Refinement())):
        for opts in (_opts(gs, f_abstol=1e-6, max_iterations=20),
                     _opts(gs, norm=gs.MaxNorm(), min_iterations=1, x_reltol=1e-4, roundoff=True)):
            x, st, xo, so = _both(gs, orc, device, m, gs.HenonHeilesMidpoint(), xk.copy(), gs.BatchParams(2.0, xk), opts)
            _assert_bitwise(x, st, xo, so)
            assert len(np.unique(st["code"])) >= 2


def test_coupled_six_codes_and_counts_f32(gs, orc, device):
    """test/nonlinear/batched.jl:263-498 in Float32: the exact dyadic root with 0 iterations, the
    counts form, and the six-code batch."""
    P = np.array([[2.125, 1.5], [1.25, 2.0], [1.5, 2.5], [1.75, 3.0]], dtype=F32)
    X0 = np.array([[2.0, 1.0], [2.5, 1.5], [3.5, 2.5], [4.0, 3.0]], dtype=F32)
    opts = _opts(gs, max_iterations=150)
    for m in (gs.Newton(), gs.Newton(globalization=gs.DogLeg()), gs.Newton(reuse=2)):
        x, st, xo, so = _both(gs, orc, device, m, gs.Coupled2(), X0.copy(), gs.PerSystem(P), opts)
        _assert_bitwise(x, st, xo, so)
        assert st["iterations"][0] == 0 and np.array_equal(x[0], X0[0]) and (st["code"] == gs.SUCCESS).all()
        _, counts = run_gpu(gs, device, m, gs.Coupled2(), X0.copy(), gs.PerSystem(P), opts, counts=True)
        assert counts["SUCCESS"] == 4
    codes = np.array([1, 2, 3, 4, 5, 6, 1, 3], dtype=np.int64)
    X0 = np.tile(np.array([2.5, 3.5], dtype=F32), (8, 1))
    x, st, xo, so = _both(gs, orc, device, gs.Newton(globalization=gs.Backtracking()), gs.Hetero2(), X0.copy(),
                          gs.PerSystem(codes), _opts(gs, max_iterations=3))
    _assert_bitwise(x, st, xo, so)
    assert list(st["code"]) == [0, 1, 2, 3, 4, 5, 0, 2]


def test_tanh_f32(gs, orc, workloads, device):
    """The n = 3 tanh family in Float32: tanhf is ROCm's on the device and glibc's on the host, so
    codes agree except within 1% of the tolerance and iterates to Float32 accuracy."""
    rec, x0 = workloads.tanh_linear3_batch(4096, 0)
    rec, x0 = rec.astype(F32), x0.astype(F32)
    tol = 1e-5
    opts = _opts(gs, f_abstol=tol, f_reltol=0.0, max_iterations=50)
    x, st, xo, so = _both(gs, orc, device, gs.Newton(), gs.TanhLinear3(), x0.copy(), gs.PerSystem(rec), opts)
    near = (np.abs(so["fnorm"] - F32(tol)) <= 0.01 * tol) | (np.abs(st["fnorm"] - F32(tol)) <= 0.01 * tol)
    assert (st["code"][~near] == so["code"][~near]).all()
    assert np.abs(x - xo)[~near].max() <= 1e-5 * max(np.abs(xo).max(), 1.0)
