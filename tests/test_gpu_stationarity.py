"""Parity of the StationarityProblem batch (src/nonlinear/stationarity.jl:642-698) with the CPU oracle,
through the C ABI (gs_solve_batch_stationarity_f64), over the Hénon–Heiles two-step discrete action.

The device evaluates the same closed-form ∇S and ∇²S as the oracle, picks the rows E and columns U
with the same index sets and adds the datum where one is given, so every form is compared bit for
bit: iterates (the full states, whose entries outside U keep their bits) and all status bytes.
"""
import numpy as np
import pytest
import torch

from test_gpu_parity import assert_parity
from test_oracle_stationarity import H, grad_cs, states

pytestmark = pytest.mark.gpu

OPTS = dict(f_abstol=1e-13, f_reltol=0.0)


def _gpu(gs, device, prob, X0, p, method, opts, b=None, counts=False, order=None):
    X = torch.as_tensor(X0, device=device).clone()
    pd = gs.PerSystem(torch.as_tensor(p.data, device=device)) if isinstance(p, gs.PerSystem) else p
    kw = {} if b is None else {"b": b}
    if counts:
        tally = gs.solve_batch(X, prob, method, pd, opts, order=order, **kw)
        return X.cpu().numpy(), tally
    S = gs.status_buffer(X.shape[0], device)
    gs.solve_batch(X, S, prob, method, pd, opts, order=order, **kw)
    return X.cpu().numpy(), gs.statuses(S)


METHODS = {
    "static": lambda gs: gs.Newton(),
    "backtracking": lambda gs: gs.Newton(globalization=gs.Backtracking()),
    "refine-regularized": lambda gs: gs.Newton(refinement=gs.Refinement(), regularization=1e-6),
    "mixed-refine": lambda gs: gs.Newton(linsolve=gs.LUFactorization("float32"), refinement=gs.Refinement()),
    "best-backtracking": lambda gs: gs.Newton(globalization=gs.Backtracking(), best=True),
    "dogleg": lambda gs: gs.Newton(globalization=gs.DogLeg()),
    "reuse3": lambda gs: gs.Newton(reuse=3),
    "mixed-reuse2-backtracking": lambda gs: gs.Newton(linsolve=gs.LUFactorization("float32"), refinement=gs.Refinement(),
                                                      globalization=gs.Backtracking(), reuse=2),
}


@pytest.mark.parametrize("name", list(METHODS))
def test_discrete_euler_lagrange_bitwise(gs, orc, device, name):
    """E = (3, 4), U = (5, 6) (|E| = 2): 8192 steps of the variational midpoint rule, λ per system,
    starts q₊ = 2q − q₋ and, for a quarter of the batch, far-off starts; L2 and MaxNorm."""
    X0 = states(8192, 11)
    X0[::4, 4:6] += np.random.default_rng(12).uniform(-3, 3, size=(2048, 2))
    lam = np.random.default_rng(13).uniform(0.5, 1.5, size=8192)
    p = gs.PerSystem(np.stack([np.full(8192, H), lam], axis=1))
    act = gs.HenonHeilesAction()
    prob = gs.StationarityProblem(act, E=(3, 4), U=(5, 6))
    method = METHODS[name](gs)
    for opts in (gs.Options(max_iterations=40, **OPTS), gs.Options(norm=gs.MaxNorm(), max_iterations=40, **OPTS)):
        x, st = _gpu(gs, device, prob, X0, p, method, opts)
        xo, so = orc.solve_batch(method, act, X0.copy(), p, opts, stationarity=prob)
        assert_parity(st, x, so, xo, bitwise=True)
    assert (x[:, 0:4] == X0[:, 0:4]).all()
    print(name, np.bincount(st["code"], minlength=6).tolist())


def test_full_state_and_datum_forms_bitwise(gs, orc, device):
    """E = U = : (|E| = 6) with a shared datum, a PerSystem datum given as the keyword, permuted index
    sets, the counts form and an `order` permutation."""
    n = 4096
    X0 = states(n, 21)
    act = gs.HenonHeilesAction()
    p = (H, 1.0)
    rng = np.random.default_rng(22)
    bs = rng.uniform(-0.05, 0.05, size=(n, 6))
    opts = gs.Options(max_iterations=30, **OPTS)
    m = gs.Newton(globalization=gs.Backtracking())
    prob = gs.StationarityProblem(act, b=(0.01, -0.02, 0.03, 0.0, 0.02, -0.01))
    x, st = _gpu(gs, device, prob, X0, p, m, opts)
    xo, so = orc.solve_batch(m, act, X0.copy(), p, opts, stationarity=prob)
    assert_parity(st, x, so, xo, bitwise=True)
    x, st = _gpu(gs, device, prob, X0, p, m, opts, b=gs.PerSystem(bs))
    xo, so = orc.solve_batch(m, act, X0.copy(), p, opts, stationarity=prob, b=gs.PerSystem(bs))
    assert_parity(st, x, so, xo, bitwise=True)
    perm = gs.StationarityProblem(act, E=(4, 3), U=(6, 5), b=0.01)
    order = rng.permutation(n) + 1
    x, counts = _gpu(gs, device, perm, X0, p, gs.Newton(), opts, counts=True, order=order)
    xo, so = orc.solve_batch(gs.Newton(), act, X0.copy(), p, opts, stationarity=perm)
    assert (x.view(np.uint64) == xo.view(np.uint64)).all()
    assert counts["SUCCESS"] == int((so["code"] == gs.SUCCESS).sum())
    for i in range(0, n, 512):
        assert np.abs(grad_cs(x[i], H, 1.0)[2:4] + 0.01).max() <= 2e-13


@pytest.mark.parametrize("name", list(METHODS))
def test_full_state_every_method_bitwise(gs, orc, device, name):
    """E = U = : (|E| = 6), the other compiled width, under every method form: the 6 × 6 LU, the
    reused factor and the DogLeg model at that size."""
    n = 2048
    X0 = states(n, 31)
    X0[::3, 4:6] += np.random.default_rng(32).uniform(-1, 1, size=(len(X0[::3]), 2))
    act = gs.HenonHeilesAction()
    p = (H, 1.0)
    prob = gs.StationarityProblem(act, b=(0.01, -0.02, 0.03, 0.0, 0.02, -0.01))
    method = METHODS[name](gs)
    opts = gs.Options(max_iterations=30, **OPTS)
    x, st = _gpu(gs, device, prob, X0, p, method, opts)
    xo, so = orc.solve_batch(method, act, X0.copy(), p, opts, stationarity=prob)
    assert_parity(st, x, so, xo, bitwise=True)


def test_refusals(gs, device):
    act = gs.HenonHeilesAction()
    X = torch.zeros((4, 6), dtype=torch.float64, device=device)
    with pytest.raises(gs.ArgumentError, match="\\|E\\| = 2 or 6"):
        gs.solve_batch(X, gs.StationarityProblem(act, E=(1, 2, 3), U=(4, 5, 6)), gs.Newton(), (H, 1.0))
    with pytest.raises(gs.ArgumentError, match="StationarityProblem"):
        gs.solve_batch(X, gs.NonlinearProblem(act), gs.Newton(), (H, 1.0))
    with pytest.raises(gs.DimensionMismatch):
        gs.solve_batch(X, gs.StationarityProblem(act, E=(3, 4), U=(5, 6)), gs.Newton(), (H, 1.0),
                       b=gs.PerSystem(np.zeros((3, 2))))


F32_METHODS = {
    "static": lambda gs: gs.Newton(),
    "backtracking": lambda gs: gs.Newton(globalization=gs.Backtracking()),
    "dogleg": lambda gs: gs.Newton(globalization=gs.DogLeg()),
    "refine-regularized": lambda gs: gs.Newton(refinement=gs.Refinement(), regularization=1e-3),
    "reuse3": lambda gs: gs.Newton(reuse=3),
    "best-backtracking": lambda gs: gs.Newton(globalization=gs.Backtracking(), best=True),
}


def _assert_f32_bitwise(x, st, xo, so):
    """Float32 iterates bit for bit and the 42 defined bytes of SolverStatus{Float32}."""
    assert x.dtype == np.float32 and st.dtype.itemsize == 44
    assert np.array_equal(x.view(np.uint32), xo.view(np.uint32))
    sb = st.view(np.uint8).reshape(len(st), 44)[:, :42]
    ob = so.view(np.uint8).reshape(len(so), 44)[:, :42]
    bad = np.nonzero((sb != ob).any(axis=1))[0]
    assert bad.size == 0, f"{bad.size} statuses differ; first {st[bad[0]]} vs {so[bad[0]]}"


@pytest.mark.parametrize("name", list(F32_METHODS))
def test_float32_batch_bitwise(gs, orc, device, name):
    """gs_solve_batch_stationarity_f32: |E| = 2 with λ per system (Float32 records) and far-off starts,
    |E| = 6 with a PerSystem Float32 datum; L2 and MaxNorm; bit for bit against the oracle's Float32 solve."""
    n = 8192
    X0 = states(n, 51)
    X0[::4, 4:6] += np.random.default_rng(52).uniform(-3, 3, size=(n // 4, 2))
    X0 = X0.astype(np.float32)
    lam = np.random.default_rng(53).uniform(0.5, 1.5, size=n)
    p = gs.PerSystem(np.stack([np.full(n, H), lam], axis=1).astype(np.float32))
    act = gs.HenonHeilesAction()
    method = F32_METHODS[name](gs)
    prob = gs.StationarityProblem(act, E=(3, 4), U=(5, 6))
    for opts in (gs.Options(eltype="float32", f_abstol=1e-5, max_iterations=40),
                 gs.Options(eltype="float32", norm=gs.MaxNorm(), f_abstol=1e-5, max_iterations=40)):
        X = torch.as_tensor(X0, device=device).clone()
        S = gs.status_buffer(n, device, "float32")
        gs.solve_batch(X, S, prob, method, gs.PerSystem(torch.as_tensor(p.data, device=device)), opts)
        xo, so = orc.solve_batch(method, act, X0.copy(), p, opts, stationarity=prob)
        _assert_f32_bitwise(X.cpu().numpy(), gs.statuses(S), xo, so)
    bs = np.random.default_rng(54).uniform(-0.05, 0.05, size=(2048, 6)).astype(np.float32)
    full = gs.StationarityProblem(act)
    opts = gs.Options(eltype="float32", f_abstol=1e-5, max_iterations=30)
    X = torch.as_tensor(X0[:2048], device=device).clone()
    S = gs.status_buffer(2048, device, "float32")
    gs.solve_batch(X, S, full, method, (H, 1.0), opts, b=gs.PerSystem(torch.as_tensor(bs, device=device)))
    xo, so = orc.solve_batch(method, act, X0[:2048].copy(), (H, 1.0), opts, stationarity=full, b=gs.PerSystem(bs))
    st = gs.statuses(S)
    _assert_f32_bitwise(X.cpu().numpy(), st, xo, so)
    print(name, np.bincount(st["code"], minlength=6).tolist())
