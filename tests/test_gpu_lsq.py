"""Parity of the least-squares batch (Newton + LevenbergMarquardt on exponential-sum fits) with the
CPU oracle, through the C ABI (gs_solve_expsum_lsq_f64).

The family's exponential is one fixed algorithm on both sides, and the device follows the oracle's
operation order (the Givens reduction of [J | −F] row by row, the damping rows, the left folds of
the gain ratio); where its column-sum ‖J‖_F could change a decision or a reported gradratio, it
re-measures with the reference's column-major fold. So every fit is compared bit for bit: the
72-byte status and the 8 doubles of its row of X.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _device_data(gs, device, K, m, off, Y):
    return gs.ExpSumData(torch.as_tensor(K, device=device), torch.as_tensor(m, device=device),
                         torch.as_tensor(off, device=device), torch.as_tensor(Y, device=device))


def _solve_gpu(gs, device, K, m, off, Y, X0, method, opts):
    X = torch.as_tensor(X0, device=device).clone()
    S = gs.status_buffer(len(X0), device)
    gs.solve_batch(X, S, gs.LeastSquaresProblem(gs.ExpSumFits()), method, _device_data(gs, device, K, m, off, Y),
                   opts)
    torch.cuda.synchronize()
    return X.cpu().numpy(), gs.statuses(S)


def _field_bits(a, name):
    v = np.ascontiguousarray(a[name])
    return v.view({1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[v.dtype.itemsize])


def _assert_bitwise(x, st, xo, so):
    """Every field of every status, and every entry of every iterate, equal in bits (NaN payloads
    included; the padding of the 72-byte layout is not compared)."""
    bad_x = np.nonzero((x.view(np.uint64) != xo.view(np.uint64)).any(axis=1))[0]
    bad = np.zeros(len(st), dtype=bool)
    for name in st.dtype.names:
        bad |= _field_bits(st, name) != _field_bits(so, name)
    bad_s = np.nonzero(bad)[0]
    msg = ""
    if bad_s.size:
        i = bad_s[0]
        msg = f"status of fit {i}: device {st[i]}, oracle {so[i]}"
    elif bad_x.size:
        i = bad_x[0]
        msg = f"x of fit {i}: device {x[i]}, oracle {xo[i]}"
    assert bad_s.size == 0 and bad_x.size == 0, (
        f"{bad_s.size} statuses and {bad_x.size} iterates differ in bits; first: {msg}")


def _lm(gs, **kw):
    return gs.Newton(globalization=gs.LevenbergMarquardt(**kw))


def test_expsum_bitwise_default(gs, orc, workloads, device):
    K, m, off, Y, X0, _ = workloads.expsum_fits(4096, seed=3)
    opts = gs.Options()
    method = _lm(gs)
    x, st = _solve_gpu(gs, device, K, m, off, Y, X0, method, opts)
    xo, so = orc.solve_expsum_batch(method, opts, X0, K, m, off, Y)
    _assert_bitwise(x, st, xo, so)
    codes = np.bincount(st["code"], minlength=6)
    print("codes", codes.tolist(), "stationary", int(st["stationary"].sum()))
    assert codes[0] > len(st) // 2


@pytest.mark.parametrize("case", [
    dict(opts=dict(norm="max")),
    dict(opts=dict(min_iterations=3, max_iterations=7)),
    dict(opts=dict(max_iterations=0)),
    dict(opts=dict(g_tol=0.0, rate_test=False, roundoff=False, max_iterations=40)),
    dict(opts=dict(f_abstol=1e-2, f_reltol=0.0)),
    dict(opts=dict(x_reltol=1e-6, kappa=0.5, max_stalls=1, roundoff=True)),
    dict(lm=dict(mu0=1e-2, rho_accept=0.25, mu_decrease=0.5, mu_increase=10.0, mu_max=1e6)),
    dict(lm=dict(mu0=1e-12, mu_min=1e-14, mu_max=1e-3)),
    dict(lm=dict(scaling="diagonal")),
    dict(lm=dict(scaling="diagonal", mu0=1e-2, mu_increase=10.0)),
])
def test_expsum_bitwise_options(gs, orc, workloads, device, case):
    K, m, off, Y, X0, _ = workloads.expsum_fits(1024, seed=11)
    o = dict(case.get("opts", {}))
    if o.get("norm") == "max":
        o["norm"] = gs.MaxNorm()
    opts = gs.Options(**o)
    method = _lm(gs, **case.get("lm", {}))
    x, st = _solve_gpu(gs, device, K, m, off, Y, X0, method, opts)
    xo, so = orc.solve_expsum_batch(method, opts, X0, K, m, off, Y)
    _assert_bitwise(x, st, xo, so)


def test_expsum_noise_free_and_truth_start(gs, orc, workloads, device):
    """Noise-free data: from the default start the fits reach the round-off floor or a stationary
    point; started at the truth, a fit ends at once or after steps at the floor."""
    K, m, off, Y, X0, truth = workloads.expsum_fits(2048, seed=5, sigma=0.0)
    opts = gs.Options()
    method = _lm(gs)
    for start in (X0, truth):
        x, st = _solve_gpu(gs, device, K, m, off, Y, start, method, opts)
        xo, so = orc.solve_expsum_batch(method, opts, start, K, m, off, Y)
        _assert_bitwise(x, st, xo, so)


@pytest.mark.parametrize("scaling", ["none", "diagonal"])
def test_expsum_hostile_starts(gs, orc, workloads, device, scaling):
    """Starts whose exponentials overflow (λ ≪ 0), NaN and Inf entries, zero amplitudes (zero
    columns of J), huge and tiny scales, and m = 2K (square fits); with scaling = :diagonal the zero
    columns take the floor and the tiny ones the scaled column norm."""
    K, m, off, Y, X0, _ = workloads.expsum_fits(512, seed=7)
    X = X0.copy()
    rng = np.random.default_rng(1)
    kinds = rng.integers(0, 7, size=len(X))
    for i, kd in enumerate(kinds):
        k = int(K[i])
        if kd == 0:
            X[i, k:2 * k] = -300.0          # e^(λ t) overflows at t = 5
        elif kd == 1:
            X[i, 0] = np.nan
        elif kd == 2:
            X[i, k] = np.inf
        elif kd == 3:
            X[i, :k] = 0.0                   # J columns of the λ block vanish
        elif kd == 4:
            X[i, :k] = 1e200
        elif kd == 5:
            X[i, :k] = 1e-200
    m2 = m.copy()
    sq = rng.random(len(X)) < 0.2
    m2[sq] = 2 * K[sq]
    opts = gs.Options()
    method = _lm(gs, scaling=scaling)
    x, st = _solve_gpu(gs, device, K, m2, off, Y, X, method, opts)
    xo, so = orc.solve_expsum_batch(method, opts, X, K, m2, off, Y)
    _assert_bitwise(x, st, xo, so)


def test_expsum_counts_form_and_order_invariance(gs, orc, workloads, device):
    """The counts form tallies the codes; the bucketing does not change any fit's result (a permuted
    batch gives the permuted results, bit for bit)."""
    K, m, off, Y, X0, _ = workloads.expsum_fits(3000, seed=9)
    opts = gs.Options()
    method = _lm(gs)
    x, st = _solve_gpu(gs, device, K, m, off, Y, X0, method, opts)
    X = torch.as_tensor(X0, device=device).clone()
    counts = gs.solve_batch(X, gs.LeastSquaresProblem(gs.ExpSumFits()), method,
                            _device_data(gs, device, K, m, off, Y), opts)
    want = np.bincount(st["code"], minlength=6)
    assert [counts[k] for k in ("SUCCESS", "STALLED", "MAXITERS", "SINGULAR", "NONFINITE",
                                "LINESEARCH_FAILED")] == want.tolist()
    perm = np.random.default_rng(2).permutation(len(X0))
    xp, sp = _solve_gpu(gs, device, K[perm], m[perm], off[perm], Y, X0[perm], method, opts)
    _assert_bitwise(xp, sp, x[perm], st[perm])


def test_expsum_rejects_bad_data(gs, workloads, device):
    K, m, off, Y, X0, _ = workloads.expsum_fits(64, seed=1)
    method = _lm(gs)
    prob = gs.LeastSquaresProblem(gs.ExpSumFits())
    for bad in ("K", "m", "off"):
        K2, m2, off2 = K.copy(), m.copy(), off.copy()
        if bad == "K":
            K2[5] = 5
        elif bad == "m":
            m2[7] = 2 * K2[7] - 1
        else:
            off2[9] = len(Y) - 1
        X = torch.as_tensor(X0, device=device).clone()
        with pytest.raises(gs.ArgumentError, match="outside"):
            gs.solve_batch(X, gs.status_buffer(len(X0), device), prob, method,
                           _device_data(gs, device, K2, m2, off2, Y))
        assert torch.equal(X.cpu(), torch.as_tensor(X0)), "no fit may be solved when the data is invalid"


def test_expsum_full_batch(gs, orc, workloads, device):
    """The 262144-fit batch (K ~ U{1..4}, m ~ U{16..128}): a sample of 4096 fits matches the oracle
    bit for bit, and every fit's reported ‖F‖ matches a host recheck of its final iterate."""
    K, m, off, Y, X0, _ = workloads.expsum_fits(262144, seed=0)
    opts = gs.Options()
    method = _lm(gs)
    x, st = _solve_gpu(gs, device, K, m, off, Y, X0, method, opts)
    sample = np.sort(np.random.default_rng(4).choice(len(X0), 4096, replace=False))
    xo, so = orc.solve_expsum_batch(method, opts, X0[sample], K[sample], m[sample], off[sample], Y)
    _assert_bitwise(x[sample], st[sample], xo, so)
    fin = np.isfinite(st["fnorm"])
    idx = np.nonzero(fin)[0][::16]
    host = workloads.expsum_residual_norms(x[idx], K[idx], m[idx], off[idx], Y)
    assert np.allclose(host, st["fnorm"][idx], rtol=1e-9, atol=0.0)
    codes = np.bincount(st["code"], minlength=6)
    print("codes", codes.tolist())


def test_expsum_graph_replays_match_eager(gs, workloads, device):
    """The captured launch sequence, replayed on the same buffers with eager work and
    synchronisations between the replays, gives the eager launch's results bit for bit every time."""
    lib = gs.load_library()
    K, m, off, Y, X0, _ = workloads.expsum_fits(6000, seed=12)
    method = _lm(gs)
    prob = gs.LeastSquaresProblem(gs.ExpSumFits())
    data = _device_data(gs, device, K, m, off, Y)
    try:
        assert lib.gs_lsq_set_graph(0) == 0
        X = torch.as_tensor(X0, device=device).clone()
        S = gs.status_buffer(len(X0), device)
        gs.solve_batch(X, S, prob, method, data)
        torch.cuda.synchronize()
        xe, se = X.cpu().numpy(), gs.statuses(S)
        assert lib.gs_lsq_set_graph(1) == 0
        Xg = torch.empty_like(X)
        Sg = gs.status_buffer(len(X0), device)
        for r in range(4):
            Xg.copy_(torch.as_tensor(X0, device=device))
            Sg.zero_()
            junk = (torch.arange(1 << 16, device=device, dtype=torch.float64) * r).sum()  # eager work
            torch.cuda.synchronize()
            gs.solve_batch(Xg, Sg, prob, method, data)
            torch.cuda.synchronize()
            _assert_bitwise(Xg.cpu().numpy(), gs.statuses(Sg), xe, se)
            assert float(junk) == float(r) * (1 << 16) * ((1 << 16) - 1) / 2
    finally:
        lib.gs_lsq_set_graph(1)


def test_tune_lsq_changes_no_result(gs, workloads, device):
    """The autotuner times candidate shapes per bucket on the batch and sets the fastest; results do
    not depend on the shape (every fit is one lane's, whichever lane takes it)."""
    K, m, off, Y, X0, _ = workloads.expsum_fits(20000, seed=13)
    method = _lm(gs)
    x0, s0 = _solve_gpu(gs, device, K, m, off, Y, X0, method, gs.Options())
    data = _device_data(gs, device, K, m, off, Y)
    try:
        best = gs.tune_lsq(torch.as_tensor(X0, device=device), data, candidates=((64, 0), (128, 1), (256, 0)),
                           repeats=1)
        assert sorted(best) == [1, 2, 3, 4]
        for k, (threads, per_cu, secs) in best.items():
            assert (threads, per_cu) == gs.autotune.get_shape(k)[:2] and secs > 0
        print("tuned", best)
        x1, s1 = _solve_gpu(gs, device, K, m, off, Y, X0, method, gs.Options())
        _assert_bitwise(x1, s1, x0, s0)
        # a shape with one resident workgroup per CU: fewer lanes than fits, every lane loops
        for k in (1, 2, 3, 4):
            gs.autotune.set_shape(k, 64, 1)
        x2, s2 = _solve_gpu(gs, device, K, m, off, Y, X0, method, gs.Options())
        _assert_bitwise(x2, s2, x0, s0)
    finally:
        for k in (1, 2, 3, 4):
            gs.autotune.set_shape(k, 64, 0)


def test_shutdown_releases_and_recreates(gs, workloads, device):
    """gs_shutdown frees the library's cached workspaces, graphs, streams and events; the next call
    creates them again and gives the same results bit for bit."""
    lib = gs.load_library()
    K, m, off, Y, X0, _ = workloads.expsum_fits(3000, seed=14)
    method = _lm(gs)
    x0, s0 = _solve_gpu(gs, device, K, m, off, Y, X0, method, gs.Options())
    assert lib.gs_shutdown() == 0
    x1, s1 = _solve_gpu(gs, device, K, m, off, Y, X0, method, gs.Options())
    _assert_bitwise(x1, s1, x0, s0)
    assert lib.gs_shutdown() == 0 and lib.gs_shutdown() == 0  # idempotent
    S = gs.status_buffer(len(X0), device)
    S.copy_(torch.as_tensor(s1.view(np.uint8).reshape(len(X0), 72), device=device))
    summ = gs.batch_summary(S)
    assert summ.failure_count == int((s1["code"] != gs.SUCCESS).sum())
