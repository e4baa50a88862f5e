"""The learned start on the GPU: its MFMA training kernels against the CPU oracle of the network
(oracle/mlp_ref.py, which emulates the device's bfloat16 rounding points), its inference kernel
against the same oracle, and what the trained network gains over the plain method (the ablation):
the start's residual, the Newton iterations, and identical converged results, with plain Newton as
the fallback."""
import numpy as np
import pytest
import torch

from oracle import mlp_ref as M

pytestmark = pytest.mark.gpu

H = 0.1
TOL = dict(f_abstol=1e-12, f_reltol=0.0)


def _pairs(gs, workloads, device, n, seed):
    xk = workloads.henon_heiles_states(n, seed=seed)
    Xk = torch.as_tensor(xk, device=device)
    X1, S = gs.generate_solutions(gs, Xk, H, gs.Options(**TOL))
    assert (gs.statuses(S)["code"] == 0).all()
    return Xk, X1


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_gradient_against_oracle(gs, workloads, device):
    """gs_mlp_grad (BF16 MFMA forward and backward) against the oracle's gradient with the same bfloat16
    rounding points: the loss to 1e-5 and every parameter block to 1e-2 (relative, FP32 vs FP64
    accumulation); this also pins the MFMA lane layouts the kernels assume."""
    Xk, X1 = _pairs(gs, workloads, device, 4096, seed=21)
    net = gs.LearnedStart(device, seed=5)
    net.fit_normalization(Xk, X1, H)
    rng = np.random.default_rng(8)
    P = gs.learned.initial_parameters(5).astype(np.float64)
    P[:] += rng.standard_normal(P.size) * 0.05  # biases and the padded rows nonzero too
    net.set_parameters(P.astype(np.float32))
    Pf = net.params.cpu().numpy().astype(np.float64)
    idx = torch.as_tensor(rng.choice(4096, 512, replace=False).astype(np.int64), device=device)
    loss, g = net.grad(Xk, X1, H, idx)
    torch.cuda.synchronize()
    loss, g = float(loss.item()), g.cpu().numpy().astype(np.float64)
    ii = idx.cpu().numpy()
    xk, x1 = Xk.cpu().numpy()[ii], X1.cpu().numpy()[ii]
    U = M.features(xk, H, net.norm)
    T = M.targets(xk, x1, H, net.norm)
    lref, gref = M.loss_and_grad(Pf, U, T, emulate=True)
    assert abs(loss - lref) <= 1e-5 * lref, (loss, lref)
    o = 0
    for name, shape in M.SHAPES:
        n = int(np.prod(shape))
        a, b = g[o:o + n], gref[o:o + n]
        if name == "W1":  # the padded input rows see zero features
            a, b = a.reshape(shape)[:5], b.reshape(shape)[:5]
            assert np.all(g[o:o + n].reshape(shape)[5:] == 0.0)
        assert _rel(a, b) <= 1e-2, (name, _rel(a, b))
        o += n


def test_inference_against_oracle(gs, workloads, device):
    """gs_mlp_start_henon_heiles against the oracle's prediction (same rounding points), on a batch
    that is not a multiple of the 16-system tile."""
    Xk, X1 = _pairs(gs, workloads, device, 4096 + 5, seed=22)
    net = gs.LearnedStart(device, seed=6)
    net.fit_normalization(Xk, X1, H)
    P = gs.learned.initial_parameters(6) + np.random.default_rng(1).standard_normal(M.NPARAM).astype(np.float32) * 0.05
    net.set_parameters(P)
    X0 = net.start(Xk, H)
    torch.cuda.synchronize()
    ref = M.predict(net.params.cpu().numpy().astype(np.float64), net.norm, Xk.cpu().numpy(), H, emulate=True)
    x0 = X0.cpu().numpy()
    base = M.rk2(Xk.cpu().numpy(), H)
    corr_dev = (x0 - base) / H ** 3
    corr_ref = (ref - base) / H ** 3
    assert np.abs(corr_dev - corr_ref).max() <= 2e-2 * np.abs(corr_ref).max()
    assert np.isfinite(x0).all()


def test_training_gains_over_plain_and_falls_back(gs, workloads, device):
    """Train on 2^17 method-generated solutions; on a held-out batch of 2^16 the learned start's
    residual is far below the default start's, Newton needs fewer iterations, every system converges
    (codes equal the plain method's) and the solutions agree with the plain method's to 1e-10. A
    network that predicts NaN falls back to plain Newton for every system, with the plain statuses."""
    Xk, X1 = _pairs(gs, workloads, device, 1 << 17, seed=31)
    net = gs.LearnedStart(device, seed=0)
    net.fit_normalization(Xk, X1, H)
    hist = net.train(Xk, X1, H, epochs=16, batch=2048, lr=3e-3, lr_final=3e-5, seed=0)
    print("loss by epoch", [f"{v:.3e}" for v in hist])
    assert hist[-1] < 0.05 * hist[0]
    xt = torch.as_tensor(workloads.henon_heiles_states(1 << 16, seed=32), device=device)
    prob = gs.NonlinearProblem(gs.HenonHeilesMidpoint())
    p = gs.BatchParams(H, xt)
    opts = gs.Options(**TOL)
    # plain
    Xp = xt.clone()
    Sp = gs.status_buffer(len(xt), device)
    gs.solve_batch(Xp, Sp, prob, gs.Newton(), p, opts)
    sp = gs.statuses(Sp)
    # learned start residuals vs the default start's
    x0 = net.start(xt, H).cpu().numpy()
    xk = xt.cpu().numpy()
    r_learned = np.linalg.norm(workloads.henon_heiles_midpoint_residual(x0, xk, H), axis=1)
    r_default = np.linalg.norm(workloads.henon_heiles_midpoint_residual(xk, xk, H), axis=1)
    print("median start residual: learned", np.median(r_learned), "default", np.median(r_default))
    assert np.median(r_learned) <= 1e-3 * np.median(r_default)
    assert np.median(r_learned) <= 1e-5
    # learned + mixed (FP32 factor, FP64 refinement)
    mixed = gs.Newton(linsolve=gs.LUFactorization(np.float32), refinement=gs.Refinement())
    for method in (gs.Newton(), mixed):
        Xl = xt.clone()
        Sl = gs.status_buffer(len(xt), device)
        nb = gs.solve_batch_learned(gs, Xl, Sl, prob, method, p, net, opts)
        sl = gs.statuses(Sl)
        print("fallbacks", nb, "mean iterations learned", sl["iterations"].mean(), "plain", sp["iterations"].mean())
        assert (sl["code"] == sp["code"]).all()
        assert sl["iterations"].mean() < sp["iterations"].mean()
        xl, xpl = Xl.cpu().numpy(), Xp.cpu().numpy()
        assert np.abs(xl - xpl).max() <= 1e-10 * np.abs(xpl).max()
        assert (np.linalg.norm(workloads.henon_heiles_midpoint_residual(xl, xk, H), axis=1) <= 1e-11).all()
    # a broken network: every system falls back, with the plain method's statuses bit for bit
    bad = gs.LearnedStart(device, seed=0)
    bad.norm = dict(net.norm)
    bad.set_parameters(np.full(M.NPARAM, np.nan, dtype=np.float32))
    Xb = xt.clone()
    Sb = gs.status_buffer(len(xt), device)
    nb = gs.solve_batch_learned(gs, Xb, Sb, prob, gs.Newton(), p, bad, opts)
    assert nb == len(xt)
    assert np.array_equal(Sb.cpu().numpy(), Sp.cpu().numpy())
    assert torch.equal(Xb, Xp)


def test_fused_solve_matches_the_three_step_path(gs, workloads, device):
    """The fused kernel (start, solve from it, fallback in one launch) gives the same statuses and
    iterates, bit for bit, as the start kernel followed by the batched solve and the fallback solve —
    for the plain and the mixed method, with a trained network and with a NaN network (every system
    falls back)."""
    Xk, X1 = _pairs(gs, workloads, device, 1 << 15, seed=41)
    net = gs.LearnedStart(device, seed=0)
    net.fit_normalization(Xk, X1, H)
    net.train(Xk, X1, H, epochs=4, batch=2048, lr=3e-3, lr_final=3e-5, seed=0)
    xt = torch.as_tensor(workloads.henon_heiles_states(20000 + 3, seed=42), device=device)
    prob = gs.NonlinearProblem(gs.HenonHeilesMidpoint())
    p = gs.BatchParams(H, xt)
    opts = gs.Options(**TOL)
    mixed = gs.Newton(linsolve=gs.LUFactorization(np.float32), refinement=gs.Refinement())
    bad = gs.LearnedStart(device, seed=0)
    bad.norm = dict(net.norm)
    bad.set_parameters(np.full(M.NPARAM, np.nan, dtype=np.float32))
    for n in (net, bad):
        for method in (gs.Newton(), mixed):
            Xa, Sa = torch.empty_like(xt), gs.status_buffer(len(xt), device)
            Xb, Sb = torch.empty_like(xt), gs.status_buffer(len(xt), device)
            na = gs.solve_batch_learned(gs, Xa, Sa, prob, method, p, n, opts, fused=True)
            nb = gs.solve_batch_learned(gs, Xb, Sb, prob, method, p, n, opts, fused=False)
            assert na == nb
            assert torch.equal(Sa, Sb)
            assert torch.equal(Xa.view(torch.int64), Xb.view(torch.int64))
        assert (gs.statuses(Sa)["code"] == 0).all()
