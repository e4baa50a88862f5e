"""The learned start at the headline configuration's full size: the network trained on 10^6
method-generated pairs (seed 100, 8 epochs), the fused solve of 2^20 Hénon–Heiles midpoint steps
(seed 0) with Newton(LUFactorization(Float32), Refinement), every system rechecked in FP64, and a
sample of 4096 systems compared bit for bit with the oracle's solve: the oracle runs the same mixed
method from the learned starts (the start kernel's output, which the fused kernel reproduces bit for
bit), and plain Newton from x_k for any system that fell back."""
import numpy as np
import pytest
import torch

from test_gpu_parity import assert_parity

pytestmark = pytest.mark.gpu

H = 0.1
NTRAIN = 1_000_000
NSYS = 1 << 20


def test_full_size_fused_solve_against_the_oracle(gs, orc, workloads, device):
    opts = gs.Options(f_abstol=1e-12, f_reltol=0.0)
    Xt = torch.as_tensor(workloads.henon_heiles_states(NTRAIN, seed=100), device=device)
    X1, S1 = gs.generate_solutions(gs, Xt, H, opts)
    assert (gs.statuses(S1)["code"] == 0).all()
    net = gs.LearnedStart(device, seed=0)
    net.fit_normalization(Xt, X1, H)
    hist = net.train(Xt, X1, H, epochs=8, batch=4096, lr=3e-3, lr_final=3e-5, seed=0)
    print("loss by epoch", [f"{v:.3e}" for v in hist], flush=True)
    del Xt, X1

    xk = workloads.henon_heiles_states(NSYS, seed=0)
    Xk = torch.as_tensor(xk, device=device)
    prob = gs.NonlinearProblem(gs.HenonHeilesMidpoint())
    mixed = gs.Newton(linsolve=gs.LUFactorization(np.float32), refinement=gs.Refinement())
    X = torch.empty_like(Xk)
    S = gs.status_buffer(NSYS, device)
    nb = gs.solve_batch_learned(gs, X, S, prob, mixed, gs.BatchParams(H, Xk), net, opts)
    st = gs.statuses(S)
    x = X.cpu().numpy()
    print("fallbacks", nb, "codes", np.bincount(st["code"], minlength=6).tolist(),
          "mean iterations", st["iterations"].mean(), flush=True)
    # every system converges, and its FP64 residual meets the tolerance
    assert (st["code"] == 0).all()
    r = np.sqrt((workloads.henon_heiles_midpoint_residual(x, xk, H) ** 2).sum(axis=1))
    assert (r <= 1e-12).all()
    assert nb < NSYS // 1000

    # a sample against the oracle, bit for bit
    starts = net.start(Xk, H).cpu().numpy()
    idx = np.sort(np.random.default_rng(5).choice(NSYS, 4096, replace=False))
    xo, so = orc.solve_batch(mixed, gs.HenonHeilesMidpoint(), starts[idx].copy(), gs.BatchParams(H, xk[idx]), opts)
    # a system whose solve from the learned start did not succeed was solved again from x_k
    fell = so["code"] != 0
    if fell.any():
        xf, sf = orc.solve_batch(gs.Newton(), gs.HenonHeilesMidpoint(), xk[idx][fell].copy(),
                                 gs.BatchParams(H, xk[idx][fell]), opts)
        xo[fell] = xf
        so[fell] = sf
    # the sample's statuses as raw records (a structured array's fancy index copies fields, not padding)
    st_idx = np.ascontiguousarray(S.cpu().numpy()[idx]).reshape(-1).view(st.dtype)
    assert_parity(st_idx, x[idx], so, xo, bitwise=True)
