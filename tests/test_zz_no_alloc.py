"""gs_solve_batch_f64_noalloc against gs_solve_batch_f64 through the C ABI: the same bits on two calls
in a row, the second smaller, for the two uses of its scratch array (the order check's bitmap, the
FpuGl3 solve's hand-over flags), and its limits refused before any device work."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

GS_EINVAL = -1


def _abi(gs):
    from importlib import import_module

    return import_module(gs.__name__ + "._abi")


def _family(gs, fam, shared):
    cf = _abi(gs).GsFamily()
    cf.id, cf.n, cf.m, cf.p_len = fam.id, fam.n, fam.n, fam.p_len
    for k, v in enumerate(shared):
        cf.shared[k] = v
    return cf


def _call(gs, entry, cm, cf, nsys, X, P, p_stride, S, order, opts):
    lib = gs.load_library()
    fn = getattr(lib, entry)
    rc = fn(ctypes.byref(cm), ctypes.byref(cf), nsys, ctypes.c_void_p(X.data_ptr()), ctypes.c_void_p(P.data_ptr()),
            p_stride, ctypes.c_void_p(S.data_ptr()), None,
            ctypes.c_void_p(order.data_ptr()) if order is not None else None, ctypes.byref(opts), 0,
            gs._stream_ptr(X))
    return rc, lib.gs_last_error().decode()


def _both(gs, cm, cf, X0, P, p_stride, order, opts):
    out = []
    for entry in ("gs_solve_batch_f64", "gs_solve_batch_f64_noalloc"):
        X = X0.clone()
        S = torch.zeros((X0.shape[0], 72), dtype=torch.uint8, device=X0.device)
        rc, msg = _call(gs, entry, cm, cf, X0.shape[0], X, P, p_stride, S, order, opts)
        assert rc == 0, f"{entry}: {msg}"
        out.append((X.cpu(), S.cpu()))
    return out


def test_noalloc_matches_the_allocating_entry_point(gs, workloads, device):
    opts = gs.Options(f_abstol=1e-12, f_reltol=0.0).to_c()
    # the lane-per-system solve with an order (the scratch array holds the order check), then a smaller batch
    fam = gs.HenonHeilesMidpoint()
    for n, seed in ((8192, 1), (1000, 2)):
        print(f"noalloc: HenonHeiles n={n} with order", flush=True)
        xk = workloads.henon_heiles_states(n, seed=seed)
        Xk = torch.as_tensor(xk, device=device)
        P, ps, sh = fam.marshal(gs.BatchParams(0.1, Xk), n, device, torch.float64)
        order = torch.as_tensor(np.random.default_rng(seed).permutation(n) + 1, dtype=torch.int64, device=device)
        for m in (gs.Newton(), gs.Newton(linsolve=gs.LUFactorization("float32"), refinement=gs.Refinement())):
            (xa, sa), (xb, sb) = _both(gs, m.to_c(), _family(gs, fam, sh), Xk, P, ps, order, opts)
            assert torch.equal(xa, xb) and torch.equal(sa, sb)
    # the FpuGl3 solve (the scratch array holds its hand-over flags), with and without order, then smaller
    fam = gs.FpuGl3()
    for n, seed in ((1024, 3), (300, 4)):
        print(f"noalloc: FpuGl3 n={n}", flush=True)
        QP = torch.as_tensor(workloads.fpu_states(n, seed=seed), device=device)
        Z0 = torch.zeros((n, 48), dtype=torch.float64, device=device)
        P, ps, sh = fam.marshal(gs.BatchParams(0.05, QP), n, device, torch.float64)
        order = torch.as_tensor(np.random.default_rng(seed).permutation(n) + 1, dtype=torch.int64, device=device)
        for m, o in ((gs.Newton(), None),
                     (gs.Newton(linsolve=gs.LUFactorization("float32"), refinement=gs.Refinement()), order)):
            (xa, sa), (xb, sb) = _both(gs, m.to_c(), _family(gs, fam, sh), Z0, P, ps, o, opts)
            assert torch.equal(xa, xb) and torch.equal(sa, sb)


def test_noalloc_refuses_what_its_array_cannot_hold(gs, workloads, device):
    """The limits are checked before any device work: nsys beyond them with a small buffer is refused
    without a launch (nothing reads past the buffers)."""
    opts = gs.Options().to_c()
    fam = gs.FpuGl3()
    QP = torch.as_tensor(workloads.fpu_states(4, seed=0), device=device)
    Z = torch.zeros((4, 48), dtype=torch.float64, device=device)
    P, ps, sh = fam.marshal(gs.BatchParams(0.05, QP), 4, device, torch.float64)
    S = torch.zeros((4, 72), dtype=torch.uint8, device=device)
    rc, msg = _call(gs, "gs_solve_batch_f64_noalloc", gs.Newton().to_c(), _family(gs, fam, sh), (1 << 20) + 1, Z, P,
                    0, S, None, opts)
    assert rc == GS_EINVAL and "at most 1048576 FpuGl3 systems" in msg
    fam = gs.HenonHeilesMidpoint()
    X = torch.zeros((4, 4), dtype=torch.float64, device=device)
    P, ps, sh = fam.marshal(gs.BatchParams(0.1, X), 4, device, torch.float64)
    order = torch.arange(1, 5, dtype=torch.int64, device=device)
    rc, msg = _call(gs, "gs_solve_batch_f64_noalloc", gs.Newton().to_c(), _family(gs, fam, sh), (1 << 24) + 1, X, P,
                    0, S, order, opts)
    assert rc == GS_EINVAL and "order of at most 16777216 systems" in msg
