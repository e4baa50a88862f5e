"""The C ABI's own checks on the device, called through ctypes with no host-side validation in front:
an `order` that is not a permutation of 1:nsys is refused with GS_EINVAL before any solve writes
(the reference's isperm, src/nonlinear/batched.jl:133-135), and a workgroup size up to 1024 runs as
the kernel's own launch shape with the results of the default (batched.jl:200-210)."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

GS_EINVAL = -1


def _abi(gs):
    from importlib import import_module

    return import_module(gs.__name__ + "._abi")


def _setup(gs, device, nsys=4096, seed=5):
    rng = np.random.default_rng(seed)
    xk = rng.uniform(-0.3, 0.3, size=(nsys, 4))
    Xk = torch.as_tensor(xk, device=device)
    fam = gs.HenonHeilesMidpoint()
    P, p_stride, shared = fam.marshal(gs.BatchParams(0.1, Xk), nsys, device, torch.float64)
    cf = _abi(gs).GsFamily()
    cf.id, cf.n, cf.m, cf.p_len = fam.id, fam.n, fam.n, fam.p_len
    for k, v in enumerate(shared):
        cf.shared[k] = v
    return xk, Xk, P, p_stride, cf


def _call(gs, entry, cm, cf, X, P, p_stride, S, order, opts, ws=0):
    lib = gs.load_library()
    fn = getattr(lib, entry)
    rc = fn(ctypes.byref(cm), ctypes.byref(cf), X.shape[0], ctypes.c_void_p(X.data_ptr()),
            ctypes.c_void_p(P.data_ptr()), p_stride, ctypes.c_void_p(S.data_ptr()), None,
            ctypes.c_void_p(order.data_ptr()) if order is not None else None, ctypes.byref(opts), ws,
            gs._stream_ptr(X))
    return rc, lib.gs_last_error().decode()


def _bad_orders(n, rng):
    perm = rng.permutation(n) + 1
    out = {}
    o = perm.copy(); o[17] = 0; out["zero"] = o
    o = perm.copy(); o[n - 1] = n + 1; out["above"] = o
    o = perm.copy(); o[3] = -5; out["negative"] = o
    o = perm.copy(); o[100] = o[200]; out["repeat"] = o
    o = perm.copy(); o[0] = np.iinfo(np.int64).min; out["int64_min"] = o
    return out


@pytest.mark.parametrize("entry", ["gs_solve_batch_f64"])
def test_bad_order_is_refused_before_any_write(gs, device, entry):
    xk, Xk, P, p_stride, cf = _setup(gs, device)
    n = xk.shape[0]
    cm, opts = gs.Newton().to_c(), gs.Options(f_abstol=1e-12, f_reltol=0.0).to_c()
    rng = np.random.default_rng(0)
    for name, o in _bad_orders(n, rng).items():
        X = Xk.clone()
        S = torch.full((n, 72), 0xAB, dtype=torch.uint8, device=device)
        order = torch.as_tensor(o, dtype=torch.int64, device=device)
        print(f"{entry}: bad order '{name}'", flush=True)
        rc, msg = _call(gs, entry, cm, cf, X, P, p_stride, S, order, opts)
        assert rc == GS_EINVAL, name
        assert msg == f"order is not a permutation of 1:{n}", msg
        assert torch.equal(X, Xk), f"{name}: X was written"
        assert bool((S == 0xAB).all()), f"{name}: S was written"


def test_valid_order_and_workgroup_sizes_agree(gs, device):
    xk, Xk, P, p_stride, cf = _setup(gs, device)
    n = xk.shape[0]
    cm, opts = gs.Newton().to_c(), gs.Options(f_abstol=1e-12, f_reltol=0.0).to_c()
    X0 = Xk.clone()
    S0 = torch.zeros((n, 72), dtype=torch.uint8, device=device)
    rc, msg = _call(gs, "gs_solve_batch_f64", cm, cf, X0, P, p_stride, S0, None, opts)
    assert rc == 0, msg
    order = torch.as_tensor(np.random.default_rng(1).permutation(n) + 1, dtype=torch.int64, device=device)
    for ws, o in ((0, order), (257, None), (512, order), (1024, None), (1000, order)):
        X = Xk.clone()
        S = torch.zeros((n, 72), dtype=torch.uint8, device=device)
        print(f"workgroupsize {ws}, order {o is not None}", flush=True)
        rc, msg = _call(gs, "gs_solve_batch_f64", cm, cf, X, P, p_stride, S, o, opts, ws)
        assert rc == 0, msg
        assert torch.equal(X, X0) and torch.equal(S, S0)


def test_bad_order_refused_by_every_batch_form(gs, device):
    """The Float32 twin, the Deflated, StationarityProblem and Reparametrized forms run the same check
    (through the Python mirror with its host check off: the raw order reaches the library)."""
    n = 1024
    rng = np.random.default_rng(2)
    bad = _bad_orders(n, rng)["repeat"]
    lib = gs.load_library()
    xk = rng.uniform(-0.3, 0.3, size=(n, 4))
    fam = gs.HenonHeilesMidpoint()
    cm, opts = gs.Newton().to_c(), gs.Options(f_abstol=1e-6, f_reltol=0.0).to_c()
    order = torch.as_tensor(bad, dtype=torch.int64, device=device)
    # Float32
    X32 = torch.as_tensor(xk, dtype=torch.float32, device=device)
    P32, st32, sh = fam.marshal(gs.BatchParams(0.1, X32.clone()), n, device, torch.float32)
    cf = _abi(gs).GsFamily()
    cf.id, cf.n, cf.m, cf.p_len = fam.id, fam.n, fam.n, fam.p_len
    cf.shared[0] = sh[0]
    S32 = torch.full((n, 44), 0xAB, dtype=torch.uint8, device=device)
    X = X32.clone()
    rc = lib.gs_solve_batch_f32(ctypes.byref(cm), ctypes.byref(cf), n, ctypes.c_void_p(X.data_ptr()),
                                ctypes.c_void_p(P32.data_ptr()), st32, ctypes.c_void_p(S32.data_ptr()), None,
                                ctypes.c_void_p(order.data_ptr()), ctypes.byref(opts), 0, gs._stream_ptr(X))
    assert rc == GS_EINVAL and torch.equal(X, X32) and bool((S32 == 0xAB).all())
    # Deflated (one root)
    X64 = torch.as_tensor(xk, device=device)
    P, ps, sh = fam.marshal(gs.BatchParams(0.1, X64.clone()), n, device, torch.float64)
    cf.shared[0] = sh[0]
    d = _abi(gs).GsDeflation()
    d.nroots, d.power, d.shift = 1, 2.0, 1.0
    for k in range(4):
        d.roots[k] = 5.0
    S = torch.full((n, 72), 0xAB, dtype=torch.uint8, device=device)
    X = X64.clone()
    rc = lib.gs_solve_batch_deflated_f64(ctypes.byref(cm), ctypes.byref(cf), ctypes.byref(d), n,
                                         ctypes.c_void_p(X.data_ptr()), ctypes.c_void_p(P.data_ptr()), ps,
                                         ctypes.c_void_p(S.data_ptr()), None, ctypes.c_void_p(order.data_ptr()),
                                         ctypes.byref(opts), 0, gs._stream_ptr(X))
    assert rc == GS_EINVAL and torch.equal(X, X64) and bool((S == 0xAB).all())
