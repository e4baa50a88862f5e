"""The learned start of the Hénon–Heiles implicit-midpoint batch, with no counterpart in the reference.

A multilayer perceptron (three hidden ReLU layers of 128 units, BF16 operands with FP32 accumulation
on MFMA) predicts the start x̂₊ = x̃ + h³ g(x_k, h) of the Newton solve of x₊ = x_k + h f((x_k + x₊)/2),
where x̃ = x_k + h f(x_k + (h/2) f(x_k)) is the explicit midpoint step; it is trained on the GPU, by
the library's own forward, backward and Adam kernels, on solutions that the plain method itself
generates. `solve_batch_learned` solves
from the learned start and re-solves every system that does not converge from there with plain
Newton from the default start x₀ = x_k: the plain method is both the fallback and, through its FP64
residual test, the judge of convergence.

    net = LearnedStart(device)
    Xk = ...; X1 = generate_solutions(Xk, h)        # plain Newton on the GPU
    net.fit_normalization(Xk, X1, h); net.train(Xk, X1, h)
    S = solve_batch_learned(X, S, prob, method, BatchParams(h, Xk), net, options)
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _abi


def _torch():
    import torch

    return torch


def _lib():
    return _abi.load_library()


# The package's stream and error helpers, bound at first use (the package imports this module).
_HELPERS = []


def _helpers():
    if not _HELPERS:
        from . import _raise_on as r, _stream_ptr as sp

        _HELPERS.extend((sp, r))
    return _HELPERS


def _stream(t):
    return _helpers()[0](t)


def _raise_on(rc, lib):
    _helpers()[1](rc, lib)


IN, HID, OUT = _abi.GS_MLP_IN, _abi.GS_MLP_HIDDEN, _abi.GS_MLP_OUT
NPARAM, NPACKED = _abi.GS_MLP_NPARAM, _abi.GS_MLP_NPACKED
SHAPES = [("W1", (IN, HID)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)), ("W3", (HID, HID)),
          ("b3", (HID,)), ("W4", (HID, OUT)), ("b4", (OUT,))]


def initial_parameters(seed: int = 0) -> np.ndarray:
    """He-normal hidden weights (fan-in 5 for the first layer, whose other inputs are padding), a small
    output layer, zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in SHAPES:
        if name == "W1":
            w = np.zeros(shape)
            w[:5] = rng.standard_normal((5, HID)) * math.sqrt(2.0 / 5.0)
        elif name in ("W2", "W3"):
            w = rng.standard_normal(shape) * math.sqrt(2.0 / HID)
        elif name == "W4":
            w = np.zeros(shape)
            w[:, :4] = rng.standard_normal((HID, 4)) * math.sqrt(1.0 / HID) * 0.1
        else:
            w = np.zeros(shape)
        parts.append(w.reshape(-1))
    return np.concatenate(parts).astype(np.float32)


class LearnedStart:
    """The network, its Adam state and its normalisation, on one device."""

    def __init__(self, device, seed: int = 0):
        torch = _torch()
        self.device = device
        self.params = torch.as_tensor(initial_parameters(seed), device=device).contiguous()
        self.packed = torch.zeros(NPACKED, dtype=torch.int16, device=device)
        self.m = torch.zeros(NPARAM, dtype=torch.float32, device=device)
        self.v = torch.zeros(NPARAM, dtype=torch.float32, device=device)
        self.grad_buf = torch.zeros(NPARAM, dtype=torch.float32, device=device)
        self.loss_buf = torch.zeros(1, dtype=torch.float32, device=device)
        self.step = 0
        self.norm = {"in_shift": [0.0] * 8, "in_scale": [1.0] * 8, "out_shift": [0.0] * 4, "out_scale": [1.0] * 4}
        # the C records of the last call: the network's (keyed by its pointers and normalisation) and
        # the method's and options' (keyed by the identity of those frozen values)
        self._c_key, self._c = None, None
        self._records = (None, None, None, None)
        self.pack()

    def pack(self):
        lib = _lib()
        _raise_on(lib.gs_mlp_pack(ctypes.c_void_p(self.params.data_ptr()), ctypes.c_void_p(self.packed.data_ptr()),
                                  _stream(self.params)), lib)

    def set_parameters(self, P):
        torch = _torch()
        self.params.copy_(torch.as_tensor(np.asarray(P, dtype=np.float32)))
        self.pack()

    def to_c(self) -> _abi.GsMlp:
        """The gs_mlp record of the network: built again only when its tensors or its normalisation
        changed since the last call (the library reads it and keeps no reference)."""
        nm = self.norm
        key = (self.params.data_ptr(), self.packed.data_ptr(), tuple(nm["in_shift"]), tuple(nm["in_scale"]),
               tuple(nm["out_shift"]), tuple(nm["out_scale"]))
        if key != self._c_key:
            c = _abi.GsMlp()
            c.params = self.params.data_ptr()
            c.packed = self.packed.data_ptr()
            for j in range(8):
                c.in_shift[j] = nm["in_shift"][j]
                c.in_scale[j] = nm["in_scale"][j]
            for j in range(4):
                c.out_shift[j] = nm["out_shift"][j]
                c.out_scale[j] = nm["out_scale"][j]
            self._c_key, self._c = key, c
        return self._c

    def method_records(self, method, opts):
        """The gs_method and gs_options_f64 records of a fused call, reused while the same (frozen)
        method and options are passed."""
        m0, o0, cm, co = self._records
        if method is not m0 or opts is not o0:
            cm, co = method.to_c(), opts.to_c()
            self._records = (method, opts, cm, co)
        return cm, co

    def fit_normalization(self, Xk, X1, h: float):
        """Feature and target shifts and scales from training pairs (means and standard deviations)."""
        xk = Xk.detach().cpu().numpy()
        x1 = X1.detach().cpu().numpy()

        def field(x):
            q1, q2, p1, p2 = x.T
            return np.stack([p1, p2, -q1 - 2.0 * q1 * q2, (-q2 - q1 * q1) + q2 * q2], axis=1)

        xt = xk + h * field(xk + (0.5 * h) * field(xk))
        corr = (x1 - xt) / ((h * h) * h)
        mu, sd = xk.mean(axis=0), xk.std(axis=0)
        cm, cs = corr.mean(axis=0), corr.std(axis=0)
        self.norm["in_shift"] = [float(v) for v in mu] + [float(h), 0.0, 0.0, 0.0]
        self.norm["in_scale"] = [float(1.0 / v) if v > 0 else 1.0 for v in sd] + [1.0, 0.0, 0.0, 0.0]
        self.norm["out_shift"] = [float(v) for v in cm]
        self.norm["out_scale"] = [float(v) if v > 0 else 1.0 for v in cs]

    def grad(self, Xk, X1, h: float, idx):
        """(loss, gradient) of the minibatch idx (a device int64 tensor, length a multiple of 32)."""
        lib = _lib()
        net = self.to_c()
        rc = lib.gs_mlp_grad(ctypes.byref(net), ctypes.c_void_p(Xk.data_ptr()), ctypes.c_void_p(X1.data_ptr()), float(h),
                             ctypes.c_void_p(idx.data_ptr()), int(len(idx)), ctypes.c_void_p(self.grad_buf.data_ptr()),
                             ctypes.c_void_p(self.loss_buf.data_ptr()), _stream(Xk))
        _raise_on(rc, lib)
        return self.loss_buf, self.grad_buf

    def adam(self, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8):
        lib = _lib()
        self.step += 1
        rc = lib.gs_mlp_adam(ctypes.c_void_p(self.params.data_ptr()), ctypes.c_void_p(self.packed.data_ptr()),
                             ctypes.c_void_p(self.grad_buf.data_ptr()), ctypes.c_void_p(self.m.data_ptr()),
                             ctypes.c_void_p(self.v.data_ptr()), self.step, lr, beta1, beta2, eps, _stream(self.params))
        _raise_on(rc, lib)

    def train(self, Xk, X1, h: float, epochs: int = 10, batch: int = 4096, lr: float = 2e-3, lr_final: float = 2e-5,
              seed: int = 0, log_every: int = 0):
        """Minibatch Adam with a cosine learning-rate schedule over shuffled epochs; returns the mean
        loss of each epoch (read back once per epoch)."""
        torch = _torch()
        n = Xk.shape[0]
        steps = n // batch
        if steps < 1:
            raise ValueError("fewer training pairs than one batch")
        gen = torch.Generator(device=Xk.device)
        gen.manual_seed(seed)
        total = epochs * steps
        history = []
        acc = torch.zeros(1, dtype=torch.float32, device=Xk.device)
        t = 0
        for _ in range(epochs):
            perm = torch.randperm(n, generator=gen, device=Xk.device)
            acc.zero_()
            for s in range(steps):
                idx = perm[s * batch:(s + 1) * batch]
                loss, _ = self.grad(Xk, X1, h, idx)
                acc += loss
                rate = lr_final + 0.5 * (lr - lr_final) * (1.0 + math.cos(math.pi * t / max(total - 1, 1)))
                self.adam(rate)
                t += 1
            history.append(float(acc.item()) / steps)
        return history

    def start(self, Xk, h: float, out=None):
        """x̂₊ of every row of Xk ((nsys, 4) float64 on the device) into out (or a new tensor)."""
        torch = _torch()
        if out is None:
            out = torch.empty_like(Xk)
        lib = _lib()
        net = self.to_c()
        rc = lib.gs_mlp_start_henon_heiles(ctypes.byref(net), Xk.shape[0], ctypes.c_void_p(Xk.data_ptr()), float(h),
                                           ctypes.c_void_p(out.data_ptr()), _stream(Xk))
        _raise_on(rc, lib)
        return out

    def state(self) -> dict:
        return {"params": self.params.detach().cpu().numpy(), "norm": {k: list(v) for k, v in self.norm.items()}}

    def load_state(self, st: dict):
        self.norm = {k: list(v) for k, v in st["norm"].items()}
        self.set_parameters(st["params"])


def generate_solutions(gs, Xk, h: float, options=None):
    """The plain method's solutions x₊ of the midpoint steps from every row of Xk (device), the
    training targets of the network."""
    X = Xk.clone()
    S = gs.status_buffer(Xk.shape[0], Xk.device)
    opts = options if options is not None else gs.Options(f_abstol=1e-12, f_reltol=0.0)
    gs.solve_batch(X, S, gs.NonlinearProblem(gs.HenonHeilesMidpoint()), gs.Newton(), gs.BatchParams(h, Xk), opts)
    return X, S


def solve_batch_learned(gs, X, S, prob, method, p, net: LearnedStart, options=None, fallback=None, fused: bool = True):
    """solve_batch from the learned start: X is overwritten with the solutions, S with the statuses.
    Every system whose solve from the learned start does not return SUCCESS is solved again from the
    default start x₀ = x_k with `fallback` (plain Newton by default), and its status is that solve's.
    Returns the number of systems that fell back.

    With `fused` (the default, for plain Newton as the fallback) one kernel does all three parts per
    system — the start, the solve from it and the fallback — and leaves the start in registers;
    otherwise the start kernel, the batched solve and the fallback solve run one after the other."""
    torch = _torch()
    if not isinstance(prob.F, gs.HenonHeilesMidpoint) or not isinstance(p, gs.BatchParams):
        raise gs.ArgumentError("the learned start is trained for HenonHeilesMidpoint with BatchParams(h, X_k)")
    h = float(p.shared)
    Xk = p.per_system
    opts = gs.resolve_options(options, {})
    if fused and fallback is None and isinstance(method.globalization, gs.Static):
        lib = _lib()
        cm, co = net.method_records(method, opts)
        net_c = net.to_c()
        nb = ctypes.c_int64(0)
        rc = lib.gs_solve_learned_henon_heiles_f64(
            ctypes.byref(net_c), ctypes.byref(cm), Xk.shape[0], ctypes.c_void_p(Xk.data_ptr()), h,
            ctypes.c_void_p(X.data_ptr()), ctypes.c_void_p(S.data_ptr()), ctypes.byref(co), ctypes.byref(nb), _stream(X))
        _raise_on(rc, lib)
        return int(nb.value)
    net.start(Xk, h, out=X)
    gs.solve_batch(X, S, prob, method, p, opts)
    bad = torch.nonzero(S[:, 0] != 0).flatten()
    nb = int(bad.numel())
    if nb:
        fb = fallback if fallback is not None else gs.Newton()
        Xb = Xk.index_select(0, bad).contiguous()
        Sb = gs.status_buffer(nb, X.device)
        gs.solve_batch(Xb, Sb, prob, fb, gs.BatchParams(h, Xb.clone()), opts)
        X.index_copy_(0, bad, Xb)
        S.index_copy_(0, bad, Sb)
    return nb
