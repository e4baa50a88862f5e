"""CPU ORACLE — TEST INFRASTRUCTURE ONLY.

FP64 numpy restatement of the learned-start network of libgsolve_hip (csrc/mlp.hip): a multilayer
perceptron with inputs padded to 32, three hidden ReLU layers of 128 units and outputs padded to 16,
predicting the correction of the explicit midpoint step x̃ = x + h f(x + (h/2) f(x)) towards the
Hénon–Heiles implicit-midpoint step,

    x̂₊ = x̃ + h³ (o_j σ_j + μ_j),  o = net(u),  u_j = (x_j − s_j) c_j (j < 4), u_4 = (h − s_4) c_4.

The network has no counterpart in the reference. With `emulate=True` the forward and backward
passes round to bfloat16 (round to nearest even) exactly where the device feeds its MFMA operands:
the features, the weights, the activations and the back-propagated gradients; everything else is
FP64 (the device accumulates in FP32). With `emulate=False` it is the plain FP64 network, whose
gradient tests/test_mlp_ref.py pins by central differences.
"""
from __future__ import annotations

import numpy as np

IN, HID, OUT = 32, 128, 16
SHAPES = [("W1", (IN, HID)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)), ("W3", (HID, HID)),
          ("b3", (HID,)), ("W4", (HID, OUT)), ("b4", (OUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def bf16(a) -> np.ndarray:
    """Round to bfloat16 (nearest, ties to even) through float32, returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    out = np.where(np.isnan(f), f, out)
    return out.astype(np.float64)


def unpack(P: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(P[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    return out


def pack(d: dict) -> np.ndarray:
    return np.concatenate([np.asarray(d[name], dtype=np.float64).reshape(-1) for name, _ in SHAPES])


def hh_field(x: np.ndarray) -> np.ndarray:
    q1, q2, p1, p2 = x[:, 0], x[:, 1], x[:, 2], x[:, 3]
    return np.stack([p1, p2, -q1 - 2.0 * q1 * q2, (-q2 - q1 * q1) + q2 * q2], axis=1)


def rk2(xk: np.ndarray, h: float) -> np.ndarray:
    """The explicit midpoint step x + h f(x + (h/2) f(x)), in the device's operation order."""
    m = xk + (0.5 * h) * hh_field(xk)
    return xk + h * hh_field(m)


def features(xk: np.ndarray, h: float, norm: dict) -> np.ndarray:
    """The device's features: FP32 arithmetic, then bfloat16."""
    B = len(xk)
    U = np.zeros((B, IN))
    sh = np.asarray(norm["in_shift"], np.float32)
    sc = np.asarray(norm["in_scale"], np.float32)
    for j in range(4):
        U[:, j] = ((xk[:, j].astype(np.float32) - sh[j]) * sc[j]).astype(np.float64)
    U[:, 4] = float((np.float32(h) - sh[4]) * sc[4])
    return bf16(U)


def targets(xk: np.ndarray, x1: np.ndarray, h: float, norm: dict) -> np.ndarray:
    B = len(xk)
    T = np.zeros((B, OUT))
    corr = (x1 - rk2(xk, h)) / ((h * h) * h)
    sh = np.asarray(norm["out_shift"], np.float32)
    sc = np.asarray(norm["out_scale"], np.float32)
    for j in range(4):
        T[:, j] = ((corr[:, j].astype(np.float32) - sh[j]) / sc[j]).astype(np.float64)
    return T


def forward(P: np.ndarray, U: np.ndarray, emulate: bool = True):
    d = unpack(P)
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    Z, A = [], []
    a = U
    for l in (1, 2, 3):
        z = a @ r(d[f"W{l}"]) + d[f"b{l}"]
        Z.append(z)
        a = r(np.maximum(z, 0.0))
        A.append(a)
    z4 = a @ r(d["W4"]) + d["b4"]
    return Z, A, z4


def loss_and_grad(P: np.ndarray, U: np.ndarray, T: np.ndarray, emulate: bool = True):
    """½·mean over the batch and the 4 used outputs of (o − t)², and its gradient (flat, as P)."""
    d = unpack(P)
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    B = len(U)
    Z, A, z4 = forward(P, U, emulate)
    diff = np.zeros_like(z4)
    diff[:, :4] = z4[:, :4] - T[:, :4]
    loss = 0.5 * float((diff ** 2).sum()) / (4 * B)
    dz = diff / (4 * B)
    g = {}
    g["W4"] = A[2].T @ r(dz)
    g["b4"] = dz.sum(axis=0)
    da = r(dz) @ r(d["W4"]).T
    for l in (3, 2, 1):
        dz = np.where(Z[l - 1] > 0.0, da, 0.0)
        prev = U if l == 1 else A[l - 2]
        g[f"W{l}"] = prev.T @ r(dz)
        g[f"b{l}"] = dz.sum(axis=0)
        if l > 1:
            da = r(dz) @ r(d[f"W{l}"]).T
    return loss, pack(g)


def predict(P: np.ndarray, norm: dict, xk: np.ndarray, h: float, emulate: bool = True) -> np.ndarray:
    """x̂₊ of every row of xk."""
    U = features(xk, h, norm)
    _, _, z4 = forward(P, U, emulate)
    o = z4[:, :4].astype(np.float32)
    corr = (o * np.asarray(norm["out_scale"], np.float32) + np.asarray(norm["out_shift"], np.float32)).astype(np.float64)
    return rk2(xk, h) + ((h * h) * h) * corr
