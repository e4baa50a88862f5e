"""Pins the CPU oracle of the learned-start network (oracle/mlp_ref.py): its bfloat16 rounding against
closed forms, and its back-propagated gradient against central differences of its own loss in FP64."""
import numpy as np

from oracle import mlp_ref as M


def test_bf16_rounding_closed_forms():
    one = 1.0
    u = 2.0 ** -7  # bfloat16 ulp of 1
    x = np.array([one, one + u / 2, one + 3 * u / 2, one + u / 4, -one - 3 * u / 4, 2.0 ** -130, np.inf, 0.0])
    want = np.array([one, one, one + 2 * u, one, -one - u, 2.0 ** -130, np.inf, 0.0])
    assert np.array_equal(M.bf16(x), want)
    assert np.isnan(M.bf16(np.nan))
    # every bfloat16 value is a fixed point
    r = np.random.default_rng(0).standard_normal(1000)
    assert np.array_equal(M.bf16(M.bf16(r)), M.bf16(r))


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(3)
    P = rng.standard_normal(M.NPARAM) * 0.2
    U = np.zeros((64, M.IN))
    U[:, :5] = rng.standard_normal((64, 5))
    T = np.zeros((64, M.OUT))
    T[:, :4] = rng.standard_normal((64, 4))
    loss, g = M.loss_and_grad(P, U, T, emulate=False)
    # the loss is ½ mean of squares of the used outputs
    _, _, z4 = M.forward(P, U, emulate=False)
    assert np.isclose(loss, 0.5 * ((z4[:, :4] - T[:, :4]) ** 2).mean())
    idx = np.concatenate([rng.choice(np.arange(0, 5 * M.HID), 6),           # W1 rows that see inputs
                          rng.choice(np.arange(M.IN * M.HID, M.NPARAM), 30)])
    eps = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += eps
        Pm[i] -= eps
        fd = (M.loss_and_grad(Pp, U, T, emulate=False)[0] - M.loss_and_grad(Pm, U, T, emulate=False)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])
    # the padded inputs (zero features) receive no gradient
    W1g = g[:M.IN * M.HID].reshape(M.IN, M.HID)
    assert np.all(W1g[5:] == 0.0)
