"""CPU: pin the numpy reference of the warm-start MLP.

The network has no counterpart in the reference library, so its checker is
pinned by mathematics: the analytic backward pass equals central finite
differences of the loss (FP64, no BF16 rounding), the loss reaches its
minimum 0 at a perfect prediction, and BF16 rounding is round-to-nearest-
even on exactly representable edge cases.
"""
import numpy as np

import mlp_ref as M


def _data(rng, n):
    X = np.zeros((n, M.IN_PAD))
    X[:, :M.IN] = rng.normal(size=(n, M.IN))
    T = np.zeros((n, 4))
    d = rng.normal(size=(n, 3))
    T[:, :3] = d / np.linalg.norm(d, axis=1, keepdims=True)
    T[:, 3] = rng.normal(size=n)
    T[0, :3] = 0  # masked row
    return X, T


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = M.init_params(1)
    p[M.OFF_B1:M.OFF_W2] = rng.normal(scale=0.1, size=M.HID)
    p[M.OFF_B4:M.OFF_B4 + 4] = rng.normal(scale=0.1, size=4)
    X, T = _data(rng, 6)

    def loss(q):
        y, _ = M.forward(q, X, rounded=False)
        return M.loss_grad(y, T)[0]

    y, saved = M.forward(p, X, rounded=False)
    _, dy = M.loss_grad(y, T)
    g = M.backward(p, saved, dy, rounded=False)
    # every parameter block, sampled entries (padding entries must be zero)
    idx = np.concatenate([rng.choice(np.arange(lo, hi), 12, replace=False) for lo, hi in
                          [(M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2), (M.OFF_B2, M.OFF_W3),
                           (M.OFF_W3, M.OFF_B3), (M.OFF_B3, M.OFF_W4)]] +
                         [M.OFF_W4 + np.arange(0, 4 * M.HID, 37), M.OFF_B4 + np.arange(4)])
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (loss(p + e) - loss(p - e)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])
    # padding columns of W1 see zero inputs; padding rows of W4 see zero output gradients
    W1g = g[M.OFF_W1:M.OFF_B1].reshape(M.HID, M.IN_PAD)
    assert np.all(W1g[:, M.IN:] == 0)
    assert np.all(g[M.OFF_W4 + 4 * M.HID:M.OFF_B4] == 0) and np.all(g[M.OFF_B4 + 4:] == 0)


def test_loss_zero_at_perfect_prediction():
    rng = np.random.default_rng(1)
    _, T = _data(rng, 8)
    y = np.zeros((8, M.OUT_PAD))
    y[:, :4] = T * np.array([3.0, 3.0, 3.0, 1.0])  # any positive scale of the direction
    loss, g = M.loss_grad(y, T)
    assert abs(loss) < 1e-12
    assert np.abs(g).max() < 1e-12


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7 + 2 ** -8, 1.0 + 2 ** -9, -3.140625, 0.0], np.float32)
    r = M.bf16(x)
    # ties to even: 1 + 2^-8 -> 1; 1 + 2^-7 + 2^-8 -> 1 + 2^-6; below half -> down
    assert list(r) == [1.0, 1.0, 1.0 + 2 ** -6, 1.0, -3.140625, 0.0]
