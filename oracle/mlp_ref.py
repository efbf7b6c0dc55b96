"""numpy reference of the warm-start MLP (TEST INFRASTRUCTURE ONLY).

The checker for the device MLP kernels: the same 39-128-128-128-4 network
with exact-erf GELU, evaluated in FP64 from the same BF16-rounded operands
the device feeds its BF16 x BF16 -> FP32 products (weights, layer inputs,
hidden activations and back-propagated gradients rounded to BF16 at the
same points).  The network has no counterpart in mind-fcl, so there is no
reference file to cite; its forward and backward formulas are pinned by
central finite differences of the unrounded FP64 variant
(tests/test_mlp_reference.py).  Product code never imports this module.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

IN, IN_PAD, HID, OUT, OUT_PAD = 39, 64, 128, 4, 32
OFF_W1 = 0
OFF_B1 = OFF_W1 + HID * IN_PAD
OFF_W2 = OFF_B1 + HID
OFF_B2 = OFF_W2 + HID * HID
OFF_W3 = OFF_B2 + HID
OFF_B3 = OFF_W3 + HID * HID
OFF_W4 = OFF_B3 + HID
OFF_B4 = OFF_W4 + OUT_PAD * HID
N_PARAMS = OFF_B4 + OUT_PAD


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (round to nearest, ties to even) and return FP64."""
    f = np.ascontiguousarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def unpack(p: np.ndarray):
    p = np.asarray(p, np.float64)
    W1 = p[OFF_W1:OFF_B1].reshape(HID, IN_PAD)
    b1 = p[OFF_B1:OFF_W2]
    W2 = p[OFF_W2:OFF_B2].reshape(HID, HID)
    b2 = p[OFF_B2:OFF_W3]
    W3 = p[OFF_W3:OFF_B3].reshape(HID, HID)
    b3 = p[OFF_B3:OFF_W4]
    W4 = p[OFF_W4:OFF_B4].reshape(OUT_PAD, HID)
    b4 = p[OFF_B4:N_PARAMS]
    return [(W1, b1), (W2, b2), (W3, b3), (W4, b4)]


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward(p, X, rounded=True):
    """X: (n, 64) features.  Returns outputs (n, 32) and the saved tensors."""
    R = bf16 if rounded else (lambda a: np.asarray(a, np.float64))
    layers = unpack(p)
    h = R(X)
    saved = {"h0": h}
    for j, (W, b) in enumerate(layers[:3], start=1):
        z = h @ R(W).T + b
        saved[f"z{j}"] = z
        h = R(gelu(z))
        saved[f"h{j}"] = h
    W4, b4 = layers[3]
    y = h @ R(W4).T + b4
    return y, saved


def loss_grad(y, T, wd=1.0):
    """Mean of (1 - cos(y[:3], t[:3])) + wd (y3 - t3)^2 over rows with a
    non-zero target direction; returns (loss, dL/dy (n, 32))."""
    n = len(y)
    g = np.zeros_like(y)
    t = T[:, :3]
    tn = np.linalg.norm(t, axis=1)
    ok = tn > 0
    yn = np.maximum(np.linalg.norm(y[:, :3], axis=1), 1e-12)
    c = np.where(ok, (y[:, :3] * t).sum(1) / (yn * np.where(ok, tn, 1)), 0.0)
    e = y[:, 3] - T[:, 3]
    loss = float(np.sum(np.where(ok, 1 - c + wd * e * e, 0.0)) / n)
    gd = -(t / np.where(ok, tn, 1)[:, None] - c[:, None] * y[:, :3] / yn[:, None]) / yn[:, None] / n
    g[:, :3] = np.where(ok[:, None], gd, 0.0)
    g[:, 3] = np.where(ok, 2 * wd * e / n, 0.0)
    return loss, g


def backward(p, saved, dy, rounded=True):
    """Gradient of the loss w.r.t. the packed parameter vector."""
    R = bf16 if rounded else (lambda a: np.asarray(a, np.float64))
    layers = unpack(p)
    grad = np.zeros(N_PARAMS)
    offs = [(OFF_W1, OFF_B1, OFF_W2), (OFF_W2, OFF_B2, OFF_W3), (OFF_W3, OFF_B3, OFF_W4), (OFF_W4, OFF_B4, N_PARAMS)]
    dz = R(dy)
    for j in (4, 3, 2, 1):
        W, _ = layers[j - 1]
        h_in = saved[f"h{j - 1}"]
        ow, ob, oe = offs[j - 1]
        grad[ow:ob] = (dz.T @ h_in).reshape(-1)
        grad[ob:oe] = dz.sum(0)
        if j > 1:
            dh = dz @ R(W)
            dz = R(dh * gelu_grad(saved[f"z{j - 1}"]))
    return grad


def init_params(seed: int) -> np.ndarray:
    """Xavier-uniform init in numpy (tests only; the device has its own)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS)
    for (ow, ob), (rows, cols, fi, fo, vr, vc) in zip(
            [(OFF_W1, OFF_B1), (OFF_W2, OFF_B2), (OFF_W3, OFF_B3), (OFF_W4, OFF_B4)],
            [(HID, IN_PAD, IN, HID, HID, IN), (HID, HID, HID, HID, HID, HID), (HID, HID, HID, HID, HID, HID),
             (OUT_PAD, HID, HID, OUT, OUT, HID)]):
        a = np.sqrt(6.0 / (fi + fo))
        w = np.zeros((rows, cols))
        w[:vr, :vc] = rng.uniform(-a, a, (vr, vc))
        p[ow:ob] = w.reshape(-1)
    return p.astype(np.float32).astype(np.float64)
