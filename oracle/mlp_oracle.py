"""FP64 restatement of the step-proposal network, for the checks only.

TEST INFRASTRUCTURE ONLY (tests/, bench.py's learned-step checks).  No
counterpart in the reference: the network is new.  It restates the GPU
kernels' arithmetic in float64 with the BF16 roundings at the same points:
  X -> bf16;  W1, W2 -> bf16;  A1 = act(Z1) -> bf16 before layer 2;
  output layer in FP32/FP64 on unrounded act(Z2);
  backward: dZ2 -> bf16 for the dW2 and dA1 products, dZ1 -> bf16 for dW1,
  bias and w3 gradients from the unrounded values.
Parameter layout: W1[64][32] (inputs 24..31 zero), b1[64], W2[64][64], b2[64],
w3[64], b3.  act(z) = z * sigmoid(1.702 z).
"""
from __future__ import annotations

import numpy as np

NIN, K1, H = 24, 32, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + H * K1
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
NPARAM = OFF_B3 + 1


def bf16(x) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    return dict(W1=p[OFF_W1:OFF_B1].reshape(H, K1), b1=p[OFF_B1:OFF_W2], W2=p[OFF_W2:OFF_B2].reshape(H, H),
                b2=p[OFF_B2:OFF_W3], w3=p[OFF_W3:OFF_B3], b3=p[OFF_B3])


def act(z):
    return z / (1.0 + np.exp(-1.702 * z))


def act_grad(z):
    s = 1.0 / (1.0 + np.exp(-1.702 * z))
    return s + 1.702 * z * s * (1.0 - s)


def init(seed: int = 0) -> np.ndarray:
    """He-uniform initial weights, zero biases; the output bias starts at the
    label mean of a typical trace set (log2(1 + 8))."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)
    W1 = np.zeros((H, K1))
    W1[:, :NIN] = rng.uniform(-1, 1, (H, NIN)) * np.sqrt(6.0 / NIN)
    p[OFF_W1:OFF_B1] = W1.reshape(-1)
    p[OFF_W2:OFF_B2] = (rng.uniform(-1, 1, (H, H)) * np.sqrt(6.0 / H)).reshape(-1)
    p[OFF_W3:OFF_B3] = rng.uniform(-1, 1, H) * np.sqrt(3.0 / H)
    p[OFF_B3] = np.log2(9.0)
    return p


def forward(p, x):
    P = unpack(p)
    x = np.asarray(x, np.float32).reshape(-1, NIN)
    X = np.zeros((x.shape[0], K1))
    X[:, :NIN] = bf16(x)
    Z1 = X @ bf16(P["W1"]).T + P["b1"]
    A1 = bf16(act(Z1))
    Z2 = A1 @ bf16(P["W2"]).T + P["b2"]
    A2 = act(Z2)
    y = A2 @ P["w3"] + P["b3"]
    return y, dict(X=X, Z1=Z1, A1=A1, Z2=Z2, A2=A2)


def loss_grad(p, x, t, w_over=4.0):
    """Loss (1/n) sum w (y - t)^2 and its gradient, FP64 with the GPU's roundings."""
    P = unpack(p)
    y, c = forward(p, x)
    t = np.asarray(t, np.float64)
    n = y.shape[0]
    d = y - t
    w = np.where(d > 0, w_over, 1.0)
    loss = float(np.sum(w * d * d) / n)
    dy = 2.0 * w * d / n
    g = np.zeros(NPARAM)
    g[OFF_W3:OFF_B3] = c["A2"].T @ dy
    g[OFF_B3] = dy.sum()
    dZ2 = (dy[:, None] * P["w3"][None, :]) * act_grad(c["Z2"])
    g[OFF_B2:OFF_W3] = dZ2.sum(0)
    dZ2b = bf16(dZ2)
    g[OFF_W2:OFF_B2] = (dZ2b.T @ c["A1"]).reshape(-1)
    dA1 = dZ2b @ bf16(P["W2"])
    dZ1 = dA1 * act_grad(c["Z1"])
    g[OFF_B1:OFF_W2] = dZ1.sum(0)
    g[OFF_W1:OFF_B1] = (bf16(dZ1).T @ c["X"]).reshape(-1)
    return loss, g


def loss_grad_exact(p, x, t, w_over=4.0):
    """Same loss without any BF16 rounding (the FP64 mathematics), for finite
    differences and autograd cross-checks."""
    P = unpack(p)
    x = np.asarray(x, np.float64).reshape(-1, NIN)
    X = np.zeros((x.shape[0], K1))
    X[:, :NIN] = x
    Z1 = X @ P["W1"].T + P["b1"]
    A1 = act(Z1)
    Z2 = A1 @ P["W2"].T + P["b2"]
    A2 = act(Z2)
    y = A2 @ P["w3"] + P["b3"]
    t = np.asarray(t, np.float64)
    n = y.shape[0]
    d = y - t
    w = np.where(d > 0, w_over, 1.0)
    loss = float(np.sum(w * d * d) / n)
    dy = 2.0 * w * d / n
    g = np.zeros(NPARAM)
    g[OFF_W3:OFF_B3] = A2.T @ dy
    g[OFF_B3] = dy.sum()
    dZ2 = (dy[:, None] * P["w3"][None, :]) * act_grad(Z2)
    g[OFF_B2:OFF_W3] = dZ2.sum(0)
    g[OFF_W2:OFF_B2] = (dZ2.T @ A1).reshape(-1)
    dZ1 = (dZ2 @ P["W2"]) * act_grad(Z1)
    g[OFF_B1:OFF_W2] = dZ1.sum(0)
    g[OFF_W1:OFF_B1] = (dZ1.T @ X).reshape(-1)
    g[OFF_W1:OFF_B1].reshape(H, K1)[:, NIN:] = 0.0
    return loss, g
