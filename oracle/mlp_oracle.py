"""FP64 restatement of the learned-proposal network (TEST INFRASTRUCTURE ONLY).

The network is 12 -> 64 -> 64 -> 1 with ReLU; parameters are packed FP32 in
the order W1[12][64], b1[64], W2[64][64], b2[64], w3[64], b3 (the layout of
libhsmc_gpu's hsmc_gpu_mlp_* entry points).  With ``bf16=True`` the forward
and backward passes round exactly where the GPU kernels keep BF16 operands
(inputs, W1, W2, layer-1 activations, dH2, dH1); everything else is FP64.
The plain FP64 gradient is checked against central finite differences in
tests/test_mlp_oracle.py.
"""
from __future__ import annotations

import numpy as np

NIN, NH = 12, 64
OW1 = 0
OB1 = OW1 + NIN * NH
OW2 = OB1 + NH
OB2 = OW2 + NH * NH
OW3 = OB2 + NH
OB3 = OW3 + NH
NPARAM = OB3 + 1


def bf16(a):
    """Round FP32 values to BF16 (nearest, ties to even), returned as FP64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OW1:OB1].reshape(NIN, NH), p[OB1:OW2], p[OW2:OB2].reshape(NH, NH), p[OB2:OW3],
            p[OW3:OB3], p[OB3])


def forward(p, x, bf=False):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    if bf:
        x, W1, W2 = bf16(x), bf16(W1), bf16(W2)
    z1 = x @ W1 + b1
    h1 = np.maximum(z1, 0.0)
    if bf:
        h1 = bf16(f32(h1))
    z2 = h1 @ W2 + b2
    h2 = np.maximum(z2, 0.0)
    y = h2 @ w3 + b3
    return y, (x, W1, h1, W2, h2, w3)


def loss_and_grad(p, x, t, bf=False):
    """Mean squared error and its gradient w.r.t. the packed parameters."""
    n = len(x)
    y, (xx, W1, h1, W2, h2, w3) = forward(p, x, bf)
    e = y - np.asarray(t, dtype=np.float64)
    loss = float((e * e).mean())
    dy = 2.0 * e / n
    g = np.zeros(NPARAM)
    g[OW3:OB3] = h2.T @ dy
    g[OB3] = dy.sum()
    dh2 = np.where(h2 > 0, dy[:, None] * w3[None, :], 0.0)
    g[OB2:OW3] = dh2.sum(0)
    if bf:
        dh2 = bf16(f32(dh2))
    g[OW2:OB2] = (h1.T @ dh2).ravel()
    dh1 = np.where(h1 > 0, dh2 @ W2.T, 0.0)
    g[OB1:OW2] = dh1.sum(0)
    if bf:
        dh1 = bf16(f32(dh1))
    g[OW1:OB1] = (xx.T @ dh1).ravel()
    return loss, g


def numeric_grad(p, x, t, h=1e-6):
    p = np.asarray(p, dtype=np.float64)
    g = np.zeros_like(p)
    for i in range(len(p)):
        a = p.copy(); a[i] += h
        b = p.copy(); b[i] -= h
        g[i] = (loss_and_grad(a, x, t)[0] - loss_and_grad(b, x, t)[0]) / (2 * h)
    return g


def nearest12(conf, L, pts):
    """Brute-force 12 smallest distances (minus one) from each point to the
    sphere centres (minimum image), FP64."""
    c = np.asarray(conf)[:, 1:]
    out = np.zeros((len(pts), NIN))
    for k, q in enumerate(np.asarray(pts, dtype=np.float64)):
        d = q - c
        d = np.where(d > L / 2, d - L, np.where(d < -L / 2, d + L, d))
        r = np.sqrt((d * d).sum(1))
        out[k] = np.sort(r)[:NIN] - 1.0
    return out
