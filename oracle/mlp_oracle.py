"""CPU restatement of the surrogate MLP's forward and backward passes (numpy).

TEST INFRASTRUCTURE ONLY: the checker of the HIP training kernels
(dwarf-p-ice3_amd/csrc/ice3_mlp.hip).  Only tests/ may use it.

It follows the kernels' arithmetic contract: bf16 weights and activations (round to
nearest even), products accumulated in high precision (float64 here, fp32 on the
matrix cores), biases and the SiLU in fp32, pre-activations kept in fp32, the loss
gradient rounded to bf16 before it enters the backward products, dZ rounded to bf16
between layers, bias gradients summed from the bf16 dZ.

The MLP has no counterpart in the reference (dwarf-p-ice3 has no learned component);
this module is pinned by the tests against finite differences of its own forward pass
(float64, no rounding) and, through the kernels, against the GPU.
"""
from __future__ import annotations

import numpy as np

IN, H, OUT, NTARG, LAYERS = 32, 128, 32, 8, 5
DIMS = [(IN, H), (H, H), (H, H), (H, H), (H, OUT)]
NW = sum(k * n for k, n in DIMS)
NB = 4 * H + OUT
NP = NW + NB


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    return np.where(np.isnan(a), a, out)


def unpack(params):
    """Split the flat parameter vector into (W_l, b_l) lists (row-major K x N)."""
    ws, bs, o = [], [], 0
    for k, n in DIMS:
        ws.append(params[o:o + k * n].reshape(k, n))
        o += k * n
    ob = NW
    for l, (k, n) in enumerate(DIMS):
        bs.append(params[ob:ob + n])
        ob += n
    return ws, bs


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, x, emulate=True):
    """x: (B, 32) features (already bf16 values).  Returns y (B, 32), [z_l], [h_l]."""
    ws, bs = unpack(np.asarray(params, dtype=np.float32 if emulate else np.float64))
    rb = bf16 if emulate else (lambda v: np.asarray(v, dtype=np.float64))
    h = np.asarray(x, dtype=np.float64)
    zs, hs = [], [h]
    for l in range(LAYERS):
        w = rb(ws[l]).astype(np.float64)
        z = (h @ w).astype(np.float32 if emulate else np.float64) + bs[l]
        if l < LAYERS - 1:
            zs.append(z)
            h = rb(silu(z)).astype(np.float64)
            hs.append(h)
        else:
            y = z
    return y, zs, hs


def loss_and_grad(params, x, target, emulate=True):
    """Mean squared error over the 8 outputs and its gradient w.r.t. the parameters."""
    B = x.shape[0]
    rb = bf16 if emulate else (lambda v: np.asarray(v, dtype=np.float64))
    y, zs, hs = forward(params, x, emulate)
    d = np.zeros((B, OUT))
    d[:, :NTARG] = y[:, :NTARG] - target
    loss = float((d ** 2).sum())
    dz = rb(2.0 * d / (B * NTARG)).astype(np.float64)
    ws, _ = unpack(np.asarray(params, dtype=np.float32 if emulate else np.float64))
    gw, gb = [None] * LAYERS, [None] * LAYERS
    for l in range(LAYERS - 1, -1, -1):
        gw[l] = hs[l].T @ dz
        gb[l] = dz.sum(axis=0)
        if l > 0:
            dh = dz @ rb(ws[l]).astype(np.float64).T
            dz = rb(dh * silu_grad(zs[l - 1])).astype(np.float64)
    grad = np.concatenate([g.ravel() for g in gw] + [g for g in gb])
    return loss, grad
