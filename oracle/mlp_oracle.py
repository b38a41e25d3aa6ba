"""TEST INFRASTRUCTURE ONLY — FP64 forward / backward of the learned-transform
encoder and its training loss, the checker for the HIP kernels.

The encoder is the library's own design (the reference's analog,
rust_umap/src/parametric.rs:89-534, is a one-hidden-layer tanh network trained
by a mean squared error to its teacher): x -> relu(x W1 + b1) ->
relu(. W2 + b2) -> . W3 + b3, inputs standardised with per-feature mean and
1/std and zero-padded to dp = ceil(d / 32) * 32 features.  Loss over a batch
of B edges with n_neg negatives each (rows: B heads, B tails, then
B * n_neg negatives):

    L = (1/B) sum_e [ log(1 + a s_ij^b) + sum_t log(1 + 1 / (a (s_ik + 1e-3)^b)) ]
        + lambda / S sum_rows |z - y|^2,     s = |z_u - z_v|^2,  S = B (2 + n_neg)

With emulate_bf16 the forward pass rounds what the GPU keeps in BF16 (the
standardised inputs, the weights, both hidden activations) and the backward
pass rounds the two hidden-activation gradients, so only FP32-vs-FP64
accumulation differs.  Only tests/ use this module.
"""
from __future__ import annotations

import numpy as np

REPULSION_EPS = 1e-3


def layout(d: int, h: int, dim: int) -> dict:
    dp = (d + 31) // 32 * 32
    off = {}
    o = 0
    for name, size in (("w1", dp * h), ("b1", h), ("w2", h * h), ("b2", h), ("w3", h * dim),
                       ("b3", dim)):
        off[name] = (o, size)
        o += size
    off["total"] = o
    off["dp"] = dp
    return off


def bf16(a: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (nearest, ties to even) and return as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params: np.ndarray, d: int, h: int, dim: int):
    L = layout(d, h, dim)
    p = np.asarray(params, dtype=np.float64)
    get = lambda k: p[L[k][0]:L[k][0] + L[k][1]]  # noqa: E731
    return (get("w1").reshape(L["dp"], h), get("b1"), get("w2").reshape(h, h), get("b2"),
            get("w3").reshape(h, dim), get("b3"))


def forward(params, x_rows, mean, inv_std, d, h, dim, emulate_bf16=False):
    W1, b1, W2, b2, W3, b3 = unpack(params, d, h, dim)
    L = layout(d, h, dim)
    xs = np.zeros((x_rows.shape[0], L["dp"]), np.float64)
    # the GPU forms (x - mean) * inv_std in f32
    xs[:, :d] = ((x_rows.astype(np.float32) - mean.astype(np.float32)) *
                 inv_std.astype(np.float32)).astype(np.float64)
    if emulate_bf16:
        xs, W1, W2, W3 = bf16(xs), bf16(W1), bf16(W2), bf16(W3)
    h1 = np.maximum(xs @ W1 + b1, 0.0)
    if emulate_bf16:
        h1 = bf16(h1)
    h2 = np.maximum(h1 @ W2 + b2, 0.0)
    if emulate_bf16:
        h2 = bf16(h2)
    z = h2 @ W3 + b3
    return z, (xs, h1, h2, W1, W2, W3)


def loss_and_dz(z, y_rows, n_edges, n_neg, a, b, lam):
    z = np.asarray(z, np.float64)
    y = np.asarray(y_rows, np.float64)
    B = n_edges
    S = B * (2 + n_neg)
    dz = np.zeros_like(z)
    loss = 0.0
    heads = np.arange(B)
    tails = B + heads
    zi, zj = z[heads], z[tails]
    s = ((zi - zj) ** 2).sum(1)
    pos = s > 0
    pb = np.where(pos, np.power(np.where(pos, s, 1.0), b), 0.0)
    P = a * pb
    loss += np.log1p(P[pos]).sum() / B
    dlds = np.where(pos, a * b * (pb / np.where(pos, s, 1.0)) / (1.0 + P), 0.0)
    g = (2.0 * dlds / B)[:, None] * (zi - zj)
    dz[heads] += g
    dz[tails] -= g
    for t in range(n_neg):
        ks = 2 * B + heads * n_neg + t
        zk = z[ks]
        se = ((zi - zk) ** 2).sum(1) + REPULSION_EPS
        P = a * np.power(se, b)
        loss += np.log1p(1.0 / P).sum() / B
        g = (2.0 * (-b / (se * (1.0 + P))) / B)[:, None] * (zi - zk)
        dz[heads] += g
        dz[ks] -= g
    diff = z[:S] - y[:S]
    loss += lam * (diff ** 2).sum() / S
    dz[:S] += 2.0 * lam / S * diff
    return loss, dz


def loss_and_grads(params, x_rows, y_rows, mean, inv_std, d, h, dim, n_edges, n_neg, a, b, lam,
                   emulate_bf16=False):
    z, (xs, h1, h2, W1, W2, W3) = forward(params, x_rows, mean, inv_std, d, h, dim, emulate_bf16)
    loss, dz = loss_and_dz(z, y_rows, n_edges, n_neg, a, b, lam)
    dW3 = h2.T @ dz
    db3 = dz.sum(0)
    dh2 = (dz @ W3.T) * (h2 > 0)
    if emulate_bf16:
        dh2 = bf16(dh2)
    dW2 = h1.T @ dh2
    db2 = dh2.sum(0)
    dh1 = (dh2 @ W2.T) * (h1 > 0)
    if emulate_bf16:
        dh1 = bf16(dh1)
    dW1 = xs.T @ dh1
    db1 = dh1.sum(0)
    g = np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dW3.ravel(), db3])
    return loss, g, z
