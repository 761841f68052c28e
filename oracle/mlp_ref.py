"""numpy reference of the surrogate MLP (forward, MSE loss, backward).

TEST INFRASTRUCTURE ONLY: the reference computation the GPU MLP kernels' gradients are checked against.
It follows the GPU's precision plan exactly where the plan is defined by
rounding points: weights and activations rounded to BF16 (round to nearest
even), products and sums in FP64 (the GPU accumulates exact BF16 products in
FP32, so results agree to FP32 accumulation error plus the odd BF16 rounding
flip of an activation).

Parameter layout per member (ppvm_mlp_get_params): for l = 0..3, W_l
[N_l][K_l] then b_l [N_l]; K = (32, H, H, H), N = (H, H, H, 32).
"""

from __future__ import annotations

import math

import numpy as np

IN_PAD = 32
OUT_PAD = 32


def bf16(x) -> np.ndarray:
    """Round float32 values to BF16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    b = a.view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) >> 16
    return (b.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def _erf(x):
    return np.vectorize(math.erf)(x)


def gelu(x):
    return 0.5 * x * (1.0 + _erf(x / math.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + _erf(x / math.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / math.sqrt(2.0 * math.pi)


def shapes(hidden: int):
    K = (IN_PAD, hidden, hidden, hidden)
    N = (hidden, hidden, hidden, OUT_PAD)
    return K, N


def unpack(params: np.ndarray, hidden: int):
    K, N = shapes(hidden)
    out, off = [], 0
    for k, n in zip(K, N):
        W = params[off:off + n * k].reshape(n, k)
        off += n * k
        b = params[off:off + n]
        off += n
        out.append((W, b))
    assert off == len(params)
    return out


def forward_backward(params: np.ndarray, X: np.ndarray, Y: np.ndarray, hidden: int,
                     rounding: bool = True):
    """Loss and gradient vector (same layout as params) of one member.  With
    rounding=False everything is float64 (used to pin the formulas against
    finite differences)."""
    if not rounding:
        return _exact(np.asarray(params, dtype=np.float64), X, Y, hidden)
    layers = unpack(np.asarray(params, dtype=np.float32), hidden)
    n, d = X.shape
    a = np.zeros((n, IN_PAD), dtype=np.float32)
    a[:, :d] = X
    a = bf16(a).astype(np.float64)
    acts, pres = [a], []
    for li, (W, b) in enumerate(layers):
        Wb = bf16(W).astype(np.float64)
        pre = (a @ Wb.T + b.astype(np.float64)).astype(np.float32).astype(np.float64)
        pres.append(pre)
        if li + 1 < len(layers):
            a = bf16(gelu(pre)).astype(np.float64)
            acts.append(a)
    out = pres[-1][:, 0]
    err = (out.astype(np.float32) - Y.astype(np.float32)).astype(np.float64)
    loss = float(np.mean(err * err))
    d = np.zeros((n, OUT_PAD))
    d[:, 0] = bf16((2.0 * err / n).astype(np.float32))
    grads = []
    for li in range(len(layers) - 1, -1, -1):
        W, b = layers[li]
        gW = d.T @ acts[li]
        gb = d.sum(axis=0)
        grads.append((gW, gb))
        if li == 0:
            break
        da = d @ bf16(W).astype(np.float64)
        d = bf16((da * gelu_grad(pres[li - 1])).astype(np.float32)).astype(np.float64)
    grads.reverse()
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat.astype(np.float64)


def _exact(params, X, Y, hidden):
    layers = unpack(params, hidden)
    n, d = X.shape
    a = np.zeros((n, IN_PAD))
    a[:, :d] = X
    acts, pres = [a], []
    for li, (W, b) in enumerate(layers):
        pre = a @ W.T + b
        pres.append(pre)
        if li + 1 < len(layers):
            a = gelu(pre)
            acts.append(a)
    err = pres[-1][:, 0] - Y
    loss = float(np.mean(err * err))
    dd = np.zeros((n, OUT_PAD))
    dd[:, 0] = 2.0 * err / n
    grads = []
    for li in range(len(layers) - 1, -1, -1):
        W, b = layers[li]
        grads.append((dd.T @ acts[li], dd.sum(axis=0)))
        if li == 0:
            break
        dd = (dd @ W) * gelu_grad(pres[li - 1])
    grads.reverse()
    return loss, np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
