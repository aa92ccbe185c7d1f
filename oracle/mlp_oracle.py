"""ORACLE — test infrastructure, not product code.

FP64 restatement of the warm-start MLP (no counterpart in the reference: the
network is an addition of this implementation).  Only tests/ use it.

Network: x -> gelu(x W1^T + b1) -> gelu(. W2^T + b2) -> gelu(. W3^T + b3)
-> . W4^T + b4, with the exact (erf) GELU.  Loss: sum over rows and over
the LEARNABLE_CONSTANT slots of (out - label)^2, divided by the number of
rows.

`bf16=True` rounds to bfloat16 (round to nearest even) at the same points as
the device kernels: the input features, the weights, each hidden activation
and each backward GEMM operand (dO, dZ).  All arithmetic itself is FP64.

Parity pins: `check_gradient` compares the analytic backward against central
differences of the FP64 loss (bf16=False), and `features` is pinned against
a direct construction in tests/test_mlp_oracle.py.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

F64 = np.float64


def bf16_round(a) -> np.ndarray:
    """Round float values to the nearest bfloat16 (ties to even); NaN stays NaN."""
    x = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    out = np.where(np.isnan(x), np.float32(np.nan), out)
    return out.astype(F64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def features(ops, Y, slots: int, vocab: int, n_pts: int, row_len: int, bf16: bool = True,
             std_block: bool = False) -> np.ndarray:
    """Input rows: one-hot (slot, opcode) over the first `slots` nodes, then
    sign(y) log1p(|y|) of the first n_pts targets (NaN -> 0); with std_block
    then (y - mean) / std clipped to [-10, 10] (0 when std <= 1e-12) and
    sign(mean) log1p|mean|, log1p(std), the statistics over the n_pts targets
    with non-finite values read as 0; zero padded."""
    B, M = ops.shape
    out = np.zeros((B, row_len), dtype=F64)
    for k in range(min(slots, M)):
        c = ops[:, k]
        ok = (c >= 0) & (c < vocab)
        out[np.nonzero(ok)[0], k * vocab + c[ok]] = 1.0
    with np.errstate(over="ignore"):
        y = np.asarray(Y, dtype=np.float32)[:, :n_pts].astype(F64)
    g = np.sign(y) * np.log1p(np.minimum(np.abs(y), 3.0e38))
    g = np.where(np.isnan(y), 0.0, g)
    base = slots * vocab
    out[:, base: base + n_pts] = g
    if std_block:
        yc = np.where(np.isfinite(y) & (np.abs(y) <= 3.0e38), y, 0.0)
        mean = yc.mean(axis=1, keepdims=True)
        sd = np.sqrt(np.maximum((yc * yc).mean(axis=1, keepdims=True) - mean * mean, 0.0))
        z = np.where(sd > 1e-12, (yc - mean) / np.where(sd > 1e-12, sd, 1.0), 0.0)
        out[:, base + n_pts: base + 2 * n_pts] = np.clip(z, -10.0, 10.0)
        out[:, base + 2 * n_pts] = np.sign(mean[:, 0]) * np.log1p(np.abs(mean[:, 0]))
        out[:, base + 2 * n_pts + 1] = np.log1p(sd[:, 0])
    return bf16_round(out) if bf16 else out


def forward(params, x, bf16: bool = True):
    """params: list of (W (out, in), b (out,)).  Returns (out, cache)."""
    rnd = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=F64))
    h = rnd(x)
    cache = {"h": [h], "z": []}
    L = len(params)
    for l, (W, b) in enumerate(params):
        z = h @ rnd(W).T + np.asarray(b, dtype=F64)
        if l < L - 1:
            cache["z"].append(z)
            h = rnd(gelu(z))
            cache["h"].append(h)
        else:
            out = z
    return out, cache


def loss_and_grads(params, x, label, mask, bf16: bool = True):
    """Loss (sum over rows and masked slots of d^2, / rows) and its gradients
    [(dW, db), ...] in the layout of params."""
    rnd = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=F64))
    out, cache = forward(params, x, bf16)
    rows = out.shape[0]
    d = np.where(mask, out - label, 0.0)
    loss = float((d * d).sum() / rows)
    dz = 2.0 * d / rows
    grads = [None] * len(params)
    for l in range(len(params) - 1, -1, -1):
        W, _ = params[l]
        dz_r = rnd(dz)
        h = cache["h"][l]
        grads[l] = (dz_r.T @ h, dz.sum(axis=0))
        if l > 0:
            dh = dz_r @ rnd(W)
            dz = dh * gelu_grad(cache["z"][l - 1])
    return loss, grads


def adam(w, g, m, v, step: int, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8,
         wd: float = 0.0):
    """One Adam update (bias-corrected moments, L2 weight decay added to the
    gradient) in float32, as the device kernel computes it."""
    f = np.float32
    w, g, m, v = (np.asarray(a, dtype=f) for a in (w, g, m, v))
    gi = g + f(wd) * w
    m2 = f(b1) * m + (f(1) - f(b1)) * gi
    v2 = f(b2) * v + (f(1) - f(b2)) * gi * gi
    bc1 = f(1) - f(np.float32(b1) ** np.float32(step))
    bc2 = f(1) - f(np.float32(b2) ** np.float32(step))
    w2 = w - f(lr) * (m2 / bc1) / (np.sqrt(v2 / bc2) + f(eps))
    return w2, m2, v2


def check_gradient(params, x, label, mask, n_probe: int = 8, h: float = 1e-6, seed: int = 0):
    """Largest relative difference between the analytic FP64 gradient and
    central differences at n_probe random entries of every parameter."""
    rng = np.random.default_rng(seed)
    _, grads = loss_and_grads(params, x, label, mask, bf16=False)
    worst = 0.0
    for l, (W, b) in enumerate(params):
        for which, arr, g in ((0, W, grads[l][0]), (1, b, grads[l][1])):
            for _ in range(n_probe):
                idx = tuple(rng.integers(0, s) for s in arr.shape)
                old = arr[idx]
                arr[idx] = old + h
                lp, _ = loss_and_grads(params, x, label, mask, bf16=False)
                arr[idx] = old - h
                lm, _ = loss_and_grads(params, x, label, mask, bf16=False)
                arr[idx] = old
                fd = (lp - lm) / (2 * h)
                scale = max(abs(fd), abs(g[idx]), 1e-8)
                worst = max(worst, abs(fd - g[idx]) / scale)
    return worst
