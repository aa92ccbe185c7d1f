"""Pins for the warm-start MLP oracle (oracle/mlp_oracle.py), CPU only.

Each check is against something the oracle's formula does not restate:
central differences of the loss, torch's own bfloat16 rounding, torch's
exact GELU and its autograd derivative, torch.optim.Adam, and a hand-built
feature row.
"""
import os
import sys

import numpy as np
import torch

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as mo  # noqa: E402


def _small_net(rng, dims):
    return [(rng.standard_normal((dims[i + 1], dims[i])) / np.sqrt(dims[i]), rng.standard_normal(dims[i + 1]) * 0.1)
            for i in range(len(dims) - 1)]


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(0)
    dims = [24, 16, 16, 16, 8]
    params = _small_net(rng, dims)
    x = rng.standard_normal((6, dims[0]))
    label = rng.standard_normal((6, dims[-1]))
    mask = rng.random((6, dims[-1])) < 0.5
    assert mo.check_gradient(params, x, label, mask, n_probe=6) < 1e-6


def test_gradient_check_catches_a_transposed_operand():
    """A plausible slip (the weight gradient transposed within a square
    layer) must fail the central-difference check."""
    rng = np.random.default_rng(1)
    dims = [8, 8, 8]
    params = _small_net(rng, dims)
    x = rng.standard_normal((5, 8))
    label = rng.standard_normal((5, 8))
    mask = np.ones((5, 8), dtype=bool)
    _, grads = mo.loss_and_grads(params, x, label, mask, bf16=False)
    h = 1e-6
    W = params[0][0]
    worst = 0.0
    for idx in [(0, 1), (2, 5), (7, 3)]:
        old = W[idx]
        W[idx] = old + h
        lp, _ = mo.loss_and_grads(params, x, label, mask, bf16=False)
        W[idx] = old - h
        lm, _ = mo.loss_and_grads(params, x, label, mask, bf16=False)
        W[idx] = old
        fd = (lp - lm) / (2 * h)
        wrong = grads[0][0].T[idx]
        worst = max(worst, abs(fd - wrong) / max(abs(fd), 1e-8))
    assert worst > 1e-2


def test_bf16_round_matches_torch():
    rng = np.random.default_rng(2)
    v = np.concatenate([
        rng.standard_normal(20000).astype(np.float32) * np.float32(1e3),
        # exact ties: low 16 bits 0x8000 with even and odd kept halves
        (np.arange(1000, dtype=np.uint32) << 16 | 0x8000).astype(np.uint32).view(np.float32),
        np.array([0.0, -0.0, np.inf, -np.inf, 3.0e38, -1.17e-38, 1e-45], dtype=np.float32),
    ])
    want = torch.from_numpy(v).to(torch.bfloat16).to(torch.float64).numpy()
    got = mo.bf16_round(v)
    assert np.array_equal(got, want)
    assert np.isnan(mo.bf16_round(np.array([np.nan], dtype=np.float32))).all()


def test_gelu_and_derivative_match_torch():
    z = torch.linspace(-8, 8, 2001, dtype=torch.float64, requires_grad=True)
    y = torch.nn.functional.gelu(z)
    (g,) = torch.autograd.grad(y.sum(), z)
    zn = z.detach().numpy()
    np.testing.assert_allclose(mo.gelu(zn), y.detach().numpy(), rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(mo.gelu_grad(zn), g.numpy(), rtol=1e-12, atol=1e-14)


def test_adam_matches_torch_optim():
    rng = np.random.default_rng(3)
    w0 = rng.standard_normal(257).astype(np.float32)
    grads = [rng.standard_normal(257).astype(np.float32) for _ in range(3)]
    p = torch.nn.Parameter(torch.from_numpy(w0.copy()))
    opt = torch.optim.Adam([p], lr=3e-3, betas=(0.9, 0.999), eps=1e-8, weight_decay=1e-2)
    w, m, v = w0.copy(), np.zeros_like(w0), np.zeros_like(w0)
    for t, g in enumerate(grads, start=1):
        p.grad = torch.from_numpy(g.copy())
        opt.step()
        w, m, v = mo.adam(w, g, m, v, t, lr=3e-3, wd=1e-2)
    np.testing.assert_allclose(w, p.detach().numpy(), rtol=2e-6, atol=1e-7)


def test_features_hand_built_row():
    ops = np.array([[50, 1, 42, 0], [1, 51, 40, 3]], dtype=np.int64)
    Y = np.array([[0.5, -2.0, np.nan], [0.0, 1e39, -1.0]], dtype=np.float64)
    vocab, slots, n_pts = 53, 4, 3
    row_len = 64 * ((slots * vocab + n_pts + 63) // 64)
    f = mo.features(ops, Y, slots, vocab, n_pts, row_len, bf16=False)
    want = np.zeros((2, row_len))
    for b in range(2):
        for k in range(4):
            want[b, k * vocab + ops[b, k]] = 1.0
    base = slots * vocab
    want[0, base:base + 3] = [np.log1p(0.5), -np.log1p(2.0), 0.0]
    want[1, base:base + 3] = [0.0, np.log1p(3.0e38), -np.log1p(1.0)]  # 1e39 -> inf in FP32 -> clamped
    np.testing.assert_allclose(f, want, rtol=1e-15)
    assert f.shape == (2, row_len) and row_len % 64 == 0


def test_features_standardised_block():
    """The standardised block against a direct computation: constant targets
    give zeros, a NaN reads as 0 in the statistics, values clip at +-10."""
    ops = np.array([[1, 50, 42]], dtype=np.int64)
    vocab, slots, n_pts = 53, 3, 4
    row_len = 64 * ((slots * vocab + 2 * n_pts + 2 + 63) // 64)
    base = slots * vocab
    Y = np.array([[1.0, 2.0, np.nan, 5.0]])
    f = mo.features(ops, Y, slots, vocab, n_pts, row_len, bf16=False, std_block=True)
    yc = np.array([1.0, 2.0, 0.0, 5.0])
    m, sd = yc.mean(), yc.std()
    np.testing.assert_allclose(f[0, base + n_pts: base + 2 * n_pts], (yc - m) / sd, rtol=1e-14)
    np.testing.assert_allclose(f[0, base + 2 * n_pts: base + 2 * n_pts + 2], [np.log1p(m), np.log1p(sd)], rtol=1e-14)
    flat = mo.features(ops, np.full((1, 4), 3.0), slots, vocab, n_pts, row_len, bf16=False, std_block=True)
    assert not flat[0, base + n_pts: base + 2 * n_pts].any()
    spike = np.zeros((1, 200))
    spike[0, 0] = 1e4
    g = mo.features(ops, spike, slots, vocab, 200, 64 * ((slots * vocab + 402 + 63) // 64), bf16=False,
                    std_block=True)
    assert g[0, base + 200] == 10.0
