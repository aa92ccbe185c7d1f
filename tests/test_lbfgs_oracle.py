"""Pins of the CPU L-BFGS restatement (oracle/lbfgs_oracle.py) and of the
package's LBFGS_batch mirror, on the CPU:

* the reference's known answers (tests/test_optimization.py:10-169):
  y = c1*x + c2 and y = c1*sin(x) + c2 are recovered to 1e-3, with either
  line search;
* a one-constant fit agrees with a brute-force scan of its loss;
* the first step follows the algorithm's rule (p = -g, alpha0 = 1, no
  scaling before the first pair): with one iteration the iterate is the
  one the rule gives, computed by hand from the closure's gradient;
* LBFGS_batch reaches the Rosenbrock minimum (1, 1)
  (reference tests/test_lbfgs_batch.py:39-75)."""
import numpy as np
import pytest
import torch

import helpers
import lbfgs_oracle
import oracle


def _linear_batch(B=3, N=64, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-2, 2, (N, 1)).astype(np.float32)
    c = rng.standard_normal((B, 2)).astype(np.float32) * 2
    Y = (c[:, :1] * X[:, 0][None, :] + c[:, 1:]).astype(np.float32)
    ops = np.tile(np.array([50, 1, 42, 1, 40], np.int64), (B, 1))
    ch = np.stack([helpers.children_of(r) for r in ops])
    C0 = helpers.constants_like(ops, rng)
    return X, Y, ops, ch, C0, c


@pytest.mark.parametrize("ls", ["strong_wolfe", "backtracking"])
def test_linear_known_answer(ls):
    X, Y, ops, ch, C0, c = _linear_batch()
    C, y, f, _ = lbfgs_oracle.optimize(X, Y, ops, ch, C0, max_iter=300, line_search_fn=ls)
    assert np.allclose(C[:, 1], c[:, 0], atol=1e-3) and np.allclose(C[:, 3], c[:, 1], atol=1e-3)
    assert np.array_equal(C[:, [0, 2, 4]], C0[:, [0, 2, 4]])  # non-constant slots untouched


def test_sin_known_answer():
    rng = np.random.default_rng(1)
    B, N = 8, 50
    X = rng.uniform(-2, 2, (N, 1)).astype(np.float32)
    c = np.stack([rng.standard_normal(B) * 2, rng.standard_normal(B)], 1).astype(np.float32)
    Y = (c[:, :1] * np.sin(X[:, 0])[None, :] + c[:, 1:]).astype(np.float32)
    ops = np.tile(np.array([50, 10, 1, 42, 1, 40], np.int64), (B, 1))
    ch = np.stack([helpers.children_of(r) for r in ops])
    C0 = helpers.constants_like(ops, rng)
    C, _, _, _ = lbfgs_oracle.optimize(X, Y, ops, ch, C0, max_iter=100)
    assert np.allclose(C[:, 2], c[:, 0], atol=1e-3) and np.allclose(C[:, 4], c[:, 1], atol=1e-3)


def test_one_constant_matches_brute_force():
    """y_hat = exp(c * x) fitted to a noisy target: the fitted c is the
    minimiser of the FP64 loss found by a dense scan."""
    rng = np.random.default_rng(2)
    N = 40
    X = rng.uniform(-1, 1, (N, 1)).astype(np.float32)
    Y = (np.exp(0.7 * X[:, 0]) + rng.normal(0, 0.05, N)).astype(np.float32)[None, :]
    ops = np.array([[1, 50, 42, 12]], np.int64)
    ch = helpers.children_of(ops[0])[None]
    C0 = np.array([[0.1, 0, 0, 0]], np.float32)
    C, _, _, _ = lbfgs_oracle.optimize(X, Y, ops, ch, C0, max_iter=100)
    grid = np.linspace(0.0, 1.5, 150001)
    loss = ((np.exp(grid[:, None] * X[None, :, 0].astype(np.float64)) - Y[0][None, :].astype(np.float64)) ** 2).mean(1)
    c_star = grid[np.argmin(loss)]
    assert abs(C[0, 0] - c_star) < 1e-4, (C[0, 0], c_star)


def test_first_iteration_rule():
    """max_iter = 1 with backtracking: x1 = x0 + alpha * (-g0) with the first
    alpha in 1, 1/2, 1/4, ... that satisfies Armijo (c = 1e-4)."""
    X, Y, ops, ch, C0, _ = _linear_batch(B=1, seed=3)
    mask = (ops == 1)
    r = oracle.fused_root(X, ops, ch, C0, Y, mode=1)
    f0, g0 = r["loss"][0], np.where(mask, r["grad_C"], 0.0).astype(np.float32)[0]
    p = -g0
    alpha = np.float32(1.0)
    for _ in range(10):
        xt = (C0[0] + alpha * p).astype(np.float32)
        ft = oracle.fused_root(X, ops, ch, xt[None], Y, mode=1)["loss"][0]
        if ft <= f0 + 1e-4 * alpha * float(np.dot(g0.astype(np.float64), p)):
            break
        alpha = np.float32(alpha * 0.5)
    C, _, _, _ = lbfgs_oracle.optimize(X, Y, ops, ch, C0, max_iter=1, line_search_fn="backtracking")
    assert np.allclose(C[0], C0[0] + alpha * p, rtol=1e-6, atol=1e-7), (C[0], C0[0] + alpha * p)


@pytest.mark.parametrize("ls", ["strong_wolfe", "backtracking"])
def test_lbfgs_batch_rosenbrock(ls):
    """The reference's own correctness test (tests/test_lbfgs_batch.py:39-75):
    10 problems from randn starts, max_iter 250, history 15."""
    from cudasymbolicregression_amd.lbfgs_batch import LBFGS_batch

    torch.manual_seed(0)
    n = 10
    x = torch.randn(2, n, requires_grad=True)
    opt = LBFGS_batch([x], max_iter=250, history_size=15, tolerance_grad=1e-7, line_search_fn=ls)

    def f_of(v):
        return (1.0 - v[0]) ** 2 + 100.0 * (v[1] - v[0] ** 2) ** 2

    def closure():
        opt.zero_grad(set_to_none=True)
        f = f_of(opt.param_groups[0]["params"][0])
        f.sum().backward()
        return f

    opt.step(closure)
    assert f_of(x).mean() < 1e-4
    assert torch.allclose(x.detach(), torch.ones(2, n), atol=1e-3)
