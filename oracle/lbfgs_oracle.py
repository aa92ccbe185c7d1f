"""ORACLE — test infrastructure, not product code.

CPU restatement of the reference's constant fitting: LBFGS_batch.step with
the strong-Wolfe bracketing + zoom and the Armijo backtracking
(symbolic_torch/lbfgs_batch.py:61-326) driven by optimize's MSE closure
(symbolic_torch/optimization.py:9-112).  Batched over problems exactly like
the reference (full-batch closures, masks, `x += s * (~done)`), in float32
numpy, with each closure evaluated by the C oracle (oracle/eval_oracle.c,
fused root-only forward/backward in mode 1).

Only tests/ and bench.py's cpu_baseline leg use it.

Parity pins: recovers the known answers of the reference's own optimize
tests (y = c1*x + c2, y = c1*sin(x) + c2) and, for one-constant problems,
agrees with a brute-force scan of the loss; the product's device L-BFGS is
then compared against it.
"""
from __future__ import annotations

import numpy as np

import oracle

F32 = np.float32


class _Closure:
    def __init__(self, X, Y, ops, ch, const_mask):
        self.X, self.Y, self.ops, self.ch = X, Y, ops, ch
        self.mask = const_mask  # (M, B) bool
        self.calls = 0

    def __call__(self, x_mb):
        """x (M, B) float32 -> (loss (B,) float32, grad (M, B) float32)."""
        self.calls += 1
        out = oracle.fused_root(self.X, self.ops, self.ch, np.ascontiguousarray(x_mb.T), self.Y, mode=1,
                                want_gx=False)
        d = (out["y"] - self.Y).astype(F32)
        loss = (d * d).mean(axis=1, dtype=F32)
        g = out["grad_C"].T.copy()
        g[~self.mask] = 0.0
        self.err = out["err"]
        return loss.astype(F32), g.astype(F32)


def _bdot(a, b):
    return (a * b).sum(axis=0, keepdims=True, dtype=F32)


def _direction(g, S, Yh, eps):
    if not S:
        return g
    q = g
    alphas = [None] * len(S)
    for i in range(len(S) - 1, -1, -1):
        rho = F32(1.0) / (_bdot(Yh[i], S[i]) + F32(eps))
        alphas[i] = rho * _bdot(S[i], q)
        q = q - alphas[i] * Yh[i]
    gamma = _bdot(S[-1], Yh[-1]) / (_bdot(Yh[-1], Yh[-1]) + F32(eps))
    z = gamma * q
    for i in range(len(S)):
        rho = F32(1.0) / (_bdot(Yh[i], S[i]) + F32(eps))
        beta = rho * _bdot(Yh[i], z)
        z = z + S[i] * (alphas[i] - beta)
    return z


def _zoom(cl, x, p, lo, hi, active, c1=1e-4, c2=0.9):
    phi0, g0 = cl(x)
    dphi0 = _bdot(g0, p)[0]
    phi_lo, _ = cl(x + lo[None, :] * p)
    alpha = np.zeros_like(lo)
    fin = np.zeros(lo.shape, bool)
    for _ in range(10):
        if np.all(fin[active]):
            break
        aj = ((lo + hi) / F32(2)).astype(F32)
        phi, gj = cl(x + aj[None, :] * p)
        dphi = _bdot(gj, p)[0]
        bad = (phi > phi0 + F32(c1) * aj * dphi0) | (phi >= phi_lo)
        good = np.abs(dphi) <= F32(c2) * np.abs(dphi0)
        take = active & ~fin & good
        alpha = np.where(take, aj, alpha)
        fin |= take
        live = active & ~fin
        t_hi = np.where(live & bad, aj, hi)
        cond3 = dphi * (t_hi - lo) >= 0
        hi = np.where(live & ~bad & cond3, lo, t_hi)
        mv = live & ~bad
        lo = np.where(mv, aj, lo)
        phi_lo = np.where(mv, phi, phi_lo)
        small = np.abs(hi - lo) < F32(1e-9)
        st = active & ~fin & small
        alpha = np.where(st, aj, alpha)
        fin |= st
    return np.where(active & ~fin, lo, alpha).astype(F32)


def _strong_wolfe(cl, x, p, max_ls, c1=1e-4, c2=0.9):
    phi0, g0 = cl(x)
    dphi0 = _bdot(g0, p)[0]
    a_prev = np.zeros_like(phi0)
    phi_prev = phi0.copy()
    a = np.ones_like(phi0)
    final = np.zeros_like(phi0)
    done = np.zeros(phi0.shape, bool)
    for i in range(max_ls):
        if np.all(done):
            break
        phi, ga = cl(x + a[None, :] * p)
        dphi = _bdot(ga, p)[0]
        bad = (phi > phi0 + F32(c1) * a * dphi0) | ((i > 0) & (phi >= phi_prev))
        good = np.abs(dphi) <= F32(c2) * np.abs(dphi0)
        passed = dphi >= 0
        ng = ~done & good
        final = np.where(ng, a, final)
        done |= ng
        nz = ~done & (bad | passed)
        if np.any(nz):
            lo = np.where(bad, a_prev, a)
            hi = np.where(bad, a, a_prev)
            final = np.where(nz, _zoom(cl, x, p, lo, hi, nz, c1, c2), final)
            done |= nz
        cont = ~done
        a_prev = np.where(cont, a, a_prev)
        phi_prev = np.where(cont, phi, phi_prev)
        a = np.where(cont, (a + (F32(20) - a) * F32(0.8)).astype(F32), a)
    return np.where(done, final, a).astype(F32)[None, :]


def _backtracking(cl, x, p, f, g, max_ls):
    slope = _bdot(g, p)[0]
    alpha = np.ones_like(f)
    for _ in range(max_ls):
        ft, _ = cl(x + alpha[None, :] * p)
        ok = ft <= f + F32(1e-4) * alpha * slope
        if ok.all():
            break
        alpha = np.where(ok, alpha, alpha * F32(0.5)).astype(F32)
    return alpha[None, :]


def optimize(X, Y, ops_bm, ch_bm, C_init_bm, max_iter=100, history_size=10, tolerance_grad=1e-7,
             line_search_fn="strong_wolfe", line_search_max_iter=10, eps=1e-10):
    """Returns (C_final (B, M), Y_final (B, N), final loss (B,), closure calls)."""
    X = np.ascontiguousarray(X, F32)
    Y = np.ascontiguousarray(Y, F32)
    mask = (np.asarray(ops_bm) == 1).T
    cl = _Closure(X, Y, np.ascontiguousarray(ops_bm), np.ascontiguousarray(ch_bm), mask)
    x = np.ascontiguousarray(np.asarray(C_init_bm, F32).T).copy()
    f, g = cl(x)
    S, Yh = [], []
    done = np.zeros(f.shape, bool)
    for _ in range(max_iter):
        with np.errstate(all="ignore"):
            gn = np.sqrt((g.astype(np.float64) ** 2).sum(axis=0))
        done |= gn < tolerance_grad
        if done.all():
            break
        with np.errstate(all="ignore"):
            p = -_direction(g, S, Yh, eps)
            if line_search_fn == "strong_wolfe":
                alpha = _strong_wolfe(cl, x, p, line_search_max_iter)
            else:
                alpha = _backtracking(cl, x, p, f, g, line_search_max_iter)
            s = (alpha * p).astype(F32)
            x = (x + s * (~done)[None, :].astype(F32)).astype(F32)
        f_n, g_n = cl(x)
        S.append(s)
        Yh.append((g_n - g).astype(F32))
        if len(S) > history_size:
            S.pop(0)
            Yh.pop(0)
        g, f = g_n, f_n
    C_final = x.T.copy()
    y = oracle.fused_root(X, np.ascontiguousarray(ops_bm), np.ascontiguousarray(ch_bm), C_final, Y, mode=1,
                          want_gx=False)["y"]
    return C_final, y, f, cl.calls
