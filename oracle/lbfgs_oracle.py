"""ORACLE — test infrastructure, not product code.

CPU restatement of the reference's constant fitting: LBFGS_batch.step with
the strong-Wolfe bracketing + zoom and the Armijo backtracking
(symbolic_torch/lbfgs_batch.py:61-326) driven by optimize's MSE closure
(symbolic_torch/optimization.py:9-112).  Batched over problems exactly like
the reference (full-batch closures, masks, `x += s * (~done)`), in float32
numpy, with each closure evaluated by the C oracle (oracle/eval_oracle.c,
fused root-only forward/backward in mode 1).

Only tests/ and bench.py's cpu_baseline leg use it.

Parity pins (tests/test_lbfgs_golden.py, tests/test_lbfgs_oracle.py):
* golden trajectories of the reference's own LBFGS_batch (written by
  tests/golden/make_lbfgs_golden.py): fed the reference's recorded losses and
  gradients, lbfgs_run asks for the same sequence of points bit for bit, for
  both line searches, through zooms, exhausted searches, history wrap-around,
  inf/NaN losses and early-finishing problems;
* hand-computed first iterations of the strong-Wolfe search on 1-D
  quadratics (acceptance at alpha 1 and 16.2, curvature before bracketing,
  a zoom, an exhausted search returning the never-evaluated next alpha);
* the known answers of the reference's optimize tests (y = c1*x + c2,
  y = c1*sin(x) + c2) and, for a one-constant problem, a brute-force scan of
  the loss.
The product's device L-BFGS is then compared against this restatement.
"""
from __future__ import annotations

import numpy as np

import oracle

F32 = np.float32


class _Closure:
    """optimize's closure (optimization.py:72-99) on the C oracle.  With
    noise = (eps, seed), every returned loss and gradient value is scaled by
    1 + eps * r, r = +-1 drawn per value: a rounding-level perturbation of
    the closure, used by the tests to tell trajectories that hinge on the
    last bits of the loss and gradient (which FP64 reductions change) from
    those that do not."""

    def __init__(self, X, Y, ops, ch, const_mask, noise=None, fp64=False):
        self.X, self.Y, self.ops, self.ch = X, Y, ops, ch
        self.mask = const_mask  # (M, B) bool
        self.calls = 0
        self.fp64 = fp64  # loss (double) and grad_C summed in double, as the device does
        self.noise = None
        if noise is not None:
            self.noise = (F32(noise[0]), np.random.default_rng(noise[1]))

    def __call__(self, x_mb):
        """x (M, B) float32 -> (loss (B,) float32, grad (M, B) float32)."""
        self.calls += 1
        out = oracle.fused_root(self.X, self.ops, self.ch, np.ascontiguousarray(x_mb.T), self.Y, mode=1,
                                want_gx=False, acc64=self.fp64)
        if self.fp64:
            loss = out["loss"]
        else:
            d = (out["y"] - self.Y).astype(F32)
            loss = (d * d).mean(axis=1, dtype=F32)
        g = out["grad_C"].T.copy()
        g[~self.mask] = 0.0
        self.err = out["err"]
        if self.noise is not None:
            eps, rng = self.noise
            loss = loss * (F32(1) + eps * rng.choice(np.array([-1, 1], F32), loss.shape))
            g = g * (F32(1) + eps * rng.choice(np.array([-1, 1], F32), g.shape))
        return loss.astype(np.float64 if self.fp64 else F32), g.astype(F32)


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


def lbfgs_run(cl, x0, max_iter=100, history_size=10, tolerance_grad=1e-7, line_search_fn="strong_wolfe",
              line_search_max_iter=10, eps=1e-10):
    """LBFGS_batch.step (symbolic_torch/lbfgs_batch.py:266-326) on any closure
    cl(x (D, B) float32) -> (f (B,) float32, g (D, B) float32).  Returns
    (x_final (D, B), f of the last closure call (B,))."""
    x = np.ascontiguousarray(np.asarray(x0, F32)).copy()
    f, g = cl(x)  # lbfgs_batch.py:277-281
    S, Yh = [], []
    done = np.zeros(f.shape, bool)
    for _ in range(max_iter):  # :289
        with np.errstate(all="ignore"):
            gn = np.sqrt((g.astype(np.float64) ** 2).sum(axis=0))
        done |= gn < tolerance_grad  # :291-293 (NaN never done)
        if done.all():
            break
        with np.errstate(all="ignore"):
            p = -_direction(g, S, Yh, eps)  # :298
            if line_search_fn == "strong_wolfe":
                alpha = _strong_wolfe(cl, x, p, line_search_max_iter)
            else:
                alpha = _backtracking(cl, x, p, f, g, line_search_max_iter)
            s = (alpha * p).astype(F32)  # :305
            x = (x + s * (~done)[None, :].astype(F32)).astype(F32)  # :307, NaN * 0 leaks into done rows
        f_n, g_n = cl(x)  # :309-312
        S.append(s)
        Yh.append((g_n - g).astype(F32))
        if len(S) > history_size:  # :318-320
            S.pop(0)
            Yh.pop(0)
        g, f = g_n, f_n
    return x, f


# ---------------------------------------------------------------------------
# The same algorithm with the device's precision choices (the north star's
# FP64 steps): losses, inner products (two-loop, slopes, norm) and the Wolfe
# tests in double; vectors, step lengths and the two-loop coefficients
# rounded to float32 where they scale a vector.  The closure returns the
# loss in double and the gradient summed in double, rounded to float32.
# Not pinned bit for bit (the device's reduction order is its own); used to
# tell where the reference's FP32 arithmetic and FP64 arithmetic take a
# problem to different ends.
# ---------------------------------------------------------------------------
def _dot64(a, b):
    return (a.astype(np.float64) * b.astype(np.float64)).sum(axis=0)


def _direction64(g, S, Yh, rho, eps):
    q = g.copy()
    al = [0.0] * len(S)
    for i in range(len(S) - 1, -1, -1):
        al[i] = rho[i] * _dot64(S[i], q)
        q = (q - al[i].astype(F32)[None, :] * Yh[i]).astype(F32)
    if not S:
        return q
    gamma = (_dot64(S[-1], Yh[-1]) / (_dot64(Yh[-1], Yh[-1]) + eps)).astype(F32)
    z = (gamma[None, :] * q).astype(F32)
    for i in range(len(S)):
        beta = rho[i] * _dot64(Yh[i], z)
        z = (z + S[i] * (al[i] - beta).astype(F32)[None, :]).astype(F32)
    return z


def _strong_wolfe64(cl, x, p, f, dphi0, max_ls, c1=1e-4, c2=0.9):
    B = f.shape[0]
    a_prev = np.zeros(B, F32)
    phi_prev = f.copy()
    a = np.ones(B, F32)
    final = np.zeros(B, F32)
    done = np.zeros(B, bool)
    for i in range(max_ls):
        if np.all(done):
            break
        phi, ga = cl((x + a[None, :] * p).astype(F32))
        dphi = _dot64(ga, p)
        bad = (phi > f + c1 * a.astype(np.float64) * dphi0) | ((i > 0) & (phi >= phi_prev))
        good = np.abs(dphi) <= c2 * np.abs(dphi0)
        ng = ~done & good
        final = np.where(ng, a, final)
        done |= ng
        nz = ~done & (bad | (dphi >= 0))
        if np.any(nz):
            lo = np.where(bad, a_prev, a)
            hi = np.where(bad, a, a_prev)
            final = np.where(nz, _zoom64(cl, x, p, f, dphi0, lo, hi, nz, c1, c2), final)
            done |= nz
        cont = ~done
        a_prev = np.where(cont, a, a_prev)
        phi_prev = np.where(cont, phi, phi_prev)
        a = np.where(cont, (a + (F32(20) - a) * F32(0.8)).astype(F32), a)
    return np.where(done, final, a).astype(F32)


def _zoom64(cl, x, p, f, dphi0, lo, hi, active, c1=1e-4, c2=0.9):
    lo, hi = lo.copy(), hi.copy()
    phi_lo, _ = cl((x + lo[None, :] * p).astype(F32))  # as the reference (lbfgs_batch.py:218-221)
    alpha = np.zeros_like(lo)
    fin = np.zeros(lo.shape, bool)
    for _ in range(10):
        if np.all(fin[active]):
            break
        aj = ((lo + hi) / F32(2)).astype(F32)
        phi, gj = cl((x + aj[None, :] * p).astype(F32))
        dphi = _dot64(gj, p)
        bad = (phi > f + c1 * aj.astype(np.float64) * dphi0) | (phi >= phi_lo)
        good = np.abs(dphi) <= c2 * np.abs(dphi0)
        take = active & ~fin & good
        alpha = np.where(take, aj, alpha)
        fin |= take
        live = active & ~fin
        t_hi = np.where(live & bad, aj, hi)
        flip = dphi * (t_hi - lo).astype(np.float64) >= 0
        hi = np.where(live & ~bad & flip, lo, t_hi)
        mv = live & ~bad
        lo = np.where(mv, aj, lo)
        phi_lo = np.where(mv, phi, phi_lo)
        st = active & ~fin & (np.abs(hi - lo) < F32(1e-9))
        alpha = np.where(st, aj, alpha)
        fin |= st
    return np.where(active & ~fin, lo, alpha).astype(F32)


def _backtracking64(cl, x, p, f, dphi0, max_ls):
    alpha = np.ones(f.shape, F32)
    for _ in range(max_ls):
        ft, _ = cl((x + alpha[None, :] * p).astype(F32))
        ok = ft <= f + 1e-4 * alpha.astype(np.float64) * dphi0
        if ok.all():
            break
        alpha = np.where(ok, alpha, alpha * F32(0.5)).astype(F32)
    return alpha


def lbfgs_run64(cl, x0, max_iter=100, history_size=10, tolerance_grad=1e-7, line_search_fn="strong_wolfe",
                line_search_max_iter=10, eps=1e-10):
    """lbfgs_run with the device's precision choices (see above); cl returns
    (f float64 (B,), g float32 (D, B)).  Problems are independent here: the
    loop runs until every problem is done, which is what the device's
    per-problem loops give too."""
    x = np.ascontiguousarray(np.asarray(x0, F32)).copy()
    f, g = cl(x)
    S, Yh, rho = [], [], []
    done = np.zeros(f.shape, bool)
    with np.errstate(all="ignore"):
        for _ in range(max_iter):
            done |= np.sqrt(_dot64(g, g)) < tolerance_grad
            if done.all():
                break
            p = (-_direction64(g, S, Yh, rho, eps)).astype(F32)
            dphi0 = _dot64(g, p)
            if line_search_fn == "strong_wolfe":
                alpha = _strong_wolfe64(cl, x, p, f, dphi0, line_search_max_iter)
            else:
                alpha = _backtracking64(cl, x, p, f, dphi0, line_search_max_iter)
            s = (alpha[None, :] * p).astype(F32)
            x = (x + s * (~done)[None, :].astype(F32)).astype(F32)
            f_n, g_n = cl(x)
            y = (g_n - g).astype(F32)
            S.append(s)
            Yh.append(y)
            rho.append(1.0 / (_dot64(y, s) + eps))
            if len(S) > history_size:
                S.pop(0)
                Yh.pop(0)
                rho.pop(0)
            g, f = g_n, f_n
    return x, f


def optimize(X, Y, ops_bm, ch_bm, C_init_bm, max_iter=100, history_size=10, tolerance_grad=1e-7,
             line_search_fn="strong_wolfe", line_search_max_iter=10, eps=1e-10, noise=None, precision="fp32"):
    """optimize (symbolic_torch/optimization.py:9-112) with the C oracle's
    closure (noise: see _Closure).  precision "fp32": the reference's
    arithmetic (lbfgs_run); "fp64": the device's (lbfgs_run64).  Returns
    (C_final (B, M), Y_final (B, N), final loss (B,), closure calls)."""
    X = np.ascontiguousarray(X, F32)
    Y = np.ascontiguousarray(Y, F32)
    mask = (np.asarray(ops_bm) == 1).T
    fp64 = precision == "fp64"
    cl = _Closure(X, Y, np.ascontiguousarray(ops_bm), np.ascontiguousarray(ch_bm), mask, noise, fp64=fp64)
    x0 = np.ascontiguousarray(np.asarray(C_init_bm, F32).T)
    run = lbfgs_run64 if fp64 else lbfgs_run
    x, f = run(cl, x0, max_iter, history_size, tolerance_grad, line_search_fn, line_search_max_iter, eps)
    C_final = x.T.copy()
    y = oracle.fused_root(X, np.ascontiguousarray(ops_bm), np.ascontiguousarray(ch_bm), C_final, Y, mode=1,
                          want_gx=False)["y"]
    return C_final, y, f, cl.calls
