"""Batched L-BFGS over independent problems for arbitrary torch closures.

Same constructor, options, errors and per-problem results as the reference's
symbolic_torch/lbfgs_batch.py:7-326 (two-loop recursion, bracketing strong
Wolfe search with a bisection zoom, Armijo backtracking, per-problem `done`
mask).  Parameters have shape (n_variables, n_problems).

This class serves closures that are not symbolic expressions (tests use the
Rosenbrock function).  `optimize` does not use it: it fits expression
constants with the device-resident L-BFGS kernel instead.
"""
from __future__ import annotations

from typing import Callable, Iterable, List, Literal

import torch
from torch import Tensor
from torch.optim.optimizer import Optimizer


class LBFGS_batch(Optimizer):
    """Batch-wise L-BFGS on a single (n_variables, n_problems) parameter."""

    def __init__(
        self,
        params: Iterable[Tensor],
        max_iter: int = 100,
        history_size: int = 10,
        tolerance_grad: float = 1e-7,
        line_search_fn: Literal["strong_wolfe", "backtracking"] = "strong_wolfe",
        line_search_max_iter: int = 10,
        eps: float = 1e-10,
    ):
        defaults = dict(
            max_iter=max_iter,
            history_size=history_size,
            tolerance_grad=tolerance_grad,
            line_search_fn=line_search_fn,
            line_search_max_iter=line_search_max_iter,
            eps=eps,
        )
        super().__init__(params, defaults)
        if len(self.param_groups) != 1:
            raise ValueError("LBFGS_batch doesn't support multiple param groups.")
        group = self.param_groups[0]
        if len(group["params"]) != 1:
            raise ValueError("LBFGS_batch is designed to work on a single parameter tensor.")
        if group["params"][0].dim() != 2:
            raise ValueError("Input params must be a 2D tensor of shape (n_variables, n_problems)")
        if line_search_fn not in ("strong_wolfe", "backtracking"):
            raise ValueError(f"Invalid line_search_fn: {line_search_fn}. Choose 'strong_wolfe' or 'backtracking'.")

    # -------------------------------------------------------------- helpers
    @property
    def _param(self) -> Tensor:
        return self.param_groups[0]["params"][0]

    @staticmethod
    def _dot(a: Tensor, b: Tensor) -> Tensor:
        """Per-problem inner product: (D, B), (D, B) -> (1, B)."""
        return (a * b).sum(dim=0, keepdim=True)

    def _eval(self, closure: Callable[[], Tensor], point: Tensor) -> Tensor:
        self._param.data.copy_(point)
        with torch.enable_grad():
            return closure()

    def _direction(self, g: Tensor, S: List[Tensor], Yh: List[Tensor], eps: float) -> Tensor:
        """Two-loop recursion; returns H*g (no scaling before the first pair)."""
        if not S:
            return g
        q = g
        alphas: List[Tensor] = [None] * len(S)  # type: ignore[list-item]
        for i in reversed(range(len(S))):
            rho = 1.0 / (self._dot(Yh[i], S[i]) + eps)
            alphas[i] = rho * self._dot(S[i], q)
            q = q - alphas[i] * Yh[i]
        gamma = self._dot(S[-1], Yh[-1]) / (self._dot(Yh[-1], Yh[-1]) + eps)
        z = gamma * q
        for i in range(len(S)):
            rho = 1.0 / (self._dot(Yh[i], S[i]) + eps)
            beta = rho * self._dot(Yh[i], z)
            z = z + S[i] * (alphas[i] - beta)
        return z

    def _backtracking(self, closure, x: Tensor, p: Tensor, f: Tensor, g: Tensor, max_ls: int) -> Tensor:
        """Armijo backtracking: halve alpha per problem until f decreases enough."""
        slope = self._dot(g, p).squeeze(0)
        alpha = torch.ones_like(f, device=x.device)
        for _ in range(max_ls):
            f_try = self._eval(closure, x + alpha.unsqueeze(0) * p)
            ok = f_try <= f + 1e-4 * alpha * slope
            if ok.all():
                break
            alpha = torch.where(ok, alpha, alpha * 0.5)
        self._param.data.copy_(x)
        return alpha.unsqueeze(0)

    def _strong_wolfe(self, closure, x: Tensor, p: Tensor, max_ls: int, alpha_max: float = 20,
                      c1: float = 1e-4, c2: float = 0.9) -> Tensor:
        """Bracketing phase; problems that bracket a minimum are zoomed."""
        phi0 = self._eval(closure, x)
        dphi0 = self._dot(self._param.grad, p).squeeze(0)
        a_prev = torch.zeros_like(phi0)
        phi_prev = phi0.clone()
        a = torch.ones_like(phi0)
        chosen = torch.zeros_like(phi0)
        settled = torch.zeros_like(phi0, dtype=torch.bool)
        for it in range(max_ls):
            if torch.all(settled):
                break
            phi = self._eval(closure, x + a.unsqueeze(0) * p)
            dphi = self._dot(self._param.grad, p).squeeze(0)
            too_high = (phi > phi0 + c1 * a * dphi0) | ((it > 0) & (phi >= phi_prev))
            flat_enough = torch.abs(dphi) <= c2 * torch.abs(dphi0)
            uphill = dphi >= 0
            accept = ~settled & flat_enough
            chosen = torch.where(accept, a, chosen)
            settled |= accept
            zoom = ~settled & (too_high | uphill)
            if torch.any(zoom):
                lo = torch.where(too_high, a_prev, a)
                hi = torch.where(too_high, a, a_prev)
                chosen = torch.where(zoom, self._zoom(closure, x, p, lo, hi, zoom, c1, c2), chosen)
                settled |= zoom
            going = ~settled
            a_prev = torch.where(going, a, a_prev)
            phi_prev = torch.where(going, phi, phi_prev)
            a = torch.where(going, a + (alpha_max - a) * 0.8, a)
        chosen = torch.where(settled, chosen, a)
        self._param.data.copy_(x)
        return chosen.unsqueeze(0)

    def _zoom(self, closure, x: Tensor, p: Tensor, lo: Tensor, hi: Tensor, active: Tensor, c1: float,
              c2: float) -> Tensor:
        """Bisection of [lo, hi] for the problems in `active` (at most 10 steps)."""
        phi0 = self._eval(closure, x)
        dphi0 = self._dot(self._param.grad, p).squeeze(0)
        phi_lo = self._eval(closure, x + lo.unsqueeze(0) * p)
        out = torch.zeros_like(lo)
        fin = torch.zeros_like(lo, dtype=torch.bool)
        for _ in range(10):
            if torch.all(fin[active]):
                break
            mid = (lo + hi) / 2
            phi = self._eval(closure, x + mid.unsqueeze(0) * p)
            dphi = self._dot(self._param.grad, p).squeeze(0)
            too_high = (phi > phi0 + c1 * mid * dphi0) | (phi >= phi_lo)
            flat_enough = abs(dphi) <= c2 * abs(dphi0)
            take = active & ~fin & flat_enough
            out = torch.where(take, mid, out)
            fin |= take
            live = active & ~fin
            new_hi = torch.where(live & too_high, mid, hi)
            flip = dphi * (new_hi - lo) >= 0
            hi = torch.where(live & ~too_high & flip, lo, new_hi)
            move_lo = live & ~too_high
            lo = torch.where(move_lo, mid, lo)
            phi_lo = torch.where(move_lo, phi, phi_lo)
            tiny = abs(hi - lo) < 1e-9
            stop = active & ~fin & tiny
            out = torch.where(stop, mid, out)
            fin |= stop
        return torch.where(active & ~fin, lo, out)

    # ----------------------------------------------------------------- step
    def step(self, closure: Callable[[], Tensor]):  # type: ignore[override]
        with torch.no_grad():
            group = self.param_groups[0]
            x_param = self._param
            with torch.enable_grad():
                f = closure()
            g = x_param.grad.clone()
            x = x_param.data.clone()
            S: List[Tensor] = []
            Yh: List[Tensor] = []
            done = torch.zeros(f.shape[0], dtype=torch.bool, device=x.device)
            for _ in range(group["max_iter"]):
                done |= torch.linalg.norm(g, dim=0) < group["tolerance_grad"]
                if done.all():
                    break
                p = -self._direction(g, S, Yh, group["eps"])
                if group["line_search_fn"] == "strong_wolfe":
                    alpha = self._strong_wolfe(closure, x, p, group["line_search_max_iter"])
                else:
                    alpha = self._backtracking(closure, x, p, f, g, group["line_search_max_iter"])
                s = alpha * p
                x += s * (~done).unsqueeze(0).float()
                f_new = self._eval(closure, x)
                g_new = x_param.grad.clone()
                S.append(s)
                Yh.append(g_new - g)
                if len(S) > group["history_size"]:
                    S.pop(0)
                    Yh.pop(0)
                g = g_new
                f = f_new
            x_param.data.copy_(x)
            return f
