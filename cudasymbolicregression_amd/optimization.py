"""Constant fitting (reference symbolic_torch/optimization.py:9-112).

`optimize` fits the LEARNABLE_CONSTANT slots of every expression to its
targets by minimising the per-expression mean squared error, like the
reference's closure + LBFGS_batch.  The whole optimisation of a problem runs
inside one device-resident L-BFGS kernel (libsr_hip.so, `sr_lbfgs_fit`): one
wavefront per problem, closures evaluated by the fused root-only
forward/backward, FP64 losses, inner products and Wolfe tests.
"""
from __future__ import annotations

import ctypes
from typing import Dict, Literal, Optional, Tuple, Union

import torch

from . import _lib, autotune
from ._C import _require_gpu, get_error_message
from .fused import fused_root, ragged_offsets

_LINE_SEARCH = {"strong_wolfe": 0, "backtracking": 1}


def _fit(X, Y, Ops, Ch, C_init, max_iter, history_size, tolerance_grad, line_search_fn, stop_loss=None,
         x_off=None, y_off=None, n_pts=None, line_search_max_iter: int = 10, eps: float = 1e-10,
         config=None, stationary_exit: bool = True, starts: int = 0):
    """One sr_lbfgs_fit call (starts = 0), or one sr_lbfgs_fit_multi call
    with C_init (B, starts, M) and per-fit outputs (B, starts, ...).
    config: launch variant per tape-row bucket (None: the autotuner's entry
    for this workload, else the defaults)."""
    dev = Ops.device
    B, M = Ops.shape
    K = max(int(starts), 1)
    lead = (B, K) if starts else (B,)
    if line_search_fn not in _LINE_SEARCH:
        raise ValueError(f"Invalid line_search_fn: {line_search_fn}. Choose 'strong_wolfe' or 'backtracking'.")
    if not 1 <= int(history_size) <= 16:
        raise ValueError("history_size must be between 1 and 16 for the device L-BFGS")
    if M > 128:
        raise RuntimeError("the gfx950 evaluator supports at most 128 nodes per expression")
    if tuple(C_init.shape) != lead + (M,):
        raise RuntimeError(f"C_init must have shape {lead + (M,)}")
    C_out = torch.empty(lead + (M,), dtype=torch.float32, device=dev)
    loss = torch.empty(lead, dtype=torch.float64, device=dev)
    iters = torch.empty(lead, dtype=torch.int32, device=dev)
    nclos = torch.empty(lead, dtype=torch.int32, device=dev)
    err = torch.empty(lead, dtype=torch.int32, device=dev)
    flags = torch.empty(2, dtype=torch.int32, device=dev)
    lib = _lib.hip()
    ws = torch.empty(max(int(lib.sr_lbfgs_workspace_bytes(B, M)), 256), dtype=torch.uint8, device=dev)
    N = X.shape[0]
    n_x = X.shape[1]
    if config is None:
        config = autotune.lookup(dev, M, N, n_x, history_size, line_search_fn)
    opts = autotune.make_options(config, stationary_exit)
    head = (_lib.ptr(X), _lib.ptr(Y), N, n_x, _lib.ptr(Ops), _lib.ptr(Ch), _lib.ptr(C_init), M, B)
    tail = (_lib.ptr(x_off), _lib.ptr(y_off), _lib.ptr(n_pts), _lib.ptr(stop_loss),
            _lib.ptr(C_out), _lib.ptr(loss), _lib.ptr(iters), _lib.ptr(nclos), _lib.ptr(err), _lib.ptr(flags),
            int(max_iter), int(history_size), float(tolerance_grad), _LINE_SEARCH[line_search_fn],
            int(line_search_max_iter), float(eps), ctypes.byref(opts), _lib.ptr(ws), ws.numel(),
            _lib.stream_of(dev))
    if starts:
        _lib.check_status(lib.sr_lbfgs_fit_multi(*head, K, *tail), "sr_lbfgs_fit_multi")
    else:
        _lib.check_status(lib.sr_lbfgs_fit(*head, *tail), "sr_lbfgs_fit")
    return C_out, loss, iters, err, flags, nclos


def optimize(
    X: torch.Tensor,
    Y: torch.Tensor,
    Ops: torch.Tensor,
    Ch: torch.Tensor,
    C_init: torch.Tensor,
    max_iter: int = 100,
    history_size: int = 10,
    tolerance_grad: float = 1e-7,
    line_search_fn: Literal["strong_wolfe", "backtracking"] = "strong_wolfe",
    *,
    stop_rel_loss: Optional[float] = None,
    raise_errors: bool = True,
    return_info: bool = False,
) -> Union[Tuple[torch.Tensor, torch.Tensor], Tuple[torch.Tensor, torch.Tensor, Dict[str, torch.Tensor]]]:
    """Optimises the constants of a batch of expressions with batched L-BFGS.

    Args (as in the reference):
        X (N, n_x), Y (B, N), Ops (B, M), Ch (B, M, 2), C_init (B, M);
        max_iter, history_size, tolerance_grad, line_search_fn.
    Extensions (keyword-only):
        stop_rel_loss: also stop a problem once its loss <= stop_rel_loss * Var(Y_b).
        raise_errors: raise like the reference on validation / backward
            errors (201, 204); if False, failing problems are reported in
            info["err"] and the others are still fitted.
        return_info: also return dict(loss (B,) FP64, iters (B,), err (B,)).

    Returns:
        C_final (B, M) and Y_final (B, N): the fitted constants (non-constant
        slots keep C_init's values) and each expression's prediction.
    """
    _require_gpu(X, "X")
    dev = X.device
    if X.dtype != torch.float32:
        raise RuntimeError("optimize fits FP32 constants: X must be a float32 tensor")
    for t, name in ((Y, "Y"), (Ops, "Ops"), (Ch, "Ch"), (C_init, "C_init")):
        if t.device != dev:
            raise RuntimeError("All input tensors must be on the same device.")
    B, M = Ops.shape
    Xc = X.contiguous()
    Yc = Y.to(torch.float32).contiguous()
    Opsc = Ops.to(torch.int64).contiguous()
    Chc = Ch.to(torch.int64).contiguous()
    Cc = C_init.to(torch.float32).contiguous()
    if Yc.shape != (B, X.shape[0]):
        raise RuntimeError("Y must have shape (B, N)")
    stop = None
    if stop_rel_loss is not None:
        var = Yc.double().var(dim=1, unbiased=False)
        var = torch.where(var > 0, var, (Yc.double() ** 2).mean(dim=1))
        stop = (float(stop_rel_loss) * var).contiguous()
    C_out, loss, iters, err, flags, nclos = _fit(Xc, Yc, Opsc, Chc, Cc, max_iter, history_size, tolerance_grad,
                                          line_search_fn, stop_loss=stop)
    if raise_errors:
        f = flags.tolist()
        if f[0]:
            raise RuntimeError(get_error_message(f[0], "input validation"))
        if f[1]:
            raise RuntimeError(get_error_message(f[1], "symbolic evaluation backward pass"))
    Y_final = fused_root(Xc, Opsc, Chc, C_out, None, mode=0, fwd_only=True, raise_errors=False)["y"]
    if return_info:
        return C_out, Y_final, dict(loss=loss, iters=iters, err=err, closures=nclos)
    return C_out, Y_final


def optimize_ragged(
    X: torch.Tensor,
    Y: torch.Tensor,
    n_pts: torch.Tensor,
    Ops: torch.Tensor,
    Ch: torch.Tensor,
    C_init: torch.Tensor,
    max_iter: int = 100,
    history_size: int = 10,
    tolerance_grad: float = 1e-7,
    line_search_fn: Literal["strong_wolfe", "backtracking"] = "strong_wolfe",
    *,
    stop_rel_loss: Optional[float] = None,
    raise_errors: bool = True,
) -> Tuple[torch.Tensor, Dict[str, torch.Tensor]]:
    """`optimize` for problems that each have their own dataset (no
    counterpart in the reference, whose optimize shares one X across the
    batch, symbolic_torch/optimization.py:9-112).

    Args:
        X: (sum_b N_b, n_x) float32, the points of problem 0, then problem 1, ...
        Y: (sum_b N_b,) float32, the targets in the same order.
        n_pts: (B,) integer N_b >= 1, the number of points of each problem.
        Ops (B, M), Ch (B, M, 2), C_init (B, M) and the options as in optimize;
        each problem's loss is the mean over its own N_b points.

    Returns:
        C_final (B, M) and info = dict(loss (B,) FP64, iters (B,), err (B,),
        closures (B,)).  One launch fits the whole batch.
    """
    _require_gpu(X, "X")
    dev = X.device
    if X.dtype != torch.float32 or X.dim() != 2:
        raise RuntimeError("X must be a (sum N_b, n_x) float32 tensor")
    B, M = Ops.shape
    n = n_pts.to(device=dev, dtype=torch.int64)
    if n.shape != (B,):
        raise RuntimeError("n_pts must have shape (B,)")
    if B and int(n.min()) < 1:
        raise RuntimeError("every problem needs at least one point")
    total = int(n.sum()) if B else 0
    if X.shape[0] != total or Y.numel() != total:
        raise RuntimeError(f"X has {X.shape[0]} rows and Y {Y.numel()} targets; n_pts sums to {total}")
    n_x = X.shape[1]
    x_off, y_off, n32 = ragged_offsets(n, n_x)
    Xc = X.contiguous()
    Yc = Y.reshape(-1).to(torch.float32).contiguous()
    stop = None
    if stop_rel_loss is not None:
        seg = torch.repeat_interleave(torch.arange(B, device=dev), n)
        Yd = Yc.double()
        mean = torch.zeros(B, dtype=torch.float64, device=dev).index_add_(0, seg, Yd) / n
        var = torch.zeros(B, dtype=torch.float64, device=dev).index_add_(0, seg, (Yd - mean[seg]) ** 2) / n
        msq = torch.zeros(B, dtype=torch.float64, device=dev).index_add_(0, seg, Yd * Yd) / n
        stop = (float(stop_rel_loss) * torch.where(var > 0, var, msq)).contiguous()
    C_out, loss, iters, err, flags, nclos = _fit(
        Xc, Yc, Ops.to(device=dev, dtype=torch.int64).contiguous(), Ch.to(device=dev, dtype=torch.int64).contiguous(),
        C_init.to(device=dev, dtype=torch.float32).contiguous(), max_iter, history_size, tolerance_grad,
        line_search_fn, stop_loss=stop, x_off=x_off, y_off=y_off, n_pts=n32)
    if raise_errors:
        f = flags.tolist()
        if f[0]:
            raise RuntimeError(get_error_message(f[0], "input validation"))
        if f[1]:
            raise RuntimeError(get_error_message(f[1], "symbolic evaluation backward pass"))
    return C_out, dict(loss=loss, iters=iters, err=err, closures=nclos)


def optimize_multistart(
    X: torch.Tensor,
    Y: torch.Tensor,
    Ops: torch.Tensor,
    Ch: torch.Tensor,
    C_init: torch.Tensor,
    max_iter: int = 100,
    history_size: int = 10,
    tolerance_grad: float = 1e-7,
    line_search_fn: Literal["strong_wolfe", "backtracking"] = "strong_wolfe",
    *,
    stop_rel_loss: Optional[float] = None,
    raise_errors: bool = True,
) -> Tuple[torch.Tensor, torch.Tensor, Dict[str, torch.Tensor]]:
    """K fits per expression from K start points, in one launch.

    The multi-start counterpart of the reference's multiple-constant
    evaluation (evaluate_multiple_constant_backend,
    src/symbolic_evaluation.cpp:171-305, K constant sets per expression)
    carried through to the fit: each expression is fitted from each of its K
    starts exactly as optimize fits it from that start, the K fits of one
    expression sharing its program, targets and points.

    Args:
        X (N, n_x), Y (B, N), Ops (B, M), Ch (B, M, 2) and the options as in
        optimize; C_init (B, K, M) the K start points of each expression.

    Returns:
        C_best (B, M): per expression the fit with the lowest final loss
        (non-finite losses last, the first start on ties); Y_best (B, N) its
        prediction; info = dict(C (B, K, M) every fit, loss (B, K) FP64,
        iters, err, closures (B, K), best (B,) the chosen start).
    """
    _require_gpu(X, "X")
    dev = X.device
    if X.dtype != torch.float32:
        raise RuntimeError("optimize fits FP32 constants: X must be a float32 tensor")
    for t in (Y, Ops, Ch, C_init):
        if t.device != dev:
            raise RuntimeError("All input tensors must be on the same device.")
    B, M = Ops.shape
    if C_init.dim() != 3 or C_init.shape[0] != B or C_init.shape[2] != M or C_init.shape[1] < 1:
        raise RuntimeError("C_init must have shape (B, K, M) with K >= 1")
    K = C_init.shape[1]
    Xc = X.contiguous()
    Yc = Y.to(torch.float32).contiguous()
    if Yc.shape != (B, X.shape[0]):
        raise RuntimeError("Y must have shape (B, N)")
    Opsc = Ops.to(torch.int64).contiguous()
    Chc = Ch.to(torch.int64).contiguous()
    stop = None
    if stop_rel_loss is not None:
        var = Yc.double().var(dim=1, unbiased=False)
        var = torch.where(var > 0, var, (Yc.double() ** 2).mean(dim=1))
        stop = (float(stop_rel_loss) * var).contiguous()
    C_all, loss, iters, err, flags, nclos = _fit(Xc, Yc, Opsc, Chc, C_init.to(torch.float32).contiguous(), max_iter,
                                                 history_size, tolerance_grad, line_search_fn, stop_loss=stop,
                                                 starts=K)
    if raise_errors:
        f = flags.tolist()
        if f[0]:
            raise RuntimeError(get_error_message(f[0], "input validation"))
        if f[1]:
            raise RuntimeError(get_error_message(f[1], "symbolic evaluation backward pass"))
    key = torch.where(torch.isfinite(loss) & (err == 0), loss, torch.full_like(loss, float("inf")))
    best = key.argmin(dim=1)  # first minimum on ties
    C_best = C_all.gather(1, best.view(B, 1, 1).expand(B, 1, M)).squeeze(1).contiguous()
    Y_best = fused_root(Xc, Opsc, Chc, C_best, None, mode=0, fwd_only=True, raise_errors=False)["y"]
    return C_best, Y_best, dict(C=C_all, loss=loss, iters=iters, err=err, closures=nclos, best=best)
