"""Root-only fused evaluation (forward + reverse sweep in one pass).

New entry point with no counterpart in the reference API: it computes what
`optimize`'s closure needs (symbolic_torch/optimization.py:72-99) without
materialising the (M, N, B) cache of `evaluate`:

    y      (B, N)  value of each expression's root node
           (root = #non-NO_OP - 1, clamped at 0, as in optimization.py:55-58)
    loss   (B,)    mean_n (y - Y)^2 in FP64                 [mode 1]
    grad_C (B, M)  dL/dC summed over the N points in FP64
    grad_X (N, n_x) dL/dX summed over the batch            [optional]
    err    (B,)    per-expression code: 101/102 (validation), 201/204 (backward)

mode 0: `dy` is the upstream gradient of y.  mode 1: `dy` is the target Y
and the upstream gradient is the MSE closure's (2 * (y - Y)) * (1 / N).
Inputs are batch-first: X (N, n_x), Ops (B, M), Ch (B, M, 2), C (B, M).
"""
from __future__ import annotations

import ctypes
from typing import Dict, Optional, Sequence

import torch

from . import _lib, autotune
from ._C import _require_gpu, get_error_message


def fused_root(
    X: torch.Tensor,
    Ops: torch.Tensor,
    Ch: torch.Tensor,
    C: torch.Tensor,
    dy: Optional[torch.Tensor],
    mode: int = 0,
    want_grad_X: bool = False,
    fwd_only: bool = False,
    raise_errors: bool = True,
    out: Optional[Dict[str, torch.Tensor]] = None,
    config: Optional[Sequence[int]] = None,
) -> Dict[str, torch.Tensor]:
    """config: launch variant per tape-row bucket (autotune.fused_variants);
    None = the tuned configuration of this workload if `autotune.tune_fused`
    stored one, else the built-in defaults.  Every variant gives the same y
    and error codes bit for bit; loss, grad_C and grad_X differ only by the
    order of their sums."""
    _require_gpu(X, "X")
    dev = X.device
    dt = X.dtype
    if dt not in (torch.float32, torch.float64):
        raise RuntimeError("X must be a floating-point tensor (float or double)")
    B, M = Ops.shape
    N, n_x = X.shape
    if Ch.shape != (B, M, 2) or C.shape != (B, M):
        raise RuntimeError("Ops (B,M), Ch (B,M,2) and C (B,M) shapes disagree")
    if M > 128:
        raise RuntimeError("the gfx950 evaluator supports at most 128 nodes per expression")
    X = X.contiguous()
    Ops = Ops.to(device=dev, dtype=torch.int64).contiguous()
    Ch = Ch.to(device=dev, dtype=torch.int64).contiguous()
    C = C.to(device=dev, dtype=dt).contiguous()
    if dy is None:
        if not (fwd_only and mode == 0):
            raise RuntimeError("dy (upstream gradient or target) is required")
        dyc = None
    else:
        dyc = dy.to(device=dev, dtype=dt).contiguous()
        if dyc.shape != (B, N):
            raise RuntimeError("dy must have shape (B, N)")
    o = {} if out is None else out
    y = o.get("y")
    if y is None:
        y = torch.empty((B, N), dtype=dt, device=dev)
    loss = o.get("loss") if mode == 1 else None
    if mode == 1 and loss is None:
        loss = torch.empty(B, dtype=torch.float64, device=dev)
    gC = None
    if not fwd_only:
        gC = o.get("grad_C")
        if gC is None:
            gC = torch.empty((B, M), dtype=dt, device=dev)
    gX = torch.empty((N, n_x), dtype=dt, device=dev) if (want_grad_X and not fwd_only) else None
    err = o.get("err")
    if err is None:
        err = torch.empty(B, dtype=torch.int32, device=dev)
    flags = torch.empty(2, dtype=torch.int32, device=dev)
    lib = _lib.hip()
    is64 = dt == torch.float64
    ws_bytes = lib.sr_fused_workspace_bytes(N, n_x, B, M, int(gX is not None), int(is64))
    ws = o.get("_ws")
    if ws is None or ws.numel() < ws_bytes or ws.device != dev:
        ws = torch.empty(max(ws_bytes, 256), dtype=torch.uint8, device=dev)
        if out is not None:
            o["_ws"] = ws
    if config is None:
        config = autotune.fused_lookup(dev, dt, M, N, n_x, gX is not None, bool(fwd_only))
    opts = autotune.make_fused_options(config, is64) if config is not None else None
    fn = lib.sr_fused_root_opts_f64 if is64 else lib.sr_fused_root_opts_f32
    _lib.check_status(
        fn(_lib.ptr(X), N, n_x, _lib.ptr(Ops), _lib.ptr(Ch), _lib.ptr(C), M, B, _lib.ptr(dyc), int(mode),
           int(bool(fwd_only)), _lib.ptr(y), _lib.ptr(loss), _lib.ptr(gC), _lib.ptr(gX), _lib.ptr(err),
           _lib.ptr(flags), ctypes.byref(opts) if opts is not None else None, _lib.ptr(ws), ws.numel(),
           _lib.stream_of(dev)),
        "fused_root",
    )
    res = dict(y=y, err=err, flags=flags)
    if loss is not None:
        res["loss"] = loss
    if gC is not None:
        res["grad_C"] = gC
    if gX is not None:
        res["grad_X"] = gX
    if raise_errors:
        f = flags.tolist()
        if f[0]:
            raise RuntimeError(get_error_message(f[0], "input validation"))
        if f[1]:
            raise RuntimeError(get_error_message(f[1], "symbolic evaluation backward pass"))
    return res


def ragged_offsets(n_pts: torch.Tensor, n_x: int):
    """(x_off, y_off, n32) for problems stored back to back: problem b's rows
    start at sum_{j<b} n_pts[j]."""
    n = n_pts.to(torch.int64)
    start = torch.cumsum(n, 0) - n
    return (start * n_x).contiguous(), start.contiguous(), n.to(torch.int32).contiguous()


def fused_root_ragged(
    X: torch.Tensor,
    n_pts: torch.Tensor,
    Ops: torch.Tensor,
    Ch: torch.Tensor,
    C: torch.Tensor,
    dy: Optional[torch.Tensor],
    mode: int = 0,
    fwd_only: bool = False,
    raise_errors: bool = True,
) -> Dict[str, torch.Tensor]:
    """`fused_root` for expressions that each have their own points: X is
    (sum_b N_b, n_x) with problem b's rows after those of problems < b, and
    dy (mode 0: upstream gradient; mode 1: targets) and the returned y are
    (sum_b N_b,) in the same order.  loss (mode 1) is each problem's mean over
    its own points.  No grad_X (there is no shared X)."""
    _require_gpu(X, "X")
    dev, dt = X.device, X.dtype
    if dt not in (torch.float32, torch.float64) or X.dim() != 2:
        raise RuntimeError("X must be a (sum N_b, n_x) floating-point tensor")
    B, M = Ops.shape
    n_x = X.shape[1]
    n = n_pts.to(device=dev)
    if n.shape != (B,):
        raise RuntimeError("n_pts must have shape (B,)")
    total = X.shape[0]
    if B and (int(n.min()) < 1 or int(n.sum()) != total):
        raise RuntimeError(f"n_pts must be >= 1 and sum to the {total} rows of X")
    x_off, y_off, n32 = ragged_offsets(n, n_x)
    Nmax = int(n.max()) if B else 1
    X = X.contiguous()
    Ops = Ops.to(device=dev, dtype=torch.int64).contiguous()
    Ch = Ch.to(device=dev, dtype=torch.int64).contiguous()
    C = C.to(device=dev, dtype=dt).contiguous()
    if dy is None:
        if not (fwd_only and mode == 0):
            raise RuntimeError("dy (upstream gradient or target) is required")
        dyc = None
    else:
        dyc = dy.reshape(-1).to(device=dev, dtype=dt).contiguous()
        if dyc.numel() != total:
            raise RuntimeError("dy must hold one value per point")
    y = torch.empty(total, dtype=dt, device=dev)
    loss = torch.empty(B, dtype=torch.float64, device=dev) if mode == 1 else None
    gC = None if fwd_only else torch.empty((B, M), dtype=dt, device=dev)
    err = torch.empty(B, dtype=torch.int32, device=dev)
    flags = torch.empty(2, dtype=torch.int32, device=dev)
    lib = _lib.hip()
    is64 = dt == torch.float64
    ws_bytes = lib.sr_fused_workspace_bytes(Nmax, n_x, B, M, 0, int(is64))
    ws = torch.empty(max(ws_bytes, 256), dtype=torch.uint8, device=dev)
    fn = lib.sr_fused_root_ragged_f64 if is64 else lib.sr_fused_root_ragged_f32
    _lib.check_status(
        fn(_lib.ptr(X), Nmax, n_x, _lib.ptr(x_off), _lib.ptr(y_off), _lib.ptr(n32), _lib.ptr(Ops), _lib.ptr(Ch),
           _lib.ptr(C), M, B, _lib.ptr(dyc), int(mode), int(bool(fwd_only)), _lib.ptr(y), _lib.ptr(loss),
           _lib.ptr(gC), _lib.ptr(err), _lib.ptr(flags), _lib.ptr(ws), ws.numel(), _lib.stream_of(dev)),
        "fused_root_ragged",
    )
    res = dict(y=y, err=err, flags=flags)
    if loss is not None:
        res["loss"] = loss
    if gC is not None:
        res["grad_C"] = gC
    if raise_errors:
        f = flags.tolist()
        if f[0]:
            raise RuntimeError(get_error_message(f[0], "input validation"))
        if f[1]:
            raise RuntimeError(get_error_message(f[1], "symbolic evaluation backward pass"))
    return res
