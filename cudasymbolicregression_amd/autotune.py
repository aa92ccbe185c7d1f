"""Autotuner for the launch variants of the fit and evaluator kernels.

The device L-BFGS (`optimize`) buckets problems by tape rows and runs one
kernel per bucket; each bucket has a few launch variants that trade
per-problem latency against throughput: W waves cooperating on one problem
(points split between them) and PPT points per lane per instruction
dispatch.  Which variant is fastest depends on the workload (program
lengths, points, iteration counts), so `tune_fit` times them on a
representative batch with coordinate descent over the buckets.

The chosen configuration is stored in a workload-keyed cache (device name,
padded length, point count, variables, history, line search); `optimize`
looks its own workload up there and otherwise runs the built-in defaults.
The configuration is a per-call option of the C-ABI (`sr_fit_options`), so
the library itself holds no tuning state and calls on several threads do
not interfere.  The cache is guarded by a lock and can be saved to / loaded
from a JSON file.

The fused evaluator (`fused_root`) has launch variants of its own per
bucket (waves per block x points per lane, `sr_fused_options`); `tune_fused`
times them the same way and stores the winner under an evaluator key
(device, precision, padded length, points, variables, grad_X, forward
only), which `fused_root` looks up.
"""
from __future__ import annotations

import ctypes
import json
import math
import threading
import time
from typing import Callable, Dict, List, Optional, Sequence, Tuple

import torch

from . import _lib

N_BUCKETS = 5  # SR_FIT_BUCKETS of include/sr_eval.h


class FusedOptions(ctypes.Structure):
    """`sr_fused_options` of include/sr_eval.h."""

    _fields_ = [("variant", ctypes.c_int * N_BUCKETS)]


class FitOptions(ctypes.Structure):
    """`sr_fit_options` of include/sr_eval.h."""

    _fields_ = [("variant", ctypes.c_int * N_BUCKETS), ("stationary_exit", ctypes.c_int)]


def make_options(config: Optional[Sequence[int]] = None, stationary_exit: bool = True) -> FitOptions:
    """Options for one sr_lbfgs_fit call (config None: the defaults)."""
    o = FitOptions()
    cfg = list(config) if config is not None else [-1] * N_BUCKETS
    if len(cfg) != N_BUCKETS:
        raise ValueError(f"a fit configuration names {N_BUCKETS} variants, got {len(cfg)}")
    for b, v in enumerate(cfg):
        o.variant[b] = int(v)
    o.stationary_exit = 1 if stationary_exit else 0
    return o


def variants() -> List[List[Dict[str, int]]]:
    """Per bucket, the available variants as dicts {waves, ppt}."""
    lib = _lib.hip()
    out = []
    for b in range(lib.sr_fit_num_buckets()):
        row = []
        for v in range(lib.sr_fit_num_variants(b)):
            w, p = ctypes.c_int(), ctypes.c_int()
            lib.sr_fit_variant_info(b, v, ctypes.byref(w), ctypes.byref(p))
            row.append(dict(waves=w.value, ppt=p.value))
        out.append(row)
    return out


def default_config() -> List[int]:
    lib = _lib.hip()
    return [int(lib.sr_fit_default_variant(b)) for b in range(lib.sr_fit_num_buckets())]


def validate_config(config: Sequence[int]) -> List[int]:
    n_var = [len(v) for v in variants()]
    cfg = [int(v) for v in config]
    if len(cfg) != len(n_var) or any(not (0 <= v < n) for v, n in zip(cfg, n_var)):
        raise ValueError(f"invalid fit configuration {cfg} (variants per bucket: {n_var})")
    return cfg


def fused_variants(is_f64: bool = False) -> List[List[Dict[str, int]]]:
    """Per bucket, the fused evaluator's variants as dicts {waves, ppt}
    (waves per block, points per lane)."""
    lib = _lib.hip()
    out = []
    for b in range(N_BUCKETS):
        row = []
        for v in range(lib.sr_fused_num_variants(b, int(is_f64))):
            w, p = ctypes.c_int(), ctypes.c_int()
            lib.sr_fused_variant_info(b, v, int(is_f64), ctypes.byref(w), ctypes.byref(p))
            row.append(dict(waves=w.value, ppt=p.value))
        out.append(row)
    return out


def validate_fused_config(config: Sequence[int], is_f64: bool = False) -> List[int]:
    n_var = [len(v) for v in fused_variants(is_f64)]
    cfg = [int(v) for v in config]
    if len(cfg) != N_BUCKETS or any(not (-1 <= v < n) for v, n in zip(cfg, n_var)):
        raise ValueError(f"invalid evaluator configuration {cfg} (variants per bucket: {n_var}; -1 = automatic)")
    return cfg


def make_fused_options(config: Sequence[int], is_f64: bool = False) -> FusedOptions:
    o = FusedOptions()
    for b, v in enumerate(validate_fused_config(config, is_f64)):
        o.variant[b] = v
    return o


# --------------------------------------------------------------- the cache
def workload_key(device: torch.device, M: int, N: int, n_x: int, history_size: int = 10,
                 line_search_fn: str = "strong_wolfe") -> Tuple:
    """Cache key of a fit workload: the device's name, the padded length, the
    point count rounded to a power of two, the variables, the history size
    and the line search."""
    device = torch.device(device)
    name = torch.cuda.get_device_name(device) if device.type == "cuda" else device.type
    n_bucket = 1 << max(0, int(math.ceil(math.log2(max(int(N), 1)))))
    return (name, int(M), n_bucket, int(n_x), int(history_size), str(line_search_fn))


def fused_workload_key(device: torch.device, dtype: torch.dtype, M: int, N: int, n_x: int, want_grad_X: bool,
                       fwd_only: bool) -> Tuple:
    """Cache key of an evaluator workload: "fused", the device's name, the
    precision, the padded length, the point count rounded to a power of two,
    the variables, and whether grad_X and the reverse sweep are wanted."""
    device = torch.device(device)
    name = torch.cuda.get_device_name(device) if device.type == "cuda" else device.type
    n_bucket = 1 << max(0, int(math.ceil(math.log2(max(int(N), 1)))))
    prec = "float64" if dtype == torch.float64 else "float32"
    return ("fused", name, prec, int(M), n_bucket, int(n_x), bool(want_grad_X), bool(fwd_only))


_cache_lock = threading.Lock()
_cache: Dict[Tuple, Dict[str, object]] = {}


def _validate_for(key: Tuple, config: Sequence[int]) -> List[int]:
    if len(key) and key[0] == "fused":
        return validate_fused_config(config, key[2] == "float64")
    return validate_config(config)


def cache_put(key: Tuple, config: Sequence[int], ms: Optional[float] = None) -> None:
    cfg = _validate_for(tuple(key), config)
    with _cache_lock:
        _cache[tuple(key)] = dict(config=cfg, ms=ms)


def cache_get(key: Tuple) -> Optional[List[int]]:
    with _cache_lock:
        e = _cache.get(tuple(key))
        return list(e["config"]) if e else None


def cache_clear() -> None:
    with _cache_lock:
        _cache.clear()


def cache_save(path: str) -> None:
    with _cache_lock:
        rows = [dict(key=list(k), **v) for k, v in _cache.items()]
    with open(path, "w") as f:
        json.dump(rows, f, indent=1)


def cache_load(path: str) -> int:
    """Adds the entries of a saved cache; returns how many were read."""
    with open(path) as f:
        rows = json.load(f)
    for r in rows:
        cache_put(tuple(r["key"]), r["config"], r.get("ms"))
    return len(rows)


def lookup(device: torch.device, M: int, N: int, n_x: int, history_size: int = 10,
           line_search_fn: str = "strong_wolfe") -> Optional[List[int]]:
    """The tuned configuration of this workload, or None (defaults)."""
    with _cache_lock:
        if not _cache:
            return None
    return cache_get(workload_key(device, M, N, n_x, history_size, line_search_fn))


def fused_lookup(device: torch.device, dtype: torch.dtype, M: int, N: int, n_x: int, want_grad_X: bool,
                 fwd_only: bool) -> Optional[List[int]]:
    """The tuned evaluator configuration of this workload, or None."""
    with _cache_lock:
        if not _cache:
            return None
    return cache_get(fused_workload_key(device, dtype, M, N, n_x, want_grad_X, fwd_only))


# ------------------------------------------------------------------ tuning
def _time(fn: Callable[[], object], reps: int) -> float:
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def tune_fit(X: torch.Tensor, Y: torch.Tensor, Ops: torch.Tensor, Ch: torch.Tensor, C_init: torch.Tensor,
             max_iter: int = 100, reps: int = 2, rounds: int = 1, verbose: bool = False, store: bool = True,
             log: Optional[List[str]] = None, **kw) -> Dict[str, object]:
    """Coordinate descent over the buckets' variants, timing the device fit
    on the given batch (keyword options: history_size, tolerance_grad,
    line_search_fn, stop_loss as in the fit).  Returns {config, ms, trials,
    key}; with store=True the fastest configuration goes into the workload
    cache, where `optimize` finds it.  Every trial is appended to `log`
    (and printed with verbose=True)."""
    from .optimization import _fit

    Xc, Yc = X.contiguous(), Y.to(torch.float32).contiguous()
    Oc, Cc = Ops.to(torch.int64).contiguous(), Ch.to(torch.int64).contiguous()
    C0 = C_init.to(torch.float32).contiguous()
    stop = kw.get("stop_loss")
    hist = kw.get("history_size", 10)
    ls = kw.get("line_search_fn", "strong_wolfe")

    def run_with(cfg):
        return lambda: _fit(Xc, Yc, Oc, Cc, C0, max_iter, hist, kw.get("tolerance_grad", 1e-7), ls,
                            stop_loss=stop, config=cfg)

    best = default_config()
    best_t = _time(run_with(best), reps)
    trials = [(list(best), best_t)]
    n_var = [len(v) for v in variants()]

    def note(cfg, t):
        line = f"[tune_fit] {cfg}: {t * 1e3:.3f} ms"
        if log is not None:
            log.append(line)
        if verbose:
            print(line, flush=True)

    note(best, best_t)
    for _ in range(rounds):
        for b in range(len(best)):
            start = best[b]  # already timed with the other buckets as they are now
            for v in range(n_var[b]):
                if v == start:
                    continue
                cfg = list(best)
                cfg[b] = v
                t = _time(run_with(cfg), reps)
                trials.append((cfg, t))
                note(cfg, t)
                if t < best_t:
                    best, best_t = cfg, t
    key = workload_key(X.device, Ops.shape[1], X.shape[0], X.shape[1], hist, ls)
    if store:
        cache_put(key, best, best_t * 1e3)
    return dict(config=best, ms=best_t * 1e3, trials=trials, key=key)


def tune_fused(X: torch.Tensor, Ops: torch.Tensor, Ch: torch.Tensor, C: torch.Tensor, dy: Optional[torch.Tensor],
               mode: int = 1, want_grad_X: bool = False, fwd_only: bool = False, reps: int = 5, rounds: int = 1,
               verbose: bool = False, store: bool = True, log: Optional[List[str]] = None) -> Dict[str, object]:
    """Coordinate descent over the fused evaluator's variants per bucket,
    timing `fused_root` on the given batch (arguments as there).  Returns
    {config, ms, trials, key}; with store=True the fastest configuration goes
    into the workload cache, where `fused_root` finds it.  Every trial is
    appended to `log` (and printed with verbose=True)."""
    from .fused import fused_root

    is64 = X.dtype == torch.float64
    out: Dict[str, torch.Tensor] = {}

    def run_with(cfg):
        return lambda: fused_root(X, Ops, Ch, C, dy, mode=mode, want_grad_X=want_grad_X, fwd_only=fwd_only,
                                  raise_errors=False, out=out, config=cfg)

    n_var = [len(v) for v in fused_variants(is64)]
    best = [-1] * N_BUCKETS
    best_t = _time(run_with(best), reps)
    trials = [(list(best), best_t)]

    def note(cfg, t):
        line = f"[tune_fused] {cfg}: {t * 1e3:.3f} ms"
        if log is not None:
            log.append(line)
        if verbose:
            print(line, flush=True)

    note(best, best_t)
    for _ in range(rounds):
        for b in range(N_BUCKETS):
            start = best[b]
            for v in range(n_var[b]):
                if v == start or (start == -1 and n_var[b] == 1):
                    continue
                cfg = list(best)
                cfg[b] = v
                t = _time(run_with(cfg), reps)
                trials.append((cfg, t))
                note(cfg, t)
                if t < best_t:
                    best, best_t = cfg, t
    key = fused_workload_key(X.device, X.dtype, Ops.shape[1], X.shape[0], X.shape[1], want_grad_X, fwd_only)
    if store:
        cache_put(key, best, best_t * 1e3)
    return dict(config=best, ms=best_t * 1e3, trials=trials, key=key)
