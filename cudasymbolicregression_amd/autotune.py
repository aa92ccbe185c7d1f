"""Autotuner for the fit kernels' launch variants.

The device L-BFGS (`optimize`) buckets problems by tape rows and runs
one kernel per bucket; each bucket has a few launch variants that trade
per-problem latency against throughput: W waves cooperating on one problem
(points split between them) and PPT points per lane per instruction dispatch.
Which variant is fastest depends on the workload (program lengths, points,
iteration counts), so `tune_fit` times them on a representative batch with
coordinate descent over the buckets and keeps the fastest for the process.
"""
from __future__ import annotations

import ctypes
import time
from typing import Callable, Dict, List, Optional

import torch

from . import _lib


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


def get_config() -> List[int]:
    lib = _lib.hip()
    return [lib.sr_fit_get_variant(b) for b in range(lib.sr_fit_num_buckets())]


def set_config(config: List[int]) -> None:
    lib = _lib.hip()
    for b, v in enumerate(config):
        if lib.sr_fit_set_variant(b, int(v)) != 0:
            raise ValueError(f"invalid variant {v} for bucket {b}")


def _time(fn: Callable[[], object], reps: int) -> float:
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def tune_fit(X: torch.Tensor, Y: torch.Tensor, Ops: torch.Tensor, Ch: torch.Tensor, C_init: torch.Tensor,
             max_iter: int = 100, reps: int = 2, rounds: int = 1, verbose: bool = False,
             **kw) -> Dict[str, object]:
    """Coordinate descent over the buckets' variants, timing `optimize` on
    the given batch (keyword options: history_size, tolerance_grad,
    line_search_fn, stop_loss as in the fit); leaves the fastest
    configuration active and returns {config, ms, trials}."""
    from .optimization import _fit

    Xc, Yc = X.contiguous(), Y.to(torch.float32).contiguous()
    Oc, Cc = Ops.to(torch.int64).contiguous(), Ch.to(torch.int64).contiguous()
    C0 = C_init.to(torch.float32).contiguous()
    stop = kw.get("stop_loss")
    run = lambda: _fit(Xc, Yc, Oc, Cc, C0, max_iter, kw.get("history_size", 10), kw.get("tolerance_grad", 1e-7),
                       kw.get("line_search_fn", "strong_wolfe"), stop_loss=stop)
    best = get_config()
    best_t = _time(run, reps)
    trials = [(list(best), best_t)]
    n_var = [len(v) for v in variants()]
    for _ in range(rounds):
        for b in range(len(best)):
            for v in range(n_var[b]):
                if v == best[b]:
                    continue
                cfg = list(best)
                cfg[b] = v
                set_config(cfg)
                t = _time(run, reps)
                trials.append((cfg, t))
                if verbose:
                    print(f"[tune_fit] {cfg}: {t * 1e3:.2f} ms", flush=True)
                if t < best_t:
                    best, best_t = cfg, t
            set_config(best)
    return dict(config=best, ms=best_t * 1e3, trials=trials)
