#!/usr/bin/env python3
"""Headline benchmark: expressions fitted per second on one MI355X, learned
warm start versus plain (randn) starts.

Workload (the warm-start test set): P problems, each a G2 expression (postfix
length <= 64, at least one learnable constant) with its own noiseless target
Y = f(X; C*) on 512 shared points X ~ U(-3, 3) (n_x = 3), true constants
C* ~ N(0, 2^2).  A step fits every problem's constants with the
device-resident batched L-BFGS (max_iter 100, history 10, strong Wolfe,
tol_grad 1e-7), each problem stopping once its loss is <= 5e-7 x Var(Y)
(the time to rel-MSE <= 1e-6).  A problem counts as fitted when its relative
MSE sum_n (f(X; C) - Y)^2 / sum_n (Y - mean Y)^2 <= 1e-6, with f evaluated in
FP64 at the fitted FP32 constants (the FP64 evaluator on the GPU).

Two methods, timed alike (inputs resident in HBM, FP64 judging after the
timed region):
  plain: L-BFGS from randn starts;
  warm:  the MLP's predicted constants (features + network on MFMA), the
         loss at both starts, the per-problem choice of the lower one, then
         the same L-BFGS.
value = fitted problems per second of the warm-start step.  The network is
trained beforehand (untimed, reported separately) on (Ops, Y, fitted C)
tuples labelled by the plain L-BFGS, from a disjoint sample of problems.

Also reported: the roofline of the fitting kernel (algorithmic FP32 slots per
closure x closures / kernel time, against the FP32 peak), the evaluate
measurement at 65536 expressions x 1024 points (fused root-only, and the API
full-cache forward/backward), the uneven-batching fit of 32768 problems with their own
datasets, and a CPU baseline (the oracle's restatement of the
reference's CPU algorithm on a bounded sample, single core).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

G2 = """E -> E + T [0.3]
E -> E - T [0.1]
E -> T [0.6]
T -> T * F [0.3]
T -> T / F [0.1]
T -> F [0.6]
F -> R ( E ) [0.15]
F -> V [0.85]
R -> sin [0.2]
R -> cos [0.2]
R -> exp [0.2]
R -> log [0.2]
R -> sqrt [0.2]
V -> X_0 [0.25]
V -> X_1 [0.25]
V -> X_2 [0.25]
V -> C [0.25]
"""

# Algorithmic FP32 slots per node per point (forward, backward) with a
# transcendental counted as T = 4 slots; interpreter overhead earns nothing.
T_COST = 4
OPCOST = {
    0: (0, 0), 1: (0, 1),
    **{c: (0, 0) for c in range(2, 9)},
    10: (T_COST + 2, T_COST + 3), 11: (T_COST + 2, T_COST + 3), 12: (T_COST + 1, 1), 13: (T_COST + 1, T_COST + 1),
    14: (1, 2), 15: (T_COST, T_COST + 2), 16: (T_COST + 2, T_COST + 3), 17: (T_COST + 2, T_COST + 3),
    18: (T_COST + 2, T_COST + 3), 19: (T_COST + 2, T_COST + 3), 20: (T_COST + 2, T_COST + 3),
    21: (T_COST + 2, T_COST + 3), 22: (T_COST + 2, T_COST + 3), 23: (1, 0), 24: (1, 0), 25: (T_COST + 1, T_COST + 1),
    26: (T_COST + 1, T_COST + 1), 27: (1, 1), 28: (T_COST, T_COST + 2), 29: (2, 3), 30: (2, 3), 31: (3, 4),
    40: (1, 2), 41: (1, 2), 42: (1, 2), 43: (T_COST + 1, T_COST + 3), 44: (2 * T_COST + 2, 2 * T_COST + 4),
}
LOSS_SLOTS = 5  # d = y - Y, d*d, accumulate, 2*d, *1/N
FP32_PEAK_TFLOPS = 157.3
VALU_LANE_PEAK = 256 * 4 * 16 * 2.4e9  # unpacked FP32 lane-instructions/s (3.93e13)
HBM_PEAK_GBS = 8000.0
FIT_TRAFFIC_PROFILED = 101.3e6  # HBM bytes per fit call, rocprofv3 FETCH_SIZE/WRITE_SIZE (calibrated)
EVAL_TRAFFIC_PROFILED = 887.6e6  # HBM bytes per fused evaluate call (65536 x 1024), same counters


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def expression_cost(ops_row) -> int:
    return sum(sum(OPCOST.get(int(o), (0, 1 if int(o) >= 50 else 0))) for o in ops_row if int(o) != 0)


def bench_evaluate(dev, steps: int, seed: int = 1, B: int = 65536, N: int = 1024):
    """Plain evaluate forward+backward at scale: B G2 expressions (padded 64,
    n_x = 3), X ~ U(-3,3) (N,3), C ~ N(0,1) on the constant slots, upstream
    dL/dy ~ N(0,1) on the root output; one fused root-only forward+backward
    with grad_X per step, timed with HIP events."""
    import torch

    import cudasymbolicregression_amd as st
    from cudasymbolicregression_amd import fused

    g = st.ProbabilisticContextFreeGrammar(G2, "E", 64, 3, torch.device("cpu"))
    torch.manual_seed(seed)
    ops, ch = g.sample(B)
    gen = torch.Generator(device=dev).manual_seed(seed)
    X = torch.rand(N, 3, generator=gen, device=dev) * 6 - 3
    ops, ch = ops.to(dev), ch.to(dev)
    C = torch.zeros(ops.shape, device=dev)
    m = ops == 1
    C[m] = torch.randn(int(m.sum()), generator=gen, device=dev)
    dy = torch.randn(B, N, generator=gen, device=dev)
    out = {}
    fused.fused_root(X, ops, ch, C, dy, mode=0, want_grad_X=True, raise_errors=False, out=out)
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(steps):
        r = fused.fused_root(X, ops, ch, C, dy, mode=0, want_grad_X=True, raise_errors=False, out=out)
    end.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(end) / steps
    M = ops.shape[1]
    nbytes = (X.numel() * 4 + ops.numel() * 8 + ch.numel() * 8 + C.numel() * 4 + dy.numel() * 4 + B * N * 4
              + B * M * 4 + X.numel() * 4)
    slots = float(sum(expression_cost(row) for row in ops.cpu().tolist())) * N
    t = ms / 1e3
    return dict(workload=f"G2 B={B}, padded 64, N={N}, n_x=3, fused fwd+bwd with grad_X", ms=round(ms, 4),
                hbm=dict(achieved_GBps=round(nbytes / t / 1e9, 1), peak_GBps=HBM_PEAK_GBS,
                         frac=round(nbytes / t / 1e9 / HBM_PEAK_GBS, 4), bytes=nbytes,
                         traffic=EVAL_TRAFFIC_PROFILED, traffic_source="profiles/evaluate_c2_hbm_traffic.md"),
                valu=dict(achieved_TFLOPs=round(slots / t / 1e12, 4), peak_TFLOPs=FP32_PEAK_TFLOPS,
                          frac=round(slots / t / 1e12 / FP32_PEAK_TFLOPS, 5),
                          frac_unpacked_lane_peak=round(slots / t / VALU_LANE_PEAK, 4), slots=slots),
                mean_len=round(float((ops != 0).sum(1).double().mean()), 3),
                errors=int((r["err"] != 0).sum()))


def bench_full_cache(dev, steps: int, seed: int = 1, B: int = 65536, N: int = 1024):
    """The API-compatible evaluate at 65536 expressions x 1024 points: the forward writes the
    (M, N, B) cache of every node (M = 64: 17.2 GB) and the backward reads an
    upstream gradient of the same shape (nonzero on each root row) and the
    cached values; both through the C-ABI, timed with HIP events.  HBM bytes:
    forward M*N*B*4 written; backward M*N*B*4 read (upstream) plus the cached
    rows of the non-padding nodes."""
    import torch

    import cudasymbolicregression_amd as st
    from cudasymbolicregression_amd import _lib

    g = st.ProbabilisticContextFreeGrammar(G2, "E", 64, 3, torch.device("cpu"))
    torch.manual_seed(seed)
    ops, ch = g.sample(B)
    gen = torch.Generator(device=dev).manual_seed(seed)
    X = torch.rand(N, 3, generator=gen, device=dev) * 6 - 3
    ops, ch = ops.to(dev), ch.to(dev)
    C = torch.zeros(ops.shape, device=dev)
    m = ops == 1
    C[m] = torch.randn(int(m.sum()), generator=gen, device=dev)
    M = ops.shape[1]
    ops_mb, ch_mb, C_mb = ops.t().contiguous(), ch.permute(1, 0, 2).contiguous(), C.t().contiguous()
    lib, s = _lib.hip(), _lib.stream_of(dev)
    cache = torch.empty((M, N, B), device=dev)
    fws = torch.empty(int(lib.sr_forward_mb_workspace_bytes(M, N, B, 1, 0)), dtype=torch.uint8, device=dev)
    fwd = lambda: _lib.check_status(lib.sr_forward_mb_ws_f32(_lib.ptr(X), N, 3, _lib.ptr(ops_mb),  # noqa
                                                             _lib.ptr(ch_mb), _lib.ptr(C_mb), M, B, _lib.ptr(cache),
                                                             _lib.ptr(fws), fws.numel(), s), "forward")
    fwd_direct = lambda: _lib.check_status(lib.sr_forward_mb_f32(_lib.ptr(X), N, 3, _lib.ptr(ops_mb),  # noqa
                                                                 _lib.ptr(ch_mb), _lib.ptr(C_mb), M, B,
                                                                 _lib.ptr(cache), s), "forward")
    ms_fd, _ = timed(fwd_direct, steps, 1)
    ms_f, _ = timed(fwd, steps, 1)  # the API's path; the backward below reads this cache
    del fws
    root = ((ops != 0).sum(1) - 1).clamp(min=0)
    gup = torch.zeros((M, N, B), device=dev)
    gup[root, :, torch.arange(B, device=dev)] = torch.randn(B, N, generator=gen, device=dev)
    gC, gX = torch.empty_like(C_mb), torch.empty_like(X)
    ws = torch.empty(max(int(lib.sr_backward_mb_workspace_bytes(B, N, 3, 0)), 256), dtype=torch.uint8, device=dev)
    rws = torch.empty(int(lib.sr_backward_mb_rows_workspace_bytes(M, N, B, 1, 0)), dtype=torch.uint8, device=dev)
    err = torch.empty(1, dtype=torch.int32, device=dev)
    bwd_direct = lambda: _lib.check_status(lib.sr_backward_mb_f32(  # noqa: E731
        _lib.ptr(gup), _lib.ptr(cache), _lib.ptr(ops_mb), _lib.ptr(ch_mb), M, B, N, 3, _lib.ptr(gC), _lib.ptr(gX),
        _lib.ptr(ws), ws.numel(), _lib.ptr(err), s), "backward")
    bwd = lambda: _lib.check_status(lib.sr_backward_mb_rows_f32(  # noqa: E731
        _lib.ptr(gup), _lib.ptr(cache), _lib.ptr(ops_mb), _lib.ptr(ch_mb), M, B, N, 3, _lib.ptr(gC), _lib.ptr(gX),
        _lib.ptr(ws), ws.numel(), _lib.ptr(rws), rws.numel(), _lib.ptr(err), s), "backward")
    ms_bd, _ = timed(bwd_direct, steps, 1)
    ms_b, _ = timed(bwd, steps, 1)  # the API's path
    del rws
    L = float((ops != 0).sum(1).double().mean())
    fb = M * N * B * 4
    bb = M * N * B * 4 + int(L * N * B * 4)
    del cache, gup
    torch.cuda.empty_cache()
    gbs = lambda nbytes, ms: round(nbytes / (ms / 1e3) / 1e9, 1)  # noqa: E731
    return dict(workload=f"G2 B={B}, padded 64, N={N}, n_x=3: API full-cache evaluate, cache (M,N,B) fp32",
                forward=dict(ms=round(ms_f, 3), bytes=fb, achieved_GBps=gbs(fb, ms_f),
                             frac=round(gbs(fb, ms_f) / HBM_PEAK_GBS, 4),
                             note="two passes through a row buffer (sr_forward_mb_ws_f32, the API's path)",
                             ms_without_workspace=round(ms_fd, 3)),
                backward=dict(ms=round(ms_b, 3), bytes=bb, achieved_GBps=gbs(bb, ms_b),
                              frac=round(gbs(bb, ms_b) / HBM_PEAK_GBS, 4),
                              note="gathered rows, then the reverse sweep (sr_backward_mb_rows_f32, the API's path)",
                              ms_without_row_workspace=round(ms_bd, 3)))


def bench_ragged(dev, steps: int, seed: int = 3, B: int = 32768):
    """Uneven batching: B problems, each its own G2 expression (postfix length
    3..127, padded 128, >= 1 constant) with its own N_i = round(exp(U(ln 64,
    ln 4096))) points X_i ~ U(-3, 3) (n_x = 3) and noiseless targets from true
    constants ~ N(0, 2^2); one ragged fit launch per step from randn starts
    (stop at loss <= 5e-7 Var(Y_i)), judged in FP64."""
    import torch

    import cudasymbolicregression_amd as st
    from cudasymbolicregression_amd import fused, warmstart
    from cudasymbolicregression_amd.optimization import _fit

    g = st.ProbabilisticContextFreeGrammar(G2, "E", 128, 3, torch.device("cpu"))
    torch.manual_seed(seed)
    gen = torch.Generator().manual_seed(seed)
    parts = []
    have = 0
    while have < B:  # candidate rounds until B problems have finite, non-constant targets
        ops, ch = g.sample(65536)
        L = (ops != 0).sum(1)
        ok = (L >= 3) & (L <= 127) & (ops == 1).any(1)
        ops, ch = ops[ok], ch[ok]
        P = ops.shape[0]
        n = torch.exp(torch.empty(P).uniform_(math.log(64), math.log(4096), generator=gen)).round().to(torch.int64)
        X = torch.rand(int(n.sum()), 3, generator=gen) * 6 - 3
        Ct = torch.zeros(ops.shape)
        m = ops == 1
        Ct[m] = torch.randn(int(m.sum()), generator=gen) * 2.0
        ops, ch, Ct, n, X = ops.to(dev), ch.to(dev), Ct.to(dev), n.to(dev), X.to(dev)
        y = fused.fused_root_ragged(X.double(), n, ops, ch, Ct.double(), None, fwd_only=True, raise_errors=False)["y"]
        seg = torch.repeat_interleave(torch.arange(P, device=dev), n)
        bad = torch.zeros(P, dtype=torch.float64, device=dev).index_add_(
            0, seg, (~torch.isfinite(y) | (y.abs() >= 1e6)).double())
        yz = torch.nan_to_num(y)
        mean = torch.zeros(P, dtype=torch.float64, device=dev).index_add_(0, seg, yz) / n
        var = torch.zeros(P, dtype=torch.float64, device=dev).index_add_(0, seg, (yz - mean[seg]) ** 2) / n
        sel = ((bad == 0) & (var > 1e-8)).nonzero().flatten()[: B - have]
        rows = torch.isin(seg, sel)
        parts.append((ops[sel], ch[sel], n[sel], var[sel], X[rows], y[rows].float()))
        have += sel.numel()
    ops, ch, n, var, X, Y = (torch.cat([p_[i] for p_ in parts]).contiguous() for i in range(6))
    Bk = ops.shape[0]
    x_off, y_off, n32 = fused.ragged_offsets(n, 3)
    C0 = warmstart.random_start(ops, seed + 1)
    stop = (5e-7 * var).contiguous()
    run = lambda: _fit(X, Y, ops, ch, C0, 100, 10, 1e-7, "strong_wolfe", stop_loss=stop, x_off=x_off, y_off=y_off,
                       n_pts=n32)
    ms, (C, loss, iters, err, _, ncl) = timed(run, steps, 1)
    yf = fused.fused_root_ragged(X.double(), n, ops, ch, C.double(), None, fwd_only=True, raise_errors=False)["y"]
    seg = torch.repeat_interleave(torch.arange(Bk, device=dev), n)
    num = torch.zeros(Bk, dtype=torch.float64, device=dev).index_add_(0, seg, (yf - Y.double()) ** 2)
    r = num / (var * n)
    fitted = int((torch.isfinite(r) & (r <= 1e-6)).sum())
    per_pt = torch.tensor([expression_cost(row) + LOSS_SLOTS for row in ops.cpu().tolist()], dtype=torch.float64)
    slots = float((per_pt * ncl.cpu().double() * n.cpu().double()).sum())
    return dict(workload=f"G2 B={Bk} own expressions (postfix 3..127, padded 128), own datasets N_i log-uniform "
                         f"[64, 4096] (mean {float(n.double().mean()):.1f}), n_x=3, randn start, stop at "
                         f"loss <= 5e-7 Var(Y_i)",
                ms=round(ms, 3), fitted=fitted, fitted_per_s=round(fitted / (ms / 1e3), 1),
                points=int(n.sum()), closures=int(ncl.sum()), mean_iters=round(float(iters.double().mean()), 2),
                problems_without_error=int((err == 0).sum()),
                valu=dict(achieved_TFLOPs=round(slots / (ms / 1e3) / 1e12, 4), peak_TFLOPs=FP32_PEAK_TFLOPS,
                          frac=round(slots / (ms / 1e3) / 1e12 / FP32_PEAK_TFLOPS, 5), slots=slots))


def fit(w, C_start, stop=True):
    from cudasymbolicregression_amd.optimization import _fit

    return _fit(w["X"], w["Y"], w["ops"], w["ch"], C_start, 100, 10, 1e-7, "strong_wolfe",
                stop_loss=w["stop"] if stop else None)


def cpu_baseline(w, sample: int):
    """The oracle's CPU restatement of the reference algorithm (randn starts,
    every iteration, no early stop) on `sample` problems."""
    import numpy as np
    import torch

    from cudasymbolicregression_amd import warmstart

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import lbfgs_oracle

    X = w["X"].cpu().numpy()
    ops = w["ops"][:sample].cpu().numpy()
    ch = w["ch"][:sample].cpu().numpy()
    Y = w["Y"][:sample].cpu().numpy()
    C0 = w["C0"][:sample].cpu().numpy()
    t0 = time.perf_counter()
    with np.errstate(all="ignore"):
        C, _, _, calls = lbfgs_oracle.optimize(X, Y, ops, ch, C0, max_iter=100)
    dt = time.perf_counter() - t0
    r = warmstart.rel_mse_fp64(w["X"], w["Y"][:sample], w["ops"][:sample], w["ch"][:sample],
                               torch.from_numpy(C).to(w["X"].device))
    fitted = int((r <= 1e-6).sum())
    return dict(value=fitted / dt, unit="expressions fitted/s", cores=1, kind="port",
                sample=f"{sample} problems of the same workload, randn starts, full fits (max_iter 100), {dt:.1f} s, "
                       f"{fitted} fitted, {calls} batched closures; oracle/lbfgs_oracle.py + oracle/eval_oracle.c, "
                       f"single thread")


def timed(fn, steps: int, warmup: int):
    """ms per call of fn() between HIP events on the current stream (every
    kernel of a call is joined back to it), after `warmup` untimed calls."""
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(steps):
        out = fn()
    end.record()
    torch.cuda.synchronize()
    return start.elapsed_time(end) / steps, out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--problems", type=int, default=16384)
    ap.add_argument("--seed", type=int, default=5, help="test problems (training uses seed - 1)")
    ap.add_argument("--train-tuples", type=int, default=2_000_000)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--cpu-sample", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-evaluate", action="store_true")
    ap.add_argument("--no-ragged", action="store_true")
    ap.add_argument("--no-warm", action="store_true",
                    help="plain fits only: no MLP training and no warm-start step (value is the plain rate)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-MiB device buffer (calibrates the HBM byte counters) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        import torch

        src = torch.ones(args.copy_probe * (1 << 20) // 4, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": src.numel() * 4}))
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): fitting problems are independent, "
              "so the path shards by problem with no collective; run one process per GPU.", file=sys.stderr)
        return 2

    import torch

    import cudasymbolicregression_amd as st
    from cudasymbolicregression_amd import warmstart

    dev = torch.device("cuda:0")
    t0 = time.time()
    g = st.ProbabilisticContextFreeGrammar(G2, "E", 64, 3, torch.device("cpu"))
    gen = torch.Generator().manual_seed(0)
    X = (torch.rand(512, 3, generator=gen) * 6 - 3).to(dev)  # the seed-fixed points of every problem
    ops, ch, Ctrue, Y = warmstart.sample_problems(g, X, args.problems, args.seed)
    var = Y.double().var(dim=1, unbiased=False)
    w = dict(X=X, ops=ops, ch=ch, Ctrue=Ctrue, Y=Y, C0=warmstart.random_start(ops, args.seed + 1),
             stop=(5e-7 * var).contiguous())
    torch.cuda.synchronize()
    log(f"test set ready: {args.problems} problems in {time.time() - t0:.1f} s")

    # ---- training (untimed): labels from the plain L-BFGS on a disjoint sample
    tstats = None
    if not args.no_warm:
        t0 = time.time()
        tops, tch, tY, label, tstats = warmstart.make_training_set(g, X, args.train_tuples, args.seed - 1)
        model = warmstart.WarmStartMLP(dev, seed=0)
        torch.cuda.synchronize()
        t1 = time.time()
        hist = warmstart.train(model, tops, tY, label, epochs=args.epochs, batch=4096, lr=1e-3, lr_final=1e-5)
        torch.cuda.synchronize()
        tstats.update(train_s=round(time.time() - t1, 2), epochs=args.epochs, batch=4096,
                      loss_first_epoch=round(hist[0], 4), loss_last_epoch=round(hist[-1], 4),
                      zero_predictor_loss=round(float((label.double() ** 2).sum()) / max(1, label.shape[0]), 4))
        log(f"training: {tstats} ({time.time() - t0:.1f} s)")
        del tops, tch, tY, label

    # ---- plain: randn starts
    ms_p, (C_p, loss_p, iters_p, err_p, _, ncl_p) = timed(lambda: fit(w, w["C0"]), args.steps, args.warmup)
    fit_p = int((warmstart.rel_mse_fp64(X, Y, ops, ch, C_p) <= 1e-6).sum())
    log(f"plain: {ms_p:.2f} ms/step, {fit_p} fitted")

    plain_value = fit_p / (ms_p / 1000.0)
    warm = None
    if not args.no_warm:
        # ---- warm: network prediction, start selection, fit
        def warm_step():
            C_start, used, _, _ = warmstart.warm_start(model, X, Y, ops, ch, w["C0"])
            return fit(w, C_start) + (used,)

        ms_w, (C_w, loss_w, iters_w, err_w, _, ncl_w, used) = timed(warm_step, args.steps, args.warmup)
        fit_w = int((warmstart.rel_mse_fp64(X, Y, ops, ch, C_w) <= 1e-6).sum())
        ms_ws, _ = timed(lambda: warmstart.warm_start(model, X, Y, ops, ch, w["C0"]), args.steps, 1)
        log(f"warm: {ms_w:.2f} ms/step ({ms_ws:.2f} ms start selection), {fit_w} fitted")
        value = fit_w / (ms_w / 1000.0)
        warm = {"ms": round(ms_w, 3), "start_selection_ms": round(ms_ws, 3), "fitted": fit_w,
                "fitted_per_s": round(value, 1), "warm_start_taken": round(float(used.double().mean()), 4),
                "mean_iters": round(float(iters_w.double().mean()), 2),
                "closures": int(ncl_w.sum()), "problems_without_error": int((err_w == 0).sum())}
        fit_ms, ncl_r, ms_step = ms_w - ms_ws, ncl_w, ms_w
    else:
        value = plain_value
        fit_ms, ncl_r, ms_step = ms_p, ncl_p, ms_p

    # roofline of the fitting kernels (warm step, or the plain step without
    # one): algorithmic slots / fit time
    per_point = torch.tensor([expression_cost(row) + LOSS_SLOTS for row in ops.cpu().tolist()], dtype=torch.float64)
    flops = float((per_point * ncl_r.cpu().double() * X.shape[0]).sum())
    achieved = flops / (fit_ms / 1000.0) / 1e12
    roofline = dict(bound="mfma", achieved=round(achieved, 4), peak=FP32_PEAK_TFLOPS, unit="TFLOP/s",
                    frac=round(achieved / FP32_PEAK_TFLOPS, 5), traffic=FIT_TRAFFIC_PROFILED,
                    traffic_source="profiles/fit_hbm_traffic.md: rocprofv3 FETCH_SIZE/WRITE_SIZE of the plain-start "
                                   "fit call (the same kernels on the same problems), not measured live",
                    kernel="sr_lbfgs_fit: the k_fit bucket kernels of one call (concurrent)",
                    flop_per_launch=flops, frac_unpacked_lane_peak=round(flops / (fit_ms / 1e3) / VALU_LANE_PEAK, 5))

    ev = None
    if not args.no_evaluate:
        ev = bench_evaluate(dev, steps=5)
        log(f"evaluate (65536 x 1024): {ev['ms']:.3f} ms, {ev['hbm']['achieved_GBps']} GB/s, VALU frac {ev['valu']['frac']}")

    fc = None
    if not args.no_evaluate:
        fc = bench_full_cache(dev, steps=3)
        log(f"full-cache evaluate (65536 x 1024): forward {fc['forward']['ms']:.2f} ms ({fc['forward']['achieved_GBps']} GB/s), "
            f"backward {fc['backward']['ms']:.2f} ms ({fc['backward']['achieved_GBps']} GB/s)")

    rg = None
    if not args.no_ragged:
        rg = bench_ragged(dev, steps=2)
        log(f"ragged (32768 own datasets): {rg['ms']:.1f} ms, {rg['fitted']} fitted, {rg['fitted_per_s']:.0f} fitted/s")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(w, args.cpu_sample)
        log(f"cpu baseline: {cpu['value']:.2f} fitted/s ({cpu['sample']})")

    out = {
        "metric": ("expressions fitted/sec (rel-MSE<=1e-6), plain L-BFGS only, 1 MI355X" if args.no_warm else
                   "expressions fitted/sec (rel-MSE<=1e-6), warm-start vs plain L-BFGS, 1 MI355X"),
        "value": round(value, 2),
        "unit": "expressions fitted/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "warm-start test set: G2 expressions (len<=64, >=1 constant), own noiseless targets on 512 "
                        "seed-fixed points X~U(-3,3), n_x=3, C*~N(0,2^2); MLP warm start with randn fallback",
            "problems": args.problems, "n_points": 512, "n_x": 3, "max_iter": 100, "history": 10,
            "line_search": "strong_wolfe", "tol_grad": 1e-7, "stop_rel_loss": 5e-7,
            "warm": warm,
            "plain": {"ms": round(ms_p, 3), "fitted": fit_p, "fitted_per_s": round(plain_value, 1),
                      "mean_iters": round(float(iters_p.double().mean()), 2), "closures": int(ncl_p.sum()),
                      "problems_without_error": int((err_p == 0).sum())},
            "warm_over_plain": round(value / plain_value, 4) if warm is not None and plain_value > 0 else None,
            "training": tstats,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "evaluate": ev,
        "evaluate_full_cache": fc,
        "ragged": rg,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
