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
full-cache forward/backward, every result checked against the FP64
evaluator), 16384 candidates fitted to one shared noisy target, the
uneven-batching fit of 32768 problems with their own datasets, the fit from
the true constants (the ceiling of any warm start), closures-per-problem
histograms, the measured cost of a transcendental in FMA slots, and a CPU
baseline: the reference's own OpenMP evaluation units with LBFGS_batch on a
bounded sample, on every host core this process may use.
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
# hardware transcendental counted as T slots; interpreter overhead earns
# nothing.  T is measured on the GPU at the start of every run
# (sr_probe_valu_rate: issue rate of FMA / issue rate of v_exp, v_log,
# v_sin, v_rcp, v_sqrt, v_rsq); 1.8 on MI355X, so the table is rebuilt with
# the measured value.
T_COST = 1.8


def opcost_table(T):
    return {
        0: (0, 0), 1: (0, 1),
        **{c: (0, 0) for c in range(2, 9)},
        10: (T + 2, T + 3), 11: (T + 2, T + 3), 12: (T + 1, 1), 13: (T + 1, T + 1),
        14: (1, 2), 15: (T, T + 2), 16: (T + 2, T + 3), 17: (T + 2, T + 3),
        18: (T + 2, T + 3), 19: (T + 2, T + 3), 20: (T + 2, T + 3),
        21: (T + 2, T + 3), 22: (T + 2, T + 3), 23: (1, 0), 24: (1, 0), 25: (T + 1, T + 1),
        26: (T + 1, T + 1), 27: (1, 1), 28: (T, T + 2), 29: (2, 3), 30: (2, 3), 31: (3, 4),
        40: (1, 2), 41: (1, 2), 42: (1, 2), 43: (T + 1, T + 3), 44: (2 * T + 2, 2 * T + 4),
    }


OPCOST = opcost_table(T_COST)
LOSS_SLOTS = 5  # d = y - Y, d*d, accumulate, 2*d, *1/N
FP32_PEAK_TFLOPS = 157.3
VALU_LANE_PEAK = 256 * 4 * 16 * 2.4e9  # unpacked FP32 lane-instructions/s (3.93e13)
HBM_PEAK_GBS = 8000.0
FIT_TRAFFIC_PROFILED = 99.0e6  # HBM bytes per fit call, rocprofv3 FETCH_SIZE/WRITE_SIZE (calibrated)
EVAL_TRAFFIC_PROFILED = 1150.5e6  # HBM bytes per fused evaluate call (65536 x 1024), same counters


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def expression_cost(ops_row) -> float:
    return sum(sum(OPCOST.get(int(o), (0, 1 if int(o) >= 50 else 0))) for o in ops_row if int(o) != 0)


def measure_trans_cost(dev) -> dict:
    """Issue rates of FP32 FMA and of the hardware transcendentals
    (sr_probe_valu_rate), and their ratio: the cost of a transcendental in
    FMA slots."""
    import torch

    from cudasymbolicregression_amd import _lib

    lib = _lib.hip()
    blocks, iters = 256 * 32, 4096
    out = torch.empty(blocks * 256, device=dev)
    rates = {}
    for k, name in enumerate(["fma", "exp", "log", "sin", "rcp", "sqrt", "rsq"]):
        for _ in range(2):  # the second launch is timed
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            _lib.check_status(lib.sr_probe_valu_rate(k, blocks, iters, _lib.ptr(out), _lib.stream_of(dev)), "probe")
            b.record()
            torch.cuda.synchronize()
        ops = blocks * 256 * lib.sr_probe_valu_ops_per_thread(iters)
        rates[name] = ops / (a.elapsed_time(b) / 1e3)
    trans = [rates[n] for n in rates if n != "fma"]
    cost = rates["fma"] / (sum(trans) / len(trans))
    return dict(trans_cost=round(cost, 2), fma_lane_ops_per_s=round(rates["fma"], -9),
                trans_lane_ops_per_s={n: round(r, -9) for n, r in rates.items() if n != "fma"})


def bench_evaluate(dev, steps: int, seed: int = 1, B: int = 65536, N: int = 1024, fp64_report: bool = True):
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
    fp64 = None
    if fp64_report:
        # every result against the FP64 evaluator on the same inputs (untimed)
        y32, gC32, gX32, e32 = r["y"].clone(), r["grad_C"].clone(), r["grad_X"].clone(), r["err"].clone()
        del out, r
        r64 = fused.fused_root(X.double(), ops, ch, C.double(), dy.double(), mode=0, want_grad_X=True,
                               raise_errors=False)
        ok = (e32 == 0) & (r64["err"] == 0)
        # grad_X sums over every expression, so one non-finite contribution
        # makes a column non-finite (an expression can have finite outputs and
        # constant gradients and still an infinite dL/dX, e.g. sqrt at 0):
        # compare it over the chunks of 64 expressions (among those whose
        # outputs and constant gradients are finite in both precisions) whose
        # grad_X is finite in both precisions
        fin = ok & torch.isfinite(y32).all(1) & torch.isfinite(r64["y"]).all(1) \
            & torch.isfinite(gC32).all(1) & torch.isfinite(r64["grad_C"]).all(1)
        cand = fin.nonzero().flatten()
        keep = []
        for c0 in range(0, cand.numel(), 64):
            part = cand[c0:c0 + 64]
            g32 = fused.fused_root(X, ops[part], ch[part], C[part], dy[part], mode=0, want_grad_X=True,
                                   raise_errors=False)["grad_X"]
            g64 = fused.fused_root(X.double(), ops[part], ch[part], C[part].double(), dy[part].double(), mode=0,
                                   want_grad_X=True, raise_errors=False)["grad_X"]
            if bool(torch.isfinite(g32).all()) and bool(torch.isfinite(g64).all()):
                keep.append(part)
        sel = torch.cat(keep) if keep else cand[:0]
        gx32 = fused.fused_root(X, ops[sel], ch[sel], C[sel], dy[sel], mode=0, want_grad_X=True,
                                raise_errors=False)["grad_X"]
        gx64 = fused.fused_root(X.double(), ops[sel], ch[sel], C[sel].double(), dy[sel].double(), mode=0,
                                want_grad_X=True, raise_errors=False)["grad_X"]
        fp64 = dict(error_codes_equal=int((e32 == r64["err"]).sum()), error_codes_fp32=int((e32 != 0).sum()),
                    error_codes_fp64=int((r64["err"] != 0).sum()),
                    codes_fp32={int(k): int(v) for k, v in zip(*torch.unique(e32[e32 != 0], return_counts=True))},
                    y=rel_error_report(y32[ok], r64["y"][ok]),
                    grad_C=rel_error_report(gC32[ok], r64["grad_C"][ok]),
                    grad_X_all=rel_error_report(gX32, r64["grad_X"]),
                    grad_X_finite_expressions=dict(expressions=int(sel.numel()), **rel_error_report(gx32, gx64)),
                    note="y and grad_C over expressions without an error code in either precision; grad_X_all sums "
                         "every expression's contribution, grad_X_finite_expressions those of the 64-expression "
                         "chunks (of expressions with finite y and grad_C in both precisions) whose grad_X is "
                         "finite in both precisions")
        del r64, y32, gC32
        r = dict(err=e32)
        torch.cuda.empty_cache()
    return dict(workload=f"G2 B={B}, padded 64, N={N}, n_x=3, fused fwd+bwd with grad_X", ms=round(ms, 4),
                hbm=dict(achieved_GBps=round(nbytes / t / 1e9, 1), peak_GBps=HBM_PEAK_GBS,
                         frac=round(nbytes / t / 1e9 / HBM_PEAK_GBS, 4), bytes=nbytes,
                         traffic=EVAL_TRAFFIC_PROFILED, traffic_source="profiles/evaluate_c2_hbm_traffic.md"),
                valu=dict(achieved_TFLOPs=round(slots / t / 1e12, 4), peak_TFLOPs=FP32_PEAK_TFLOPS,
                          frac=round(slots / t / 1e12 / FP32_PEAK_TFLOPS, 5),
                          frac_unpacked_lane_peak=round(slots / t / VALU_LANE_PEAK, 4), slots=slots),
                mean_len=round(float((ops != 0).sum(1).double().mean()), 3),
                errors=int((r["err"] != 0).sum()), fp64_error=fp64)


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
    bwd_read = lambda: _lib.check_status(lib.sr_backward_mb_rows_f32(  # noqa: E731
        _lib.ptr(gup), _lib.ptr(cache), _lib.ptr(ops_mb), _lib.ptr(ch_mb), M, B, N, 3, _lib.ptr(gC), _lib.ptr(gX),
        _lib.ptr(ws), ws.numel(), _lib.ptr(rws), rws.numel(), _lib.ptr(err), s), "backward")
    bwd = lambda: _lib.check_status(lib.sr_backward_mb_rows_recompute_f32(  # noqa: E731
        _lib.ptr(gup), _lib.ptr(cache), _lib.ptr(X), _lib.ptr(C_mb), _lib.ptr(ops_mb), _lib.ptr(ch_mb), M, B, N, 3,
        _lib.ptr(gC), _lib.ptr(gX), _lib.ptr(ws), ws.numel(), _lib.ptr(rws), rws.numel(), _lib.ptr(err), s),
        "backward")
    ms_bd, _ = timed(bwd_direct, steps, 1)
    ms_br, _ = timed(bwd_read, steps, 1)
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
                              note="gathered gradient rows, then the reverse sweep recomputing the forward values "
                                   "(sr_backward_mb_rows_recompute_f32, the API's path); bytes are the reference's "
                                   "(upstream gradient and cached values read)",
                              ms_reading_cached_values=round(ms_br, 3),
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


def rel_error_report(a, b, tol: float = 1e-5) -> dict:
    """Error of FP32 results `a` against FP64 results `b` (same shape):
    where both are finite, quantiles of |a - b| / |b| (|b| > 0) and the
    share within `tol` relative; where either is not, whether the
    non-finite masks agree."""
    import torch

    a, b = a.double().flatten(), b.double().flatten()
    fa, fb = torch.isfinite(a), torch.isfinite(b)
    both = fa & fb
    d = (a[both] - b[both]).abs()
    ref = b[both].abs()
    nz = ref > 0
    rel = torch.sort(d[nz] / ref[nz])[0]
    q = lambda f: float(rel[min(rel.numel() - 1, int(f * rel.numel()))]) if rel.numel() else 0.0  # noqa: E731
    return dict(values=int(a.numel()), finite_both=int(both.sum()),
                within_rel_tol=round(float((d <= tol * ref).double().mean()) if d.numel() else 1.0, 7),
                rel_tol=tol, rel_p50=q(0.5), rel_p99=q(0.99), rel_p999=q(0.999), rel_max=q(1.0),
                abs_max=float(d.max()) if d.numel() else 0.0,
                exact_zero_ref_nonzero_value=int(((ref == 0) & (d > 0)).sum()),
                nonfinite_mask_mismatch=int((fa != fb).sum()),
                nan_mask_mismatch=int((torch.isnan(a) != torch.isnan(b)).sum()))


def closure_histogram(ncl) -> dict:
    """Problems per closures-per-problem bin (bin k holds [edge_k, edge_k+1))."""
    import torch

    edges = [1, 2, 3, 5, 9, 17, 33, 65, 129, 257, 513, 1025]
    c = ncl.to(torch.int64).cpu()
    out = {}
    for lo, hi in zip(edges, edges[1:] + [1 << 40]):
        n = int(((c >= lo) & (c < hi)).sum())
        out[f"{lo}-{hi - 1}" if hi < (1 << 40) else f">={lo}"] = n
    out["0"] = int((c == 0).sum())
    return out


def bench_c3(dev, steps: int, seed: int = 2, B: int = 16384, N: int = 512):
    """Plain batched L-BFGS, B candidates fitted to one shared noisy target:
    G2 expressions (padded 64) on N points X ~ U(-3, 3)^3, the target a
    seed-drawn G2 expression with >= 1 constant, constants ~ N(0, 2^2), plus
    N(0, 0.01^2) noise; randn starts, history 10, max_iter 100, strong Wolfe,
    tol_grad 1e-7, every problem run to the reference's own stop rules.  The
    FP32 loss each fit ends with is compared with the FP64 evaluation of the
    same constants."""
    import torch

    import cudasymbolicregression_amd as st
    from cudasymbolicregression_amd import fused
    from cudasymbolicregression_amd.optimization import _fit

    gen = torch.Generator().manual_seed(seed)
    X = (torch.rand(N, 3, generator=gen) * 6 - 3).to(dev)
    g = st.ProbabilisticContextFreeGrammar(G2, "E", 64, 3, torch.device("cpu"))
    torch.manual_seed(seed)
    while True:
        o, c = g.sample(1)
        if not (o == 1).any():
            continue
        Ct = torch.where(o == 1, torch.randn(o.shape, generator=gen) * 2, torch.zeros(o.shape))
        y = fused.fused_root(X.double(), o.to(dev), c.to(dev), Ct.to(dev).double(), None, fwd_only=True,
                             raise_errors=False)["y"][0]
        if torch.isfinite(y).all() and y.var() > 1e-3 and y.abs().max() < 1e4:
            break
    target = (y + 0.01 * torch.randn(N, generator=gen, dtype=torch.float64).to(dev)).float()
    tstr = g.to_string(g.postfix_to_infix(o, 128))[0]
    ops, ch = g.sample(B)
    ops, ch = ops.to(dev), ch.to(dev)
    C0 = st.create_constants(ops, method="randn")
    Y = target.unsqueeze(0).expand(B, -1).contiguous()
    run = lambda: _fit(X, Y, ops, ch, C0, 100, 10, 1e-7, "strong_wolfe")  # noqa: E731
    ms, (C, loss, iters, err, _, ncl) = timed(run, steps, 1)
    y64 = fused.fused_root(X.double(), ops, ch, C.double(), None, fwd_only=True, raise_errors=False)["y"]
    mse64 = ((y64 - Y.double()) ** 2).mean(1)
    ok = (err == 0) & torch.isfinite(loss) & torch.isfinite(mse64)
    per_pt = torch.tensor([expression_cost(row) + LOSS_SLOTS for row in ops.cpu().tolist()], dtype=torch.float64)
    slots = float((per_pt * ncl.cpu().double()).sum()) * N
    noise_var = 1e-4
    best = float(mse64[ok].min()) if bool(ok.any()) else float("nan")
    return dict(workload=f"G2 B={B} candidates (padded 64) fitted to one shared noisy target on N={N} points, n_x=3, "
                         f"randn starts, max_iter 100, history 10, strong Wolfe, tol_grad 1e-7",
                target=tstr, ms=round(ms, 3), fits_per_s=round(B / (ms / 1e3), 1),
                mean_iters=round(float(iters.double().mean()), 2), closures=int(ncl.sum()),
                problems_without_error=int((err == 0).sum()),
                fp64_loss_error=rel_error_report(loss[ok], mse64[ok], tol=1e-5),
                best_mse_fp64=best, noise_variance=noise_var,
                within_2x_noise=int((mse64[ok] <= 2 * noise_var).sum()),
                closures_per_problem=closure_histogram(ncl),
                valu=dict(achieved_TFLOPs=round(slots / (ms / 1e3) / 1e12, 4), peak_TFLOPs=FP32_PEAK_TFLOPS,
                          frac=round(slots / (ms / 1e3) / 1e12 / FP32_PEAK_TFLOPS, 5), slots=slots))


def fit(w, C_start, stop=True):
    from cudasymbolicregression_amd.optimization import _fit

    return _fit(w["X"], w["Y"], w["ops"], w["ch"], C_start, 100, 10, 1e-7, "strong_wolfe",
                stop_loss=w["stop"] if stop else None)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform

    return platform.processor() or "unknown"


def host_threads() -> int:
    """The host cores this process may use: its affinity set, capped by
    OMP_NUM_THREADS where that is set (the machine's share for this GPU)."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def reference_cpu_fit(X, Y, ops_bm, ch_bm, C0_bm, stop=None, max_iter: int = 100, threads: int = 1):
    """The reference's CPU algorithm: its OpenMP evaluation units
    (src/evaluation_kernels_cpu.cpp:8-662, validation, forward over the full
    (M, N, B) cache, atomic reverse sweep) compiled from its own sources into
    oracle/_ref/libsrref.so, driven as symbolic_torch/optimization.py:9-112
    drives them (root-row gather, mean-squared error, dense cache gradient,
    masked dL/dC) by LBFGS_batch in torch CPU FP32.  stop (B,) float64: a
    problem whose loss reaches it gets a zero gradient from then on, so the
    reference's own done test (gradient norm below tol) ends it: the same
    stop rule as the device step.

    A backward error (204, a divisor exactly 0 under a nonzero adjoint) sets
    the reference's one error flag, after which its sweep skips every
    remaining (point, expression) of the batch and its optimize raises.  To
    keep the batch running as the device path does (errors per problem),
    the closure then finds the failing problems with one backward call each,
    zeroes their upstream gradient and their dL/dC from then on (so the
    done test ends them) and repeats the batch backward.
    Returns (C (B, M), seconds, closures, failed problems)."""
    import ctypes

    import numpy as np
    import torch

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    from cudasymbolicregression_amd.lbfgs_batch import LBFGS_batch

    ref = oracle.ref()
    torch.set_num_threads(threads)
    ctypes.CDLL("libgomp.so.1").omp_set_num_threads(threads)
    B, M = ops_bm.shape
    N, n_x = X.shape
    ops = np.ascontiguousarray(ops_bm.T)
    ch = np.ascontiguousarray(ch_bm.transpose(1, 0, 2))
    mask = torch.from_numpy(ops == 1)
    root = np.maximum((ops != 0).sum(0) - 1, 0)
    bidx = np.arange(B)
    Xc = np.ascontiguousarray(X, np.float32)
    Yt = torch.from_numpy(np.ascontiguousarray(Y, np.float32))
    stop_t = None if stop is None else torch.from_numpy(np.asarray(stop, np.float64))
    x = torch.from_numpy(np.ascontiguousarray(C0_bm.T, np.float32)).clone().requires_grad_()
    opt = LBFGS_batch([x], max_iter=max_iter, history_size=10, tolerance_grad=1e-7, line_search_fn="strong_wolfe")
    cache = np.empty((M, N, B), np.float32)
    gcache = np.empty((M, N, B), np.float32)
    gC = np.empty((M, B), np.float32)
    gX = np.empty((N, n_x), np.float32)
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    calls = [0]
    failed = np.zeros(B, bool)

    def find_failing(gup):
        out = np.zeros(B, bool)
        for b in np.nonzero(~failed & (gup != 0).any(1))[0]:
            cb = np.ascontiguousarray(cache[:, :, b:b + 1])
            gb = np.zeros((M, N, 1), np.float32)
            gb[root[b], :, 0] = gup[b]
            ob, cb_ = np.ascontiguousarray(ops[:, b:b + 1]), np.ascontiguousarray(ch[:, b:b + 1])
            gc1, gx1 = np.zeros((M, 1), np.float32), np.zeros((N, n_x), np.float32)
            out[b] = ref.ref_backward_f32(P(gb), P(cb), P(ob), P(cb_), M, 1, N, n_x, P(gc1), P(gx1)) != 0
        return out

    def closure():
        opt.zero_grad(set_to_none=True)
        calls[0] += 1
        C = np.ascontiguousarray(x.detach().numpy())
        ref.ref_validate(P(ops), P(ch), M, B)
        ref.ref_forward_f32(P(Xc), N, n_x, P(ops), P(ch), P(C), M, B, P(cache))
        yr = torch.from_numpy(cache[root, :, bidx].copy()).requires_grad_()
        loss = ((yr - Yt) ** 2).mean(1)
        loss.sum().backward()
        gup = yr.grad.numpy()
        gup[failed] = 0.0
        while True:
            gcache.fill(0.0)
            gcache[root, :, bidx] = gup
            gC.fill(0.0)
            gX.fill(0.0)
            if ref.ref_backward_f32(P(gcache), P(cache), P(ops), P(ch), M, B, N, n_x, P(gC), P(gX)) == 0:
                break
            new = find_failing(gup)
            if not new.any():
                raise RuntimeError("reference backward error that no single problem reproduces")
            failed[new] = True
            gup[new] = 0.0
        g = torch.from_numpy(gC.copy())
        g[~mask] = 0.0
        g[:, torch.from_numpy(failed)] = 0.0
        if stop_t is not None:
            g[:, loss.detach().double() <= stop_t] = 0.0
        x.grad = g
        return loss.detach()

    t0 = time.perf_counter()
    with np.errstate(all="ignore"):
        opt.step(closure)
    return x.detach().numpy().T.copy(), time.perf_counter() - t0, calls[0], int(failed.sum())


def cpu_baseline(w, sample: int):
    """The reference's CPU path on a bounded sample of the plain workload
    (randn starts) on this host's cores, timed twice: with the device step's
    stop rule (value) and as the reference runs it, to max_iter or its own
    tolerances.  Without the reference's compiled units (oracle/_ref), the
    oracle's single-thread restatement (oracle/lbfgs_oracle.py +
    oracle/eval_oracle.c) is timed instead."""
    import numpy as np
    import torch

    from cudasymbolicregression_amd import warmstart

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    Xn = w["X"].cpu().numpy()
    ops = w["ops"][:sample].cpu().numpy()
    ch = w["ch"][:sample].cpu().numpy()
    Y = w["Y"][:sample].cpu().numpy()
    C0 = w["C0"][:sample].cpu().numpy()
    judge = lambda C: int((warmstart.rel_mse_fp64(  # noqa: E731
        w["X"], w["Y"][:sample], w["ops"][:sample], w["ch"][:sample],
        torch.from_numpy(np.ascontiguousarray(C)).to(w["X"].device)) <= 1e-6).sum())
    where = f"{cpu_model()}"
    if oracle.ref_available():
        thr = host_threads()
        stop = w["stop"][:sample].cpu().numpy()
        torch_threads = torch.get_num_threads()
        try:
            C_s, dt_s, calls_s, fail_s = reference_cpu_fit(Xn, Y, ops, ch, C0, stop=stop, threads=thr)
            C_f, dt_f, calls_f, fail_f = reference_cpu_fit(Xn, Y, ops, ch, C0, stop=None, threads=thr)
        finally:
            torch.set_num_threads(torch_threads)
        fit_s, fit_f = judge(C_s), judge(C_f)
        return dict(value=round(fit_s / dt_s, 3), unit="expressions fitted/s", cores=thr, kind="reference",
                    cpu_model=where,
                    sample=f"{sample} problems of the same workload (randn starts, max_iter 100, strong Wolfe): the "
                           f"reference's evaluation units (src/evaluation_kernels_cpu.cpp, OpenMP, {thr} threads) "
                           f"under LBFGS_batch in torch CPU FP32 ({thr} threads); with the device step's stop rule "
                           f"{dt_s:.1f} s, {fit_s} fitted, {calls_s} batched closures",
                    full_fit=dict(value=round(fit_f / dt_f, 3), seconds=round(dt_f, 2), fitted=fit_f,
                                  closures=calls_f, backward_error_problems=fail_f,
                                  note="the reference's own stop rules only (max_iter, tol_grad)"),
                    backward_error_problems=fail_s,
                    error_note="the reference's optimize raises on a backward error anywhere in the batch; these "
                               "problems were frozen instead (as the device path reports errors per problem)",
                    seconds=round(dt_s, 2), fitted=fit_s, problems=sample)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import lbfgs_oracle

    t0 = time.perf_counter()
    with np.errstate(all="ignore"):
        C, _, _, calls = lbfgs_oracle.optimize(Xn, Y, ops, ch, C0, max_iter=100)
    dt = time.perf_counter() - t0
    fitted = judge(C)
    return dict(value=round(fitted / dt, 3), unit="expressions fitted/s", cores=1, kind="port", cpu_model=where,
                sample=f"{sample} problems of the same workload, randn starts, full fits (max_iter 100), {dt:.1f} s, "
                       f"{fitted} fitted, {calls} batched closures; oracle/lbfgs_oracle.py + oracle/eval_oracle.c, "
                       f"single thread (the reference's compiled CPU units are absent here)")


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
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--cpu-sample", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-evaluate", action="store_true")
    ap.add_argument("--no-ragged", action="store_true")
    ap.add_argument("--no-c3", action="store_true")
    ap.add_argument("--no-ceiling", action="store_true", help="no fit from the true constants")
    ap.add_argument("--no-probe", action="store_true",
                    help="no transcendental-rate probe: price a transcendental at the last measured T")
    ap.add_argument("--no-fp64-report", action="store_true")
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
    global OPCOST
    trans = None
    if not args.no_probe:
        trans = measure_trans_cost(dev)
        OPCOST = opcost_table(trans["trans_cost"])
        log(f"transcendental cost: {trans['trans_cost']} FMA slots ({trans})")
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
    roofline = dict(bound="valu", achieved=round(achieved, 4), peak=FP32_PEAK_TFLOPS, unit="TFLOP/s",
                    frac=round(achieved / FP32_PEAK_TFLOPS, 5), traffic=FIT_TRAFFIC_PROFILED,
                    traffic_source="profiles/fit_hbm_traffic.md: rocprofv3 FETCH_SIZE/WRITE_SIZE of the plain-start "
                                   "fit call (the same kernels on the same problems), not measured live",
                    kernel="sr_lbfgs_fit: the k_fit bucket kernels of one call (concurrent)",
                    flop_per_launch=flops, frac_unpacked_lane_peak=round(flops / (fit_ms / 1e3) / VALU_LANE_PEAK, 5),
                    slot_unit="one FP32 VALU lane operation; a transcendental counts trans_cost slots",
                    trans_cost=trans["trans_cost"] if trans else T_COST,
                    trans_cost_source="measured in this run (sr_probe_valu_rate)" if trans else "last measurement")

    # ---- ceiling: the same fit from the true constants
    ceiling = None
    if not args.no_ceiling:
        ms_t, (C_t, loss_t, iters_t, err_t, _, ncl_t) = timed(lambda: fit(w, w["Ctrue"]), args.steps, args.warmup)
        fit_t = int((warmstart.rel_mse_fp64(X, Y, ops, ch, C_t) <= 1e-6).sum())
        ceiling = {"ms": round(ms_t, 3), "fitted": fit_t, "fitted_per_s": round(fit_t / (ms_t / 1e3), 1),
                   "over_plain": round(fit_t / (ms_t / 1e3) / plain_value, 3) if plain_value > 0 else None,
                   "mean_iters": round(float(iters_t.double().mean()), 2), "closures": int(ncl_t.sum()),
                   "problems_without_error": int((err_t == 0).sum()),
                   "note": "fit started from the true constants: what a perfect warm start would reach"}
        log(f"ceiling (true constants): {ms_t:.2f} ms/step, {fit_t} fitted")
    hist = {"plain": closure_histogram(ncl_p)}
    if warm is not None:
        hist["warm"] = closure_histogram(ncl_w)
    if ceiling is not None:
        hist["true_constants"] = closure_histogram(ncl_t)

    ev = None
    if not args.no_evaluate:
        ev = bench_evaluate(dev, steps=5, fp64_report=not args.no_fp64_report)
        log(f"evaluate (65536 x 1024): {ev['ms']:.3f} ms, {ev['hbm']['achieved_GBps']} GB/s, VALU frac {ev['valu']['frac']}")

    fc = None
    if not args.no_evaluate:
        fc = bench_full_cache(dev, steps=3)
        log(f"full-cache evaluate (65536 x 1024): forward {fc['forward']['ms']:.2f} ms ({fc['forward']['achieved_GBps']} GB/s), "
            f"backward {fc['backward']['ms']:.2f} ms ({fc['backward']['achieved_GBps']} GB/s)")

    c3 = None
    if not args.no_c3:
        c3 = bench_c3(dev, steps=2)
        log(f"C3 (16384 candidates, shared noisy target): {c3['ms']:.1f} ms, {c3['fits_per_s']:.0f} fits/s, "
            f"best FP64 MSE {c3['best_mse_fp64']:.3g}")

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
            "ceiling": ceiling,
            "closures_per_problem": hist,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "trans_cost": trans,
        "evaluate": ev,
        "evaluate_full_cache": fc,
        "shared_target_c3": c3,
        "ragged": rg,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
