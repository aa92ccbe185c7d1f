"""Run-to-run bit stability, thread safety and large-batch work lists.

* grad_X of the fused evaluator and of the full-cache backward is summed
  without float atomics (wave-private partial rows, then a fixed-order FP64
  sum), so two calls give the same bits; checked for few (3) and many (6,
  10 > the 8 register slots) variables, and against the FP64 oracle.
* sr_lbfgs_fit and sr_fused_root_* fork side streams per caller stream; two
  host threads calling at once on their own streams get the results of
  sequential calls.
* Batches above 4096 problems (more than 64 classification runs, so the
  work-list scan carries across waves) give the same per-problem results as
  the same problems in slices of at most 4096, and every bucket list is a
  permutation of its problems.
"""
import ctypes
import threading

import numpy as np
import pytest
import torch

import helpers
import oracle

pytestmark = pytest.mark.gpu


def _dev():
    return torch.device("cuda:0")


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(_dev())


def _bits(t):
    return {torch.float32: lambda: t.view(torch.int32), torch.float64: lambda: t.view(torch.int64)}.get(
        t.dtype, lambda: t)()


def _g2_case(seed, B, M, n_x, N):
    rng = np.random.default_rng(seed)
    ops, ch = helpers.random_batch(seed, B, M, n_x, unary=helpers.G2_UNARY, binary=helpers.G2_BINARY, depth=6)
    C = helpers.constants_like(ops, rng)
    X = rng.uniform(-3, 3, (N, n_x)).astype(np.float32)
    dy = rng.standard_normal((B, N)).astype(np.float32)
    return X, ops, ch, C, dy


@pytest.mark.parametrize("n_x,N", [(3, 512), (6, 300), (3, 9000)])
def test_fused_grad_X_is_bit_stable_and_accurate(n_x, N):
    """(3, 9000): 27000 values per row do not fit the LDS rows, so the wave
    rows live in the workspace."""
    from cudasymbolicregression_amd import fused

    print(f"case fused grad_X n_x={n_x} N={N}", flush=True)
    X, ops, ch, C, dy = _g2_case(7 + n_x, 400, 32, n_x, N)
    outs = [fused.fused_root(_t(X), _t(ops), _t(ch), _t(C), _t(dy), mode=0, want_grad_X=True, raise_errors=False)
            for _ in range(3)]
    torch.cuda.synchronize()
    for o in outs[1:]:
        assert torch.equal(_bits(o["grad_X"]), _bits(outs[0]["grad_X"]))
        assert torch.equal(_bits(o["grad_C"]), _bits(outs[0]["grad_C"]))
        assert torch.equal(_bits(o["y"]), _bits(outs[0]["y"]))
    r32 = oracle.fused_root(X, ops, ch, C, dy, mode=0)
    r64 = oracle.fused_root(X.astype(np.float64), ops, ch, C.astype(np.float64), dy.astype(np.float64), mode=0)
    good = (r32["err"] == 0) & (r64["err"] == 0)
    # grad_X sums every expression: compare on the expressions without errors
    if not good.all():
        keep = np.nonzero(good)[0]
        ops, ch, C, dy = ops[keep], ch[keep], C[keep], dy[keep]
        o = fused.fused_root(_t(X), _t(ops), _t(ch), _t(C), _t(dy), mode=0, want_grad_X=True, raise_errors=False)
        r32 = oracle.fused_root(X, ops, ch, C, dy, mode=0)
        r64 = oracle.fused_root(X.astype(np.float64), ops, ch, C.astype(np.float64), dy.astype(np.float64), mode=0)
    else:
        o = outs[0]
    ok, msg = helpers.check_vs_f64(o["grad_X"].cpu().numpy(), r32["grad_X"], r64["grad_X"], atol=1e-4)
    print(f"grad_X n_x={n_x} N={N}: {msg}", flush=True)
    assert ok, msg


@pytest.mark.parametrize("n_x", [3, 10])
def test_full_cache_grad_X_is_bit_stable_and_accurate(n_x):
    """Full-cache backward: n_x = 10 exceeds the 8 variable registers, so the
    variable adjoints are summed after the sweep, 8 variables at a time."""
    from cudasymbolicregression_amd.evaluation import evaluate

    print(f"case full-cache grad_X n_x={n_x}", flush=True)
    X, ops, ch, C, _ = _g2_case(30 + n_x, 64, 16, n_x, 150)
    Xt = _t(X).requires_grad_()
    res = []
    for _ in range(2):
        Xt.grad = None
        Ct = _t(C).requires_grad_()
        y = evaluate(Xt, _t(ops), _t(ch), Ct)  # (B, M, N)
        root = torch.from_numpy(np.maximum((ops != 0).sum(1) - 1, 0)).to(_dev())
        yr = y.gather(1, root.view(-1, 1, 1).expand(-1, 1, y.shape[2])).squeeze(1)
        w = torch.from_numpy(np.random.default_rng(1).standard_normal(yr.shape).astype(np.float32)).to(_dev())
        ok_rows = torch.isfinite(yr).all(1)
        (torch.where(ok_rows[:, None], yr, torch.zeros_like(yr)) * w).sum().backward()
        res.append((Xt.grad.clone(), Ct.grad.clone()))
    torch.cuda.synchronize()
    assert torch.equal(_bits(res[0][0]), _bits(res[1][0]))
    assert torch.equal(_bits(res[0][1]), _bits(res[1][1]))
    # oracle: the root-only upstream gradient on the rows kept above
    keep = ok_rows.cpu().numpy()
    dy = np.where(keep[:, None], np.random.default_rng(1).standard_normal(yr.shape).astype(np.float32), 0)
    dy = dy.astype(np.float32)
    r32 = oracle.fused_root(X, ops, ch, C, dy, mode=0)
    r64 = oracle.fused_root(X.astype(np.float64), ops, ch, C.astype(np.float64), dy.astype(np.float64), mode=0)
    if (r32["err"] == 0).all() and (r64["err"] == 0).all():
        ok, msg = helpers.check_vs_f64(res[0][0].cpu().numpy(), r32["grad_X"], r64["grad_X"], atol=1e-4)
        print(f"full-cache grad_X n_x={n_x}: {msg}", flush=True)
        assert ok, msg


def _fit_problem(seed, B, N=128):
    import cudasymbolicregression_amd as st
    from cudasymbolicregression_amd import warmstart

    g = st.ProbabilisticContextFreeGrammar(helpers.G2, "E", 32, 3, torch.device("cpu"))
    gen = torch.Generator().manual_seed(seed)
    X = (torch.rand(N, 3, generator=gen) * 6 - 3).to(_dev())
    ops, ch, Ct, Y = warmstart.sample_problems(g, X, B, seed=seed)
    torch.manual_seed(seed + 1)
    C0 = torch.where(ops == 1, torch.randn_like(Ct), Ct).contiguous()
    return X, Y, ops, ch, C0


def test_concurrent_calls_from_two_threads_match_sequential():
    from cudasymbolicregression_amd import fused
    from cudasymbolicregression_amd.optimization import _fit

    print("case two threads", flush=True)
    probs = [_fit_problem(40 + i, 1500) for i in range(2)]
    seq = []
    for X, Y, ops, ch, C0 in probs:
        f = _fit(X, Y, ops, ch, C0, 30, 10, 1e-7, "strong_wolfe")
        e = fused.fused_root(X, ops, ch, C0, Y, mode=1, raise_errors=False)
        torch.cuda.synchronize()
        seq.append((f[0].clone(), f[1].clone(), e["loss"].clone(), e["grad_C"].clone()))
    out = [None, None]
    errs = []

    def work(i):
        try:
            X, Y, ops, ch, C0 = probs[i]
            s = torch.cuda.Stream(device=_dev())
            s.wait_stream(torch.cuda.current_stream(_dev()))
            with torch.cuda.stream(s):
                for _ in range(3):
                    f = _fit(X, Y, ops, ch, C0, 30, 10, 1e-7, "strong_wolfe")
                    e = fused.fused_root(X, ops, ch, C0, Y, mode=1, raise_errors=False)
                s.synchronize()
            out[i] = (f[0], f[1], e["loss"], e["grad_C"])
        except Exception as exc:  # reported below
            errs.append(exc)

    th = [threading.Thread(target=work, args=(i,)) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errs, errs
    for i in range(2):
        for a, b in zip(out[i], seq[i]):
            assert torch.equal(_bits(a), _bits(b))


def _lists_from_workspace(ws, B):
    """The five bucket lists and their counts at the start of a fit or fused
    workspace: lists [5][B] int32, then counts [5]."""
    raw = ws[: 5 * B * 4 + 20].cpu().numpy().view(np.int32)
    counts = raw[5 * B: 5 * B + 5]
    return [raw[k * B: k * B + counts[k]] for k in range(5)], counts


def test_large_batches_equal_slices_and_lists_are_permutations():
    from cudasymbolicregression_amd import _lib, fused
    from cudasymbolicregression_amd.optimization import _fit

    B = 20000
    print(f"case large batch B={B}", flush=True)
    X, Y, ops, ch, C0 = _fit_problem(50, B, N=96)
    # the whole batch at once, with the fused evaluator's workspace kept
    lib = _lib.hip()
    M, N, n_x = ops.shape[1], X.shape[0], X.shape[1]
    y = torch.empty((B, N), dtype=torch.float32, device=_dev())
    loss = torch.empty(B, dtype=torch.float64, device=_dev())
    gC = torch.empty((B, M), dtype=torch.float32, device=_dev())
    gX = torch.empty((N, n_x), dtype=torch.float32, device=_dev())
    err = torch.empty(B, dtype=torch.int32, device=_dev())
    flags = torch.empty(2, dtype=torch.int32, device=_dev())
    for want_gx in (0, 1):
        ws = torch.empty(lib.sr_fused_workspace_bytes(N, n_x, B, M, want_gx, 0), dtype=torch.uint8, device=_dev())
        assert lib.sr_fused_root_f32(_lib.ptr(X), N, n_x, _lib.ptr(ops), _lib.ptr(ch), _lib.ptr(C0), M, B,
                                     _lib.ptr(Y), 1, 0, _lib.ptr(y), _lib.ptr(loss), _lib.ptr(gC),
                                     _lib.ptr(gX) if want_gx else None, _lib.ptr(err), _lib.ptr(flags),
                                     _lib.ptr(ws), ws.numel(), _lib.stream_of(_dev())) == 0
        torch.cuda.synchronize()
        lists, counts = _lists_from_workspace(ws, B)
        assert int(counts.sum()) == B
        allv = np.sort(np.concatenate(lists))
        assert np.array_equal(allv, np.arange(B))
        for lst in lists:  # each list in item order
            assert np.all(np.diff(lst) > 0)
        whole = (y.clone(), loss.clone(), gC.clone(), err.clone())
        for s0 in range(0, B, 4096):
            sl = slice(s0, min(B, s0 + 4096))
            part = fused.fused_root(X, ops[sl], ch[sl], C0[sl], Y[sl], mode=1, want_grad_X=bool(want_gx),
                                    raise_errors=False)
            assert torch.equal(_bits(part["y"]), _bits(whole[0][sl]))
            assert torch.equal(_bits(part["loss"]), _bits(whole[1][sl]))
            assert torch.equal(_bits(part["grad_C"]), _bits(whole[2][sl]))
            assert torch.equal(part["err"], whole[3][sl])
    # the fit: whole batch vs slices
    full = _fit(X, Y, ops, ch, C0, 20, 10, 1e-7, "strong_wolfe")
    for s0 in range(0, B, 4096):
        sl = slice(s0, min(B, s0 + 4096))
        part = _fit(X, Y[sl], ops[sl], ch[sl], C0[sl], 20, 10, 1e-7, "strong_wolfe")
        for a, b in zip(part[:4], full[:4]):
            assert torch.equal(_bits(a), _bits(b[sl]))
