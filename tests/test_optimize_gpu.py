"""GPU tests of `optimize` (device-resident L-BFGS) against the reference's
known answers and against the CPU restatement of LBFGS_batch (oracle).

Stated parity tolerance: per problem |MSE_gpu - MSE_oracle| <= 1e-3 *
MSE_oracle + 1e-7 (helpers.compare_fit_to_oracle).  The oracle runs with the
reference's FP32 arithmetic (its trajectory pinned bit for bit to the
reference's own optimizer by tests/test_lbfgs_golden.py), with the device's
FP64 sums, inner products and Wolfe tests, and with the latter on X scaled by
1 +- 2^-21 and with its closure values perturbed.  On the problems the
algorithm determines (every perturbed run ends within the tolerance) the
device must meet the tolerance against the device-precision run, all but 2%
(at least one, listed: chaotic trajectories the probes miss), and against
the reference-precision run wherever the two precisions agree.  Problems
whose outcome hinges on the last ulp of the node arithmetic (where device
transcendentals and the host libm may differ) are counted (at most 15%) and
listed."""
import numpy as np
import pytest
import torch

import helpers
import lbfgs_oracle

pytestmark = pytest.mark.gpu


def _dev():
    return torch.device("cuda:0")


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(_dev())


@pytest.mark.parametrize("line_search", ["strong_wolfe", "backtracking"])
def test_optimize_linear_known_answer(line_search):
    """y = c1*X0 + c2 (reference tests/test_optimization.py:10-80)."""
    import cudasymbolicregression_amd as st

    torch.manual_seed(0)
    B, N, M = 2, 100, 5
    c1 = torch.tensor([2.5, -1.5], device=_dev()).view(B, 1)
    c2 = torch.tensor([0.5, 3.0], device=_dev()).view(B, 1)
    X = torch.randn(N, 1, device=_dev())
    Y = c1 * X[:, 0].T + c2
    Ops = torch.full((B, M), 0, dtype=torch.int64, device=_dev())
    Ops[:, 0] = 50
    Ops[:, 1] = 1
    Ops[:, 2] = 42
    Ops[:, 3] = 1
    Ops[:, 4] = 40
    Ch = torch.full((B, M, 2), -1, dtype=torch.int64, device=_dev())
    Ch[:, 2, 0], Ch[:, 2, 1] = 0, 1
    Ch[:, 4, 0], Ch[:, 4, 1] = 2, 3
    C0 = st.create_constants(Ops, method="randn")
    Cf, Yf = st.optimize(X, Y, Ops, Ch, C0, line_search_fn=line_search, max_iter=300)
    assert torch.allclose(Cf[:, 1], c1.squeeze(), atol=1e-3)
    assert torch.allclose(Cf[:, 3], c2.squeeze(), atol=1e-3)
    mask = Ops == 1
    assert torch.equal(Cf[~mask], C0[~mask])
    assert torch.allclose(Yf, Y, atol=1e-3)


def test_optimize_sin_batch_reduces_loss():
    """c1*sin(X0) + c2 over 2000 problems (reference tests/test_optimization.py:82-169)."""
    import cudasymbolicregression_amd as st

    torch.manual_seed(1)
    B, N, M = 2000, 50, 6
    c1 = torch.randn(B, 1, device=_dev()) * 2
    c2 = torch.randn(B, 1, device=_dev())
    X = torch.rand(N, 1, device=_dev()) * 4 - 2
    Y = c1 * torch.sin(X[:, 0]).T + c2
    Ops = torch.zeros((B, M), dtype=torch.int64, device=_dev())
    Ops[:, :5] = torch.tensor([50, 10, 1, 42, 1], device=_dev())
    Ops[:, 5] = 40
    Ch = torch.full((B, M, 2), -1, dtype=torch.int64, device=_dev())
    Ch[:, 1, 0] = 0
    Ch[:, 3, 0], Ch[:, 3, 1] = 1, 2
    Ch[:, 5, 0], Ch[:, 5, 1] = 3, 4
    C0 = st.create_constants(Ops)
    y0 = st.evaluation.evaluate(X, Ops, Ch, C0)[:, 5, :]
    init = ((y0 - Y) ** 2).mean(1)
    Cf, Yf, info = st.optimize(X, Y, Ops, Ch, C0, return_info=True)
    final = ((Yf - Y) ** 2).mean(1)
    assert (final < 0.1 * init + 1e-10).all()
    assert torch.allclose(Cf[:, 2], c1.squeeze(), atol=1e-3)
    assert torch.allclose(Cf[:, 4], c2.squeeze(), atol=1e-3)
    assert (info["err"] == 0).all()


@pytest.mark.parametrize("seed,M", [(0, 16), (1, 32), (2, 64)])
def test_optimize_matches_oracle_lbfgs(seed, M):
    """Every stable problem matches the oracle's final MSE to 1e-3 relative
    (helpers.compare_fit_to_oracle); unstable ones are counted and listed."""
    import cudasymbolicregression_amd as st

    rng = np.random.default_rng(seed)
    B, N, n_x = 96, 64, 2
    ops, ch = helpers.random_batch(seed, B, M, n_x, unary=[10, 11, 12, 14], binary=[40, 41, 42],
                                   leaf_const_p=0.4, depth=5)
    X = rng.uniform(-2, 2, (N, n_x)).astype(np.float32)
    C_true = helpers.constants_like(ops, rng, scale=1.0)
    Y = lbfgs_oracle.oracle.fused_root(X, ops, ch, C_true, np.zeros((B, N), np.float32), mode=0)["y"]
    keep = np.isfinite(Y).all(1)
    ops, ch, Y, C_true = ops[keep], ch[keep], Y[keep], C_true[keep]
    C0 = helpers.constants_like(ops, rng, scale=1.0)
    fits = helpers.oracle_fits(X, Y, ops, ch, C0, max_iter=50)
    Cg, Yg, info = st.optimize(_t(X), _t(Y), _t(ops), _t(ch), _t(C0), max_iter=50, return_info=True,
                               raise_errors=False)
    with np.errstate(all="ignore"):
        mse_g = ((Yg.cpu().numpy().astype(np.float64) - Y) ** 2).mean(1)
    msg = helpers.compare_fit_to_oracle(mse_g, fits, info["err"].cpu().numpy(), f"M={M}")
    print(msg, flush=True)


def test_optimize_raises_div_by_zero():
    import cudasymbolicregression_amd as st

    # 1 / (X0 - X0) has an exact zero denominator at every point -> 204
    X = torch.randn(16, 1, device=_dev())
    Y = torch.randn(1, 16, device=_dev())
    Ops = torch.tensor([[1, 50, 50, 41, 43]], dtype=torch.int64, device=_dev())
    Ch = torch.full((1, 5, 2), -1, dtype=torch.int64, device=_dev())
    Ch[0, 3, 0], Ch[0, 3, 1] = 1, 2
    Ch[0, 4, 0], Ch[0, 4, 1] = 0, 3
    C0 = torch.ones(1, 5, device=_dev())
    with pytest.raises(RuntimeError, match="a/b where b = 0"):
        st.optimize(X, Y, Ops, Ch, C0)
    Cf, Yf, info = st.optimize(X, Y, Ops, Ch, C0, raise_errors=False, return_info=True)
    assert int(info["err"][0]) == 204


def _bucket_batch(seed):
    """Programs in every length bucket of the fit kernels (<= 16, <= 32,
    <= 64 and > 64 nodes) plus a 64-constant sum (65 live slots incl. the
    shared non-constant one), padded to M = 128."""
    rng = np.random.default_rng(seed)
    M, n_x = 128, 2
    rows = []
    for lo, hi, depth, count in ((3, 16, 4, 10), (17, 32, 6, 10), (33, 64, 7, 10), (65, 127, 8, 8)):
        for _ in range(count):
            o, c = helpers.random_postfix(rng, M, n_x, max_len=hi, min_len=lo, unary=[10, 11, 14],
                                          binary=[40, 41, 42], leaf_const_p=0.4, hard_const_p=0.0, depth=depth)
            rows.append((o, c))
    o = np.zeros(M, np.int64)
    o[0] = 1
    for k in range(1, 64):
        o[2 * k - 1] = 1
        o[2 * k] = 40
    rows.append((o, helpers.children_of(o)))
    ops = np.stack([r[0] for r in rows])
    ch = np.stack([r[1] for r in rows])
    return ops, ch, rng


def test_optimize_length_buckets_match_oracle():
    """Short runs (3 iterations) keep the two L-BFGS trajectories close, so
    the fitted constants of every bucket are compared directly:
    |C_gpu - C_oracle| <= 1e-3 * (1 + |C_oracle|) (helpers.compare_constants_to_oracle:
    against the device-precision oracle, and against the reference-precision
    one wherever the two precisions agree), and exact agreement on
    non-constant slots."""
    import cudasymbolicregression_amd as st

    ops, ch, rng = _bucket_batch(11)
    B, M = ops.shape
    lengths = (ops != 0).sum(1)
    print("lengths per bucket:", [(lengths <= 16).sum(), ((lengths > 16) & (lengths <= 32)).sum(),
                                  ((lengths > 32) & (lengths <= 64)).sum(), (lengths > 64).sum()])
    N = 70  # a full and a partial tile of 64 points
    X = rng.uniform(-2, 2, (N, 2)).astype(np.float32)
    C_true = helpers.constants_like(ops, rng, scale=1.0)
    Y = lbfgs_oracle.oracle.fused_root(X, ops, ch, C_true, np.zeros((B, N), np.float32), mode=0)["y"]
    C0 = helpers.constants_like(ops, rng, scale=1.0)
    C0[ops != 1] = 0.25  # non-constant slots must come back unchanged
    fits = helpers.oracle_fits(X, Y, ops, ch, C0, max_iter=3)
    C_o = fits["C32"]
    Cg, Yg, info = st.optimize(_t(X), _t(Y), _t(ops), _t(ch), _t(C0), max_iter=3, return_info=True,
                               raise_errors=False)
    Cg = Cg.cpu().numpy()
    assert (info["err"].cpu().numpy() == 0).all()
    nonc = ops != 1
    np.testing.assert_array_equal(Cg[nonc], C_o[nonc])  # C_init, or NaN where the reference's go NaN
    fin = np.isfinite(C_o).all(1) & np.isfinite(Cg).all(1)
    print(helpers.compare_constants_to_oracle(Cg, fits, "buckets, 3 iterations"), flush=True)
    print("per-problem max |dC| vs the reference-precision oracle:", np.round(np.abs(Cg - C_o).max(1), 6).tolist())
    assert fin.sum() >= B * 3 // 4
    close = (np.abs(Cg - C_o) <= 1e-3 * (1 + np.abs(C_o))).all(1)
    # the 64-constant sum: every constant gets the same gradient
    assert close[-1]


@pytest.mark.parametrize("line_search", ["strong_wolfe", "backtracking"])
def test_stationary_exit_matches_running_every_iteration(line_search):
    """Problems whose history is all zero pairs sit at an exact fixed point of
    the iteration; ending them there must give the same constants, loss,
    iteration count and error code, bit for bit, as running all max_iter
    iterations (badly scaled G2 problems started near their optimum stall
    like that under the reference's unscaled first step)."""
    import cudasymbolicregression_amd as st
    from cudasymbolicregression_amd import warmstart
    from cudasymbolicregression_amd.optimization import _fit

    g = st.ProbabilisticContextFreeGrammar(helpers.G2, "E", 64, 3, torch.device("cpu"))
    gen = torch.Generator().manual_seed(0)
    X = (torch.rand(512, 3, generator=gen) * 6 - 3).to(_dev())
    ops, ch, Ct, Y = warmstart.sample_problems(g, X, 4096, seed=21)
    torch.manual_seed(22)
    C0 = (Ct + 0.01 * torch.randn_like(Ct) * (ops == 1)).contiguous()
    res = {}
    for on in (0, 1):
        res[on] = _fit(X, Y, ops, ch, C0, 100, 10, 1e-7, line_search, stationary_exit=bool(on))
        torch.cuda.synchronize()
    (C_a, l_a, it_a, e_a, f_a, n_a), (C_b, l_b, it_b, e_b, f_b, n_b) = res[0], res[1]
    assert torch.equal(C_a.view(torch.int32), C_b.view(torch.int32))
    assert torch.equal(l_a.view(torch.int64), l_b.view(torch.int64))
    assert torch.equal(it_a, it_b) and torch.equal(e_a, e_b) and torch.equal(f_a, f_b)
    saved = int(n_a.sum() - n_b.sum())
    ended = int((n_b < n_a).sum())
    print(f"{line_search}: {ended} problems ended early, {saved} of {int(n_a.sum())} closures saved", flush=True)
    assert (n_b <= n_a).all()
    assert ended > 0


def _bits(t):
    """Bit pattern of a tensor (NaN-safe equality)."""
    return {torch.float32: lambda: t.view(torch.int32), torch.float64: lambda: t.view(torch.int64)}.get(
        t.dtype, lambda: t)()


def _ragged_case(seed, B=24, buckets=True):
    """Problems with their own datasets: N_b in [20, 300], targets from random
    true constants.  buckets=True: programs from every length bucket (padded
    to 128); else depth-5 programs of the dense oracle comparison (M = 32)."""
    if buckets:
        ops, ch, rng = _bucket_batch(seed)
        ops, ch = ops[:B], ch[:B]
    else:
        rng = np.random.default_rng(seed)
        ops, ch = helpers.random_batch(seed, B, 32, 2, unary=[10, 11, 12, 14], binary=[40, 41, 42],
                                       leaf_const_p=0.4, depth=5)
    n = rng.integers(20, 301, size=B)
    Xs = [rng.uniform(-2, 2, (int(k), 2)).astype(np.float32) for k in n]
    C_true = helpers.constants_like(ops, rng, scale=1.0)
    Ys = [lbfgs_oracle.oracle.fused_root(Xs[b], ops[b:b + 1], ch[b:b + 1], C_true[b:b + 1],
                                         np.zeros((1, int(n[b])), np.float32), mode=0)["y"][0] for b in range(B)]
    C0 = helpers.constants_like(ops, rng, scale=1.0)
    return ops, ch, n, Xs, Ys, C0


def test_optimize_ragged_equals_per_problem_fits():
    """One ragged launch gives, bit for bit, what fitting every problem alone
    on its own dataset gives."""
    import cudasymbolicregression_amd as st

    ops, ch, n, Xs, Ys, C0 = _ragged_case(31)
    B = ops.shape[0]
    C_r, info = st.optimize_ragged(_t(np.concatenate(Xs)), _t(np.concatenate(Ys)), _t(n), _t(ops), _t(ch), _t(C0),
                                   max_iter=20, raise_errors=False)
    torch.cuda.synchronize()
    for b in range(B):
        Cb, _, ib = st.optimize(_t(Xs[b]), _t(Ys[b][None, :]), _t(ops[b:b + 1]), _t(ch[b:b + 1]), _t(C0[b:b + 1]),
                                max_iter=20, return_info=True, raise_errors=False)
        assert torch.equal(_bits(C_r[b:b + 1]), _bits(Cb)), b
        assert torch.equal(_bits(info["loss"][b:b + 1]), _bits(ib["loss"])), b
        assert int(info["iters"][b]) == int(ib["iters"][0])


def test_optimize_ragged_matches_oracle():
    """Against the CPU oracle fitting each problem on its own dataset (the
    program family of the shared-X comparison above; the per-problem path
    itself is pinned bit for bit by the test before).

    After 3 iterations the constants agree within 1e-3 (1 + |C|)
    (helpers.compare_constants_to_oracle).  After 50 iterations the stated
    optimize tolerance |MSE_gpu - MSE_oracle| <= 1e-3 MSE_oracle + 1e-7
    holds as in helpers.compare_fit_to_oracle, widened by how far the two
    FP32 evaluators are apart on the problem
    (e_b = mean_n (y_gpu(C_oracle) - y_oracle(C_oracle))^2 adds
    2 sqrt(e_b MSE_oracle) + e_b)."""
    import cudasymbolicregression_amd as st
    from cudasymbolicregression_amd import fused

    ops, ch, n, Xs, Ys, C0 = _ragged_case(32, B=64, buckets=False)
    keep = np.array([np.isfinite(y).all() for y in Ys])
    ops, ch, n, C0 = ops[keep], ch[keep], n[keep], C0[keep]
    Xs = [x for x, k in zip(Xs, keep) if k]
    Ys = [y for y, k in zip(Ys, keep) if k]
    B = ops.shape[0]
    Xc, Yc = _t(np.concatenate(Xs)), _t(np.concatenate(Ys))
    C3, info3 = st.optimize_ragged(Xc, Yc, _t(n), _t(ops), _t(ch), _t(C0), max_iter=3, raise_errors=False)
    C3 = C3.cpu().numpy()
    C_r, info = st.optimize_ragged(Xc, Yc, _t(n), _t(ops), _t(ch), _t(C0), max_iter=50, raise_errors=False)
    mse_g = info["loss"].cpu().numpy()
    m32, m64, m64x, floor = np.empty(B), np.empty(B), np.empty(B), np.empty(B)
    C3_32, C3_64, C3_64x = np.empty_like(C3), np.empty_like(C3), np.empty_like(C3)
    for b in range(B):
        args = (Xs[b], Ys[b][None, :], ops[b:b + 1], ch[b:b + 1], C0[b:b + 1])
        with np.errstate(all="ignore"):
            f3 = helpers.oracle_fits(*args, max_iter=3)
            f50 = helpers.oracle_fits(*args, max_iter=50)
        C3_32[b], C3_64[b], C3_64x[b] = f3["C32"][0], f3["C64"][0], f3["C64x"][0]
        m32[b], m64[b], m64x[b] = f50["mse32"][0], f50["mse64"][0], f50["mse64x"][0]
        C_o = f50["C32"]
        y_o = lbfgs_oracle.oracle.fused_root(Xs[b], ops[b:b + 1], ch[b:b + 1], C_o, Ys[b][None, :],
                                             mode=1)["y"][0].astype(np.float64)
        y_g = fused.fused_root(_t(Xs[b]), _t(ops[b:b + 1]), _t(ch[b:b + 1]), _t(C_o), None, fwd_only=True,
                               raise_errors=False)["y"][0].cpu().numpy().astype(np.float64)
        floor[b] = ((y_g - y_o) ** 2).mean()
    print(helpers.compare_constants_to_oracle(C3, dict(C32=C3_32, C64=C3_64, C64x=C3_64x), "ragged, 3 iterations"),
          flush=True)
    with np.errstate(invalid="ignore"):
        extra = np.nan_to_num(2 * np.sqrt(floor * m32) + floor)
    msg = helpers.compare_fit_to_oracle(mse_g, dict(mse32=m32, mse64=m64, mse64x=m64x), info["err"].cpu().numpy(),
                                        "ragged, 50 iterations", extra_tol=extra)
    print(msg, flush=True)


def test_optimize_ragged_rejects_bad_sizes():
    import cudasymbolicregression_amd as st

    ops, ch, n, Xs, Ys, C0 = _ragged_case(33, B=4)
    X, Y = _t(np.concatenate(Xs)), _t(np.concatenate(Ys))
    with pytest.raises(RuntimeError, match="n_pts sums"):
        st.optimize_ragged(X[1:], Y[1:], _t(n), _t(ops), _t(ch), _t(C0))
    n0 = n.copy()
    n0[1] = 0
    with pytest.raises(RuntimeError, match="at least one point"):
        st.optimize_ragged(X, Y, _t(n0), _t(ops), _t(ch), _t(C0))


def test_fit_replays_from_a_captured_graph():
    """The whole fit (classification, fork to side streams, bucket kernels,
    join) captured once in a hipGraph and replayed three times, with eager
    work between replays: every replay equals the eager call bit for bit."""
    from cudasymbolicregression_amd import warmstart
    from cudasymbolicregression_amd.optimization import _fit
    import cudasymbolicregression_amd as st

    g = st.ProbabilisticContextFreeGrammar(helpers.G2, "E", 64, 3, torch.device("cpu"))
    gen = torch.Generator().manual_seed(0)
    X = (torch.rand(512, 3, generator=gen) * 6 - 3).to(_dev())
    ops, ch, _, Y = warmstart.sample_problems(g, X, 2048, seed=41)
    C0 = warmstart.random_start(ops, 42)
    eager = _fit(X, Y, ops, ch, C0, 30, 10, 1e-7, "strong_wolfe")
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        _fit(X, Y, ops, ch, C0, 30, 10, 1e-7, "strong_wolfe")  # warm the side streams and caches
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        captured = _fit(X, Y, ops, ch, C0, 30, 10, 1e-7, "strong_wolfe")
    for rep in range(3):
        for t in captured:
            t.fill_(-7)  # scribble over the outputs between replays
        junk = (Y * 2).sum()  # eager work between replays
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for a, b in zip(eager, captured):
            assert torch.equal(_bits(a), _bits(b)), rep
        del junk


def test_empty_and_single_point_batches():
    """B = 0 and a single point through the fit, the fused evaluator, the
    ragged forms and the warm start: empty outputs or the right values, no
    errors."""
    import cudasymbolicregression_amd as st
    from cudasymbolicregression_amd import fused, warmstart
    from cudasymbolicregression_amd.optimization import _fit

    dev = _dev()
    X = torch.rand(1, 3, device=dev)
    ops = torch.tensor([[1, 50, 42]], device=dev)
    ch = torch.tensor([[[-1, -1], [-1, -1], [0, 1]]], device=dev)
    C0 = torch.tensor([[0.5, 0.0, 0.0]], device=dev)
    Y = (2.0 * X[:, 0])[None, :].contiguous()
    # B = 0
    e_ops, e_ch, e_C = ops[:0], ch[:0], C0[:0]
    C, loss, iters, err, flags, ncl = _fit(X, Y[:0], e_ops, e_ch, e_C, 10, 10, 1e-7, "strong_wolfe")
    assert C.shape == (0, 3) and int(flags.abs().sum()) == 0
    r = fused.fused_root(X, e_ops, e_ch, e_C, Y[:0], mode=1)
    assert r["y"].shape == (0, 1) and r["loss"].shape == (0,)
    rr = fused.fused_root_ragged(X[:0], torch.zeros(0, dtype=torch.int64, device=dev), e_ops, e_ch, e_C, None,
                                 fwd_only=True)
    assert rr["y"].numel() == 0
    # one point: C * X0 fits 2 * X0 exactly
    Cf, Yf = st.optimize(X, Y, ops, ch, C0, max_iter=50)
    assert abs(float(Cf[0, 0]) - 2.0) < 1e-4
    Cr, info = st.optimize_ragged(X, Y.reshape(-1), torch.tensor([1], device=dev), ops, ch, C0, max_iter=50)
    assert torch.equal(Cr, Cf)
    # warm start of one problem (the network reads 256 target points)
    model = warmstart.WarmStartMLP(dev, seed=0)
    Y256 = torch.randn(1, 256, device=dev)
    Cw = model.predict(ops, Y256)
    assert Cw.shape == (1, 3) and Cw[0, 1] == 0 and Cw[0, 2] == 0
