"""GPU parity of the evaluation kernels against the C oracle.

* Every operator alone (forward and backward): FP32 within
  |a-b| <= 1e-5*max(|a|,|b|) + 1e-6 of the reference's FP32 CPU algorithm,
  identical NaN / inf masks and error codes at every point; FP64 within 1e-12.
* Composed random programs: the same strict test at every point where the
  reference's own result is stable (helpers.unstable_points: scaling X or C
  by 1 +- 2^-21 leaves its class and value within the tolerance).  At the
  rare unstable points (near-cancellations, exp of large arguments) last-ulp
  differences between math libraries legitimately change the result, so they
  are counted (at most 5%) and skipped.  For root-only results a point is
  also unstable when any inner node is (helpers.unstable_tape_points, 1e-3),
  and the tolerance adds the first-order rounding bound
  8 * 2^-24 * sum_k |dy/dv_k| |v_k| (helpers.rounding_bound, FP64 oracle).
* Sums over points or expressions (grad_C, grad_X, loss) reduce in FP64 on
  the GPU and in FP32 in the reference: they must be at least as accurate as
  the reference against the FP64 oracle (helpers.check_vs_f64).
"""
import numpy as np
import pytest
import torch

import helpers
import oracle

pytestmark = pytest.mark.gpu

RTOL32, ATOL32 = 1e-5, 1e-6
ALL_OPS = helpers.UNARY + helpers.BINARY


def _dev():
    return torch.device("cuda:0")


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(_dev())


def _case(seed, B=16, M=24, n_x=3, N=100, dtype=np.float32, **kw):
    rng = np.random.default_rng(seed)
    ops_bm, ch_bm = helpers.random_batch(seed, B, M, n_x, **kw)
    C_bm = helpers.constants_like(ops_bm, rng, dtype)
    X = rng.uniform(-3, 3, (N, n_x)).astype(dtype)
    return X, ops_bm, ch_bm, C_bm


def test_device_limits_reported():
    p = torch.cuda.get_device_properties(0)
    print(f"device {p.name}: {p.multi_processor_count} CUs, shared/block {p.shared_memory_per_block}, "
          f"shared/CU {getattr(p, 'shared_memory_per_multiprocessor', 'n/a')}", flush=True)


def _single_op_program(op, with_const):
    """[X0 (, X1)] op  or  [C (, X0)] op  as a node-first (M, 1) program."""
    ar = 1 if op < 40 else 2
    leaves = ([1] if with_const else [50]) + ([51 if not with_const else 50] if ar == 2 else [])
    ops = np.array(leaves + [op], np.int64)[:, None]
    ch = np.full((len(ops), 1, 2), -1, np.int64)
    ch[-1, 0, :ar] = np.arange(ar)
    return ops, ch


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("op", ALL_OPS)
def test_single_op_forward_backward(op, dtype):
    from cudasymbolicregression_amd.evaluation import evaluate_backend

    rng = np.random.default_rng(op)
    N = 257
    X = (rng.standard_normal((N, 2)) * 1.5).astype(dtype)
    X[:8, 0] = [0.0, -0.0, 1.0, -1.0, 0.5, -2.0, 1e-30, 3.0]
    for with_const in (False, True):
        ops, ch = _single_op_program(op, with_const)
        C = np.zeros(ops.shape, dtype)
        C[0, 0] = dtype(0.75) if with_const else 0
        ref = oracle.forward(X, ops, ch, C)
        Xt = _t(X).requires_grad_(True)
        Ct = _t(C).requires_grad_(True)
        out = evaluate_backend(Xt, _t(ops), _t(ch), Ct)
        tol = (1e-12, 1e-13) if dtype == np.float64 else (RTOL32, ATOL32)
        ok, msg = helpers.allclose_nan(out.detach().cpu().numpy(), ref, *tol)
        assert ok, f"forward op {op}: {msg}"
        g = np.zeros_like(ref)
        g[-1] = rng.standard_normal((N, 1)).astype(dtype)
        gc, gx, e = oracle.backward(g, ref, ops, ch, 2)
        if e:
            with pytest.raises(RuntimeError, match="backward pass"):
                out.backward(_t(g))
            continue
        out.backward(_t(g))
        ok, msg = helpers.allclose_nan(Xt.grad.cpu().numpy(), gx, *tol)
        assert ok, f"grad_X op {op}: {msg}"
        # grad_C sums N terms: compare against FP64
        ref64 = oracle.backward(g.astype(np.float64), oracle.forward(X.astype(np.float64), ops, ch,
                                                                      C.astype(np.float64)), ops, ch, 2)[0]
        ok, msg = helpers.check_vs_f64(Ct.grad.cpu().numpy(), gc, ref64)
        assert ok, f"grad_C op {op}: {msg}"


@pytest.mark.parametrize("seed", range(4))
def test_backend_forward_composed(seed):
    from cudasymbolicregression_amd.evaluation import evaluate_backend

    X, ops_bm, ch_bm, C_bm = _case(seed, N=130)
    ops, ch, C = helpers.to_mb(ops_bm, ch_bm, C_bm)
    got = evaluate_backend(_t(X), _t(ops), _t(ch), _t(C)).cpu().numpy()
    r32 = oracle.forward(X, ops, ch, C)
    uns = helpers.unstable_points(lambda x, c: oracle.forward(x, ops, ch, c), X, C, r32, 2.0 ** -21, RTOL32, ATOL32)
    ok, msg = helpers.check_stable(got, r32, uns)
    print(f"forward seed {seed}: {msg}")
    assert ok, msg
    X64, C64 = X.astype(np.float64), C.astype(np.float64)
    r64 = oracle.forward(X64, ops, ch, C64)
    got64 = evaluate_backend(_t(X64), _t(ops), _t(ch), _t(C64)).cpu().numpy()
    uns64 = helpers.unstable_points(lambda x, c: oracle.forward(x, ops, ch, c), X64, C64, r64, 2.0 ** -45, 1e-12, 1e-13)
    ok, msg = helpers.check_stable(got64, r64, uns64, 1e-12, 1e-13)
    assert ok, "f64 " + msg


@pytest.mark.parametrize("seed", range(4))
def test_backend_backward_composed(seed):
    from cudasymbolicregression_amd.evaluation import evaluate_backend

    X, ops_bm, ch_bm, C_bm = _case(seed, N=70, M=32, unary=helpers.G2_UNARY + [14, 27, 29])
    ops, ch, C = helpers.to_mb(ops_bm, ch_bm, C_bm)
    cache = oracle.forward(X, ops, ch, C)
    cache64 = oracle.forward(X.astype(np.float64), ops, ch, C.astype(np.float64))
    rng = np.random.default_rng(seed + 50)
    g = rng.standard_normal(cache.shape).astype(np.float32)
    g[np.isnan(cache) | np.isnan(cache64)] = 0
    gc32, gx32, e32 = oracle.backward(g, cache, ops, ch, X.shape[1])
    gc64, gx64, e64 = oracle.backward(g.astype(np.float64), cache64, ops, ch, X.shape[1])
    Xt = _t(X).requires_grad_(True)
    Ct = _t(C).requires_grad_(True)
    out = evaluate_backend(Xt, _t(ops), _t(ch), Ct)
    if e32:
        with pytest.raises(RuntimeError, match="backward pass"):
            out.backward(_t(g))
        return
    out.backward(_t(g))

    def grads(x, c):
        cc = oracle.forward(x, ops, ch, c)
        gc, gx, _ = oracle.backward(g, cc, ops, ch, X.shape[1])
        return np.concatenate([gc.ravel(), gx.ravel()])

    both32 = np.concatenate([gc32.ravel(), gx32.ravel()])
    uns = helpers.unstable_points(grads, X, C, both32, 2.0 ** -21, 1e-5, 1e-5)
    gpu = np.concatenate([Ct.grad.cpu().numpy().ravel(), Xt.grad.cpu().numpy().ravel()])
    ref64 = np.concatenate([gc64.ravel(), gx64.ravel()])
    ok, msg = helpers.check_vs_f64(gpu[~uns], both32[~uns], ref64[~uns], rtol=1e-5, atol=1e-5)
    print(f"backward seed {seed}: {uns.mean():.4%} unstable sums; {msg}")
    assert uns.mean() <= 0.05 and ok, msg


def test_validation_errors_gpu():
    from cudasymbolicregression_amd.evaluation import evaluate_backend

    X = torch.randn(10, 2, device=_dev())
    ops = torch.tensor([[50], [40]], dtype=torch.int64, device=_dev())
    ch = -torch.ones(2, 1, 2, dtype=torch.int64, device=_dev())
    ch[1, 0, 0] = 0
    ch[1, 0, 1] = 2
    C = torch.zeros(2, 1, device=_dev())
    with pytest.raises(RuntimeError, match="Error during input validation: Validation failed: A child's index is "
                                           "greater than or equal to its parent's."):
        evaluate_backend(X, ops, ch, C)
    ch[1, 0, 1] = -1
    with pytest.raises(RuntimeError, match="Validation failed: Operator arity does not match the number of children."):
        evaluate_backend(X, ops, ch, C)
    for op in helpers.UNARY + helpers.BINARY:
        o = torch.tensor([[op]], dtype=torch.int64, device=_dev())
        c = -torch.ones(1, 1, 2, dtype=torch.int64, device=_dev())
        with pytest.raises(RuntimeError, match="Operator arity does not match the number of children."):
            evaluate_backend(X, o, c, torch.zeros(1, 1, device=_dev()))


def test_backward_error_codes_gpu():
    from cudasymbolicregression_amd.evaluation import evaluate_backend

    X = torch.zeros(1, 2, device=_dev(), requires_grad=True)
    ops = torch.tensor([[2], [50], [43]], dtype=torch.int64, device=_dev())
    ch = -torch.ones(3, 1, 2, dtype=torch.int64, device=_dev())
    ch[2, 0, 0] = 0
    ch[2, 0, 1] = 1
    C = torch.zeros(3, 1, device=_dev(), requires_grad=True)
    y = evaluate_backend(X, ops, ch, C)
    with pytest.raises(RuntimeError, match="a/b where b = 0"):
        y[-1].backward(torch.ones(1, 1, device=_dev()))
    ops = torch.tensor([[0]], dtype=torch.int64, device=_dev())
    ch = -torch.ones(1, 1, 2, dtype=torch.int64, device=_dev())
    C = torch.zeros(1, 1, device=_dev(), requires_grad=True)
    y = evaluate_backend(X, ops, ch, C)
    assert torch.isnan(y).all()
    with pytest.raises(RuntimeError, match="passed to a NO_OP node"):
        y[-1].backward(torch.ones(1, 1, device=_dev()))
    # zero upstream on a NO_OP is fine and gives zero gradients
    X2 = torch.randn(10, 2, device=_dev(), requires_grad=True)
    C2 = torch.zeros(1, 1, device=_dev(), requires_grad=True)
    y = evaluate_backend(X2, ops, ch, C2)
    y[-1].backward(torch.zeros(10, 1, device=_dev()))
    assert (X2.grad == 0).all() and (C2.grad == 0).all()


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("M", [12, 64, 128])
def test_fused_root_matches_oracle(M, mode):
    from cudasymbolicregression_amd import fused

    print(f"case fused M={M} mode={mode}", flush=True)
    seed = M + mode
    X, ops_bm, ch_bm, C_bm = _case(seed, B=300, M=M, N=200, unary=helpers.G2_UNARY, binary=helpers.G2_BINARY,
                                   depth=7)
    rng = np.random.default_rng(seed)
    dy = rng.standard_normal((300, 200)).astype(np.float32)
    r32 = oracle.fused_root(X, ops_bm, ch_bm, C_bm, dy, mode=mode)
    r64 = oracle.fused_root(X.astype(np.float64), ops_bm, ch_bm, C_bm.astype(np.float64), dy.astype(np.float64),
                            mode=mode)
    out = fused.fused_root(_t(X), _t(ops_bm), _t(ch_bm), _t(C_bm), _t(dy), mode=mode, want_grad_X=True,
                           raise_errors=False)
    pert = [oracle.fused_root(x, ops_bm, ch_bm, c, dy, mode=mode)
            for x, c in ((X * (1 + 2.0 ** -21), C_bm), (X * (1 - 2.0 ** -21), C_bm),
                         (X, C_bm * (1 + 2.0 ** -21)), (X, C_bm * (1 - 2.0 ** -21)))]
    pert = [{k: v for k, v in p.items()} for p in pert]

    def unstable(key, rtol, atol):
        it = iter(pert)
        return helpers.unstable_points(lambda x, c: next(it)[key], X, C_bm, r32[key], 0.0, rtol, atol)

    uy = unstable("y", RTOL32, ATOL32) | helpers.unstable_tape_points(X, ops_bm, ch_bm, C_bm, 2.0 ** -21, 1e-3, 1e-6)
    rb = helpers.rounding_bound(X, ops_bm, ch_bm, C_bm)
    ok, msg = helpers.check_stable(out["y"].cpu().numpy(), r32["y"], uy, extra_tol=rb)
    print(f"fused M={M} mode={mode} y: {msg}")
    assert ok, "y " + msg
    err = out["err"].cpu().numpy()
    # an expression's error code can only differ where one of its points is unstable
    err_ok = (err == r32["err"]) | uy.any(axis=1)
    assert err_ok.all(), np.nonzero(~err_ok)
    good = (r32["err"] == 0) & (r64["err"] == 0) & (err == 0)
    ug = unstable("grad_C", 1e-5, 1e-5)
    sel = good[:, None] & ~ug
    ok, msg = helpers.check_vs_f64(out["grad_C"].cpu().numpy()[sel], r32["grad_C"][sel], r64["grad_C"][sel], atol=1e-5)
    assert ug[good].mean() <= 0.05 and ok, "grad_C " + msg
    if mode == 1:
        ul = unstable("loss", 1e-5, 1e-9)
        sel = ~ul
        ok, msg = helpers.check_vs_f64(out["loss"].cpu().numpy()[sel], r32["loss"][sel], r64["loss"][sel])
        assert ok, "loss " + msg
    if good.all():
        ok, msg = helpers.check_vs_f64(out["grad_X"].cpu().numpy(), r32["grad_X"], r64["grad_X"], atol=1e-4)
        assert ok, "grad_X " + msg
    # FP64 path against the FP64 oracle
    out64 = fused.fused_root(_t(X.astype(np.float64)), _t(ops_bm), _t(ch_bm), _t(C_bm.astype(np.float64)),
                             _t(dy.astype(np.float64)), mode=mode, raise_errors=False)
    X64, C64 = X.astype(np.float64), C_bm.astype(np.float64)
    uy64 = helpers.unstable_tape_points(X64, ops_bm, ch_bm, C64, 2.0 ** -50, 1e-6, 1e-12)
    rb64 = helpers.rounding_bound(X, ops_bm, ch_bm, C_bm, eps=2.0 ** -53)
    ok, msg = helpers.check_stable(out64["y"].cpu().numpy(), r64["y"], uy64, 1e-10, 1e-12, extra_tol=rb64)
    assert ok, "y64 " + msg


def test_public_evaluate_shape_and_view():
    from cudasymbolicregression_amd.evaluation import evaluate

    X, ops_bm, ch_bm, C_bm = _case(3, B=7, M=10, N=33, unary=[10, 11, 14, 27], binary=[40, 41, 42])
    y = evaluate(_t(X), _t(ops_bm), _t(ch_bm), _t(C_bm))
    assert y.shape == (7, 10, 33)
    ops, ch, C = helpers.to_mb(ops_bm, ch_bm, C_bm)
    ref = oracle.forward(X, ops, ch, C).transpose(2, 0, 1)
    ok, msg = helpers.allclose_nan(y.cpu().numpy(), ref, RTOL32, ATOL32)
    assert ok, msg


@pytest.mark.parametrize("K", [1, 3])
def test_multiple_constants_match_oracle(K):
    from cudasymbolicregression_amd.evaluation import evaluate_multiple

    X, ops_bm, ch_bm, C_bm = _case(21, B=9, M=16, N=90, unary=[10, 11, 12, 14], binary=[40, 41, 42])
    rng = np.random.default_rng(4)
    Ck = np.repeat(C_bm[:, :, None], K, axis=2) + ((ops_bm == 1)[:, :, None] *
                                                   rng.standard_normal((*C_bm.shape, K))).astype(np.float32)
    Ck = Ck.astype(np.float32)
    Xt = _t(X).requires_grad_(True)
    Ct = _t(Ck).requires_grad_(True)
    y = evaluate_multiple(Xt, _t(ops_bm), _t(ch_bm), Ct)
    assert y.shape == (9, 16, 90, K)
    ops, ch = helpers.to_mb(ops_bm, ch_bm)
    Cmbk = np.ascontiguousarray(Ck.transpose(1, 0, 2))
    r32 = oracle.multi_forward(X, ops, ch, Cmbk)
    r64 = oracle.multi_forward(X.astype(np.float64), ops, ch, Cmbk.astype(np.float64))
    ok, msg = helpers.check_vs_f64(y.detach().cpu().numpy(), r32.transpose(2, 0, 1, 3), r64.transpose(2, 0, 1, 3))
    assert ok, msg
    g = rng.standard_normal(r32.shape).astype(np.float32)
    g[np.isnan(r32) | np.isnan(r64)] = 0
    gc32, gx32, e32 = oracle.multi_backward(g, r32, ops, ch, 3)
    gc64, gx64, _ = oracle.multi_backward(g.astype(np.float64), r64, ops, ch, 3)
    y.backward(_t(np.ascontiguousarray(g.transpose(2, 0, 1, 3))))
    assert e32 == 0
    ok, msg = helpers.check_vs_f64(Ct.grad.cpu().numpy(), gc32.transpose(1, 0, 2), gc64.transpose(1, 0, 2), atol=1e-5)
    assert ok, "grad_C " + msg
    ok, msg = helpers.check_vs_f64(Xt.grad.cpu().numpy(), gx32, gx64, atol=1e-5)
    assert ok, "grad_X " + msg


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_fused_root_ragged_equals_per_problem_calls(mode, dtype):
    """Each expression on its own points (N_b in [1, 300]) in one ragged call
    gives, bit for bit, what a dense call on that expression alone gives with
    the same bucket kernels: y, loss (mode 1), grad_C and error codes.  (The
    ragged form runs the one-wave-per-program kernels, variant 2 of buckets 0
    and 1; a default dense FP32 call runs the planned wide kernels, whose
    FP64 sums group the points differently.)"""
    from cudasymbolicregression_amd import fused

    rng = np.random.default_rng(50 + mode)
    B, M, n_x = 40, 96, 3
    ops, ch = helpers.random_batch(51 + mode, B, M, n_x, unary=helpers.G2_UNARY, binary=helpers.G2_BINARY,
                                   depth=7)
    C = helpers.constants_like(ops, rng, dtype)
    n = rng.integers(1, 301, size=B)
    n[0] = 1
    Xs = [rng.uniform(-3, 3, (int(k), n_x)).astype(dtype) for k in n]
    ds = [rng.standard_normal(int(k)).astype(dtype) for k in n]
    td = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(_dev())  # noqa: E731
    r = fused.fused_root_ragged(td(np.concatenate(Xs)), td(n), td(ops), td(ch), td(C), td(np.concatenate(ds)),
                                mode=mode, raise_errors=False)
    bits = lambda t: t.view(torch.int32) if t.dtype == torch.float32 else t.view(torch.int64)  # noqa: E731
    start = np.concatenate([[0], np.cumsum(n)[:-1]])
    lengths = (ops != 0).sum(1)
    print("program lengths:", sorted(lengths.tolist()), flush=True)
    same_kernels = [2, 2, 0, 0, 0] if dtype == np.float32 else None
    for b in range(B):
        d = fused.fused_root(td(Xs[b]), td(ops[b:b + 1]), td(ch[b:b + 1]), td(C[b:b + 1]), td(ds[b][None, :]),
                             mode=mode, raise_errors=False, config=same_kernels)
        seg = r["y"][start[b]:start[b] + n[b]]
        assert torch.equal(bits(seg), bits(d["y"][0])), b
        assert torch.equal(bits(r["grad_C"][b]), bits(d["grad_C"][0])), b
        assert int(r["err"][b]) == int(d["err"][0]), b
        if mode == 1:
            assert torch.equal(r["loss"][b:b + 1].view(torch.int64), d["loss"].view(torch.int64)), b


def test_fused_root_ragged_rejects_bad_sizes():
    from cudasymbolicregression_amd import fused

    X = torch.zeros(10, 3, device=_dev())
    ops = torch.tensor([[50, 1, 40]], device=_dev())
    ch = torch.tensor([[[-1, -1], [-1, -1], [0, 1]]], device=_dev())
    C = torch.zeros(1, 3, device=_dev())
    with pytest.raises(RuntimeError, match="sum to"):
        fused.fused_root_ragged(X, torch.tensor([9], device=_dev()), ops, ch, C, None, fwd_only=True)
    with pytest.raises(RuntimeError, match="sum to"):
        fused.fused_root_ragged(X[:0], torch.tensor([0], device=_dev()), ops, ch, C, None, fwd_only=True)


@pytest.mark.parametrize("M", [32, 96])
def test_backward_leaves_the_upstream_gradient_unchanged(M):
    """The caller's dL/dcache is not modified by backward, whether the kernel
    of this size reads it directly (M <= 64) or works on a copy (M = 96)."""
    import cudasymbolicregression_amd as st

    X, ops_bm, ch_bm, C_bm = _case(60 + M, B=24, M=M, N=90, unary=[10, 11, 14], binary=[40, 41, 42])
    Xt = _t(X).requires_grad_(True)
    Ct = _t(C_bm).requires_grad_(True)
    out = st.evaluation.evaluate(Xt, _t(ops_bm), _t(ch_bm), Ct)  # (B, M, N)
    torch.cuda.synchronize()
    ops = ops_bm
    g = torch.zeros_like(out)
    root = np.maximum((ops != 0).sum(1) - 1, 0)
    for b in range(ops.shape[0]):
        g[b, root[b]] = torch.randn(out.shape[2], device=_dev())
    g0 = g.clone()
    out.backward(g)
    torch.cuda.synchronize()
    assert torch.equal(g, g0)
    assert Ct.grad is not None and torch.isfinite(Ct.grad).any()
