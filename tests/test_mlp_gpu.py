"""GPU tests of the warm-start MLP kernels (csrc/mlp_kernels.hip).

The BF16 GEMM is checked against an FP64 product of the same BF16 operands
(bound: K * 2^-24 * sum |a b| + 1e-6, the worst case of FP32 accumulation);
transposes and BF16 conversions are bit-exact; the network's loss and every
gradient are checked against the FP64 oracle (oracle/mlp_oracle.py, BF16
rounding at the same points) to 2e-3 relative Frobenius norm; Adam matches
the oracle's FP32 restatement to 1e-6 relative.  The last tests train the
network on labels from the plain device L-BFGS and compare warm starts with
randn starts on held-out problems.
"""
import math

import numpy as np
import pytest
import torch

import helpers
import mlp_oracle as mo

pytestmark = pytest.mark.gpu



def _dev():
    return torch.device("cuda:0")


def _lib():
    from cudasymbolicregression_amd import _lib as L

    return L


def _s():
    L = _lib()
    return L.stream_of(_dev())


def _bf16_dev(a64):
    return torch.from_numpy(np.ascontiguousarray(a64, dtype=np.float32)).to(torch.bfloat16).to(_dev())


def _np(t):
    return t.detach().float().cpu().numpy().astype(np.float64)


@pytest.mark.parametrize("M,N,K", [(32, 32, 16), (96, 544, 48), (64, 64, 3648), (128, 512, 512)])
@pytest.mark.parametrize("epi", [0, 1, 2])
def test_gemm_nt_matches_fp64(M, N, K, epi):
    L = _lib()
    rng = np.random.default_rng(M * 7 + N * 3 + K + epi)
    A = mo.bf16_round(rng.standard_normal((M, K)))
    Bt = mo.bf16_round(rng.standard_normal((N, K)))
    bias = rng.standard_normal(N).astype(np.float32).astype(np.float64)
    Ad, Bd = _bf16_dev(A), _bf16_dev(Bt)
    bd = torch.from_numpy(bias.astype(np.float32)).to(_dev())
    out = torch.empty(M, N, dtype=torch.float32, device=_dev())
    outb = torch.empty(M, N, dtype=torch.bfloat16, device=_dev())
    pre = torch.empty(M, N, dtype=torch.float32, device=_dev())
    print(f"gemm M={M} N={N} K={K} epi={epi}", flush=True)
    rc = L.hip().sr_mlp_gemm_nt(L.ptr(Ad), L.ptr(Bd), L.ptr(bd), M, N, K, K, K, epi, L.ptr(out), L.ptr(outb),
                                L.ptr(pre) if epi == 2 else None, _s())
    assert rc == 0
    torch.cuda.synchronize()
    z = A @ Bt.T + (bias if epi else 0.0)
    bound = K * 2.0 ** -24 * (np.abs(A) @ np.abs(Bt).T + (np.abs(bias) if epi else 0.0)) + 1e-6
    if epi == 2:
        assert np.all(np.abs(_np(pre) - z) <= bound)
        h = mo.gelu(z)
        assert np.all(np.abs(_np(out) - h) <= bound + 1e-6 * np.abs(h))
    else:
        assert np.all(np.abs(_np(out) - z) <= bound)
    # the BF16 copy is the round-to-nearest-even image of the FP32 result
    assert torch.equal(outb, out.to(torch.bfloat16))


def test_gemm_rejects_unaligned_shapes():
    L = _lib()
    a = torch.zeros(64 * 64 + 8, dtype=torch.bfloat16, device=_dev())
    out = torch.empty(64 * 64, dtype=torch.float32, device=_dev())
    ok = (L.ptr(a), L.ptr(a), None, 32, 32, 16, 16, 16, 0, L.ptr(out), None, None, _s())
    assert L.hip().sr_mlp_gemm_nt(*ok) == 0
    for M, N, K, lda in [(33, 32, 16, 16), (32, 40, 16, 16), (32, 32, 24, 24), (0, 32, 16, 16), (32, 32, 16, 12),
                         (32, 32, 16, 8)]:
        assert L.hip().sr_mlp_gemm_nt(L.ptr(a), L.ptr(a), None, M, N, K, lda, K, 0, L.ptr(out), None, None,
                                      _s()) == -1
    # misaligned operand, no output, missing bias for a bias epilogue
    assert L.hip().sr_mlp_gemm_nt(L.ptr(a) + 2, L.ptr(a), None, 32, 32, 16, 16, 16, 0, L.ptr(out), None, None,
                                  _s()) == -1
    assert L.hip().sr_mlp_gemm_nt(L.ptr(a), L.ptr(a), None, 32, 32, 16, 16, 16, 0, None, None, None, _s()) == -1
    assert L.hip().sr_mlp_gemm_nt(L.ptr(a), L.ptr(a), None, 32, 32, 16, 16, 16, 1, L.ptr(out), None, None,
                                  _s()) == -1
    torch.cuda.synchronize()


def test_gemm_strided_operands():
    """A and Bt read as column blocks of wider matrices (row strides lda, ldb)."""
    L = _lib()
    rng = np.random.default_rng(17)
    M, N, K, lda, ldb, off = 64, 96, 48, 200, 136, 24
    A = mo.bf16_round(rng.standard_normal((M, lda)))
    Bt = mo.bf16_round(rng.standard_normal((N, ldb)))
    Ad, Bd = _bf16_dev(A), _bf16_dev(Bt)
    out = torch.empty(M, N, dtype=torch.float32, device=_dev())
    assert L.hip().sr_mlp_gemm_nt(Ad.data_ptr() + 2 * off, Bd.data_ptr() + 2 * off, None, M, N, K, lda, ldb, 0,
                                  L.ptr(out), None, None, _s()) == 0
    torch.cuda.synchronize()
    a, b = A[:, off:off + K], Bt[:, off:off + K]
    bound = K * 2.0 ** -24 * (np.abs(a) @ np.abs(b).T) + 1e-6
    assert np.all(np.abs(_np(out) - a @ b.T) <= bound)


@pytest.mark.parametrize("rows,cols", [(70, 33), (64, 3648), (4096, 64)])
def test_transpose_bit_exact(rows, cols):
    L = _lib()
    a = torch.randn(rows, cols, device=_dev()).to(torch.bfloat16)
    out = torch.empty(cols, rows, dtype=torch.bfloat16, device=_dev())
    assert L.hip().sr_mlp_transpose(L.ptr(a), L.ptr(out), rows, cols, _s()) == 0
    assert torch.equal(out, a.t().contiguous())


def test_col_sum_gelu_bwd_to_bf16():
    L = _lib()
    lib = L.hip()
    rng = np.random.default_rng(5)
    M, N = 4160, 96
    dz = torch.from_numpy(rng.standard_normal((M, N)).astype(np.float32)).to(_dev())
    db = torch.empty(N, dtype=torch.float32, device=_dev())
    ws = torch.empty(lib.sr_mlp_col_sum_workspace_bytes(M, N), dtype=torch.uint8, device=_dev())
    assert lib.sr_mlp_col_sum(L.ptr(dz), M, N, L.ptr(db), L.ptr(ws), ws.numel(), _s()) == 0
    want = _np(dz).sum(axis=0).astype(np.float32)
    np.testing.assert_allclose(db.cpu().numpy(), want, rtol=2e-7, atol=1e-6)
    assert lib.sr_mlp_col_sum(L.ptr(dz), M, N, L.ptr(db), L.ptr(ws), ws.numel() - 8, _s()) == -2

    z = torch.from_numpy((rng.standard_normal(10000) * 3).astype(np.float32)).to(_dev())
    dh = torch.from_numpy(rng.standard_normal(10000).astype(np.float32)).to(_dev())
    dzf = torch.empty_like(z)
    dzb = torch.empty(10000, dtype=torch.bfloat16, device=_dev())
    assert lib.sr_mlp_gelu_bwd(L.ptr(dh), L.ptr(z), 10000, L.ptr(dzf), L.ptr(dzb), _s()) == 0
    want = _np(dh) * mo.gelu_grad(_np(z))
    np.testing.assert_allclose(_np(dzf), want, rtol=1e-5, atol=1e-6)
    assert torch.equal(dzb, dzf.to(torch.bfloat16))

    x = torch.from_numpy(np.concatenate([rng.standard_normal(5000).astype(np.float32) * 100,
                                         np.array([np.inf, -np.inf, 0.0, -0.0], dtype=np.float32)])).to(_dev())
    xb = torch.empty(x.numel(), dtype=torch.bfloat16, device=_dev())
    assert lib.sr_mlp_to_bf16(L.ptr(x), x.numel(), L.ptr(xb), _s()) == 0
    assert torch.equal(xb, x.to(torch.bfloat16))


def test_adam_kernel_matches_oracle():
    L = _lib()
    rng = np.random.default_rng(6)
    n = 1000
    w0 = rng.standard_normal(n).astype(np.float32)
    w = torch.from_numpy(w0.copy()).to(_dev())
    m = torch.zeros(n, device=_dev())
    v = torch.zeros(n, device=_dev())
    wb = torch.empty(n, dtype=torch.bfloat16, device=_dev())
    wr, mr, vr = w0.copy(), np.zeros(n, np.float32), np.zeros(n, np.float32)
    for t in range(1, 4):
        g = rng.standard_normal(n).astype(np.float32)
        gd = torch.from_numpy(g).to(_dev())
        assert L.hip().sr_mlp_adam(L.ptr(w), L.ptr(gd), L.ptr(m), L.ptr(v), n, 2e-3, 0.9, 0.999, 1e-8, t, 1e-2,
                                   L.ptr(wb), _s()) == 0
        wr, mr, vr = mo.adam(wr, g, mr, vr, t, lr=2e-3, wd=1e-2)
    np.testing.assert_allclose(w.cpu().numpy(), wr, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(m.cpu().numpy(), mr, rtol=1e-6, atol=1e-8)
    np.testing.assert_allclose(v.cpu().numpy(), vr, rtol=1e-6, atol=1e-10)
    assert torch.equal(wb, w.to(torch.bfloat16))


def _problems(P, seed, n_pts=512):
    import cudasymbolicregression_amd as st
    from cudasymbolicregression_amd import warmstart

    g = st.ProbabilisticContextFreeGrammar(helpers.G2, "E", 64, 3, torch.device("cpu"))
    gen = torch.Generator().manual_seed(0)
    X = (torch.rand(n_pts, 3, generator=gen) * 6 - 3).to(_dev())
    ops, ch, Ct, Y = warmstart.sample_problems(g, X, P, seed)
    return g, X, ops, ch, Ct, Y


@pytest.mark.parametrize("B,std", [(64, False), (50, False), (50, True)])
def test_loss_and_grads_match_fp64_oracle(B, std):
    from cudasymbolicregression_amd import warmstart

    _, _, ops, _, Ct, Y = _problems(B, seed=11)
    model = warmstart.WarmStartMLP(_dev(), seed=3, std_features=std)
    # larger biases than the default init so every GELU region is visited
    for b in model.b:
        b.normal_(0.0, 0.5)
    print(f"mlp loss_and_grads B={B} std_features={std}", flush=True)
    rows = model.loss_and_grads(ops, Y, Ct)
    torch.cuda.synchronize()
    params = [(_np(W), _np(b)) for W, b in zip(model.W, model.b)]
    ops_n = ops.cpu().numpy()
    x = mo.features(ops_n, Y.cpu().numpy(), model.slots, model.vocab, model.n_pts, model.row_len,
                    std_block=bool(model.std_block))
    mask = ops_n == 1
    loss, grads = mo.loss_and_grads(params, x, Ct.cpu().numpy().astype(np.float64), mask)
    got_loss = float(rows.double().sum()) / B
    assert abs(got_loss - loss) <= 1e-4 * loss
    out_ref, _ = mo.forward(params, x)
    out = _np(model.forward(ops, Y))
    assert np.linalg.norm(out - out_ref) <= 2e-3 * np.linalg.norm(out_ref)
    for l in range(model.layers):
        for got, want, what in ((_np(model.gW[l]), grads[l][0], "W"), (_np(model.gb[l]), grads[l][1], "b")):
            err = np.linalg.norm(got - want) / max(np.linalg.norm(want), 1e-30)
            print(f"layer {l} d{what}: rel err {err:.2e}", flush=True)
            assert err <= 2e-3, (l, what, err)
            assert np.max(np.abs(got - want)) <= 2e-2 * np.max(np.abs(want))


@pytest.mark.parametrize("std", [False, True])
def test_features_predict_and_select_start(std):
    from cudasymbolicregression_amd import _lib as L
    from cudasymbolicregression_amd import warmstart

    _, _, ops, _, _, Y = _problems(40, seed=12)
    model = warmstart.WarmStartMLP(_dev(), seed=1, std_features=std)
    B, M = ops.shape
    Bp = 64
    xb = torch.empty(Bp, model.row_len, dtype=torch.bfloat16, device=_dev())
    assert L.hip().sr_mlp_features(L.ptr(ops), M, B, Bp, model.slots, model.vocab, L.ptr(Y), Y.shape[1],
                                   model.n_pts, model.std_block, model.row_len, L.ptr(xb), _s()) == 0
    want = mo.features(ops.cpu().numpy(), Y.cpu().numpy(), model.slots, model.vocab, model.n_pts, model.row_len,
                    std_block=bool(model.std_block))
    got = _np(xb)
    onehot = model.slots * model.vocab
    assert np.array_equal(got[:B, :onehot], want[:, :onehot])
    # FP32 device arithmetic vs FP64 rounded twice: at most one BF16 step apart
    # (plus a small absolute floor for standardised values near 0)
    assert np.all(np.abs(got[:B, onehot:] - want[:, onehot:]) <= 2.0 ** -7 * np.abs(want[:, onehot:]) + 1e-4)
    assert not got[B:].any()

    out = model.forward(ops, Y)
    C = model.predict(ops, Y)
    assert torch.equal(C, torch.where(ops == 1, out[:, :M], torch.zeros_like(C)))

    Cw = torch.full((4, 3), 1.0, device=_dev())
    Cr = torch.full((4, 3), 2.0, device=_dev())
    lw = torch.tensor([float("nan"), 1.0, 2.0, 1.0], dtype=torch.float64, device=_dev())
    lr = torch.tensor([1.0, 1.0, 1.0, float("nan")], dtype=torch.float64, device=_dev())
    Cs = torch.empty(4, 3, device=_dev())
    used = torch.empty(4, dtype=torch.int32, device=_dev())
    assert L.hip().sr_mlp_select_start(L.ptr(Cw), L.ptr(Cr), L.ptr(lw), L.ptr(lr), 3, 4, L.ptr(Cs), L.ptr(used),
                                       _s()) == 0
    assert used.tolist() == [0, 1, 0, 1]
    assert Cs[:, 0].tolist() == [2.0, 1.0, 2.0, 1.0]


def test_training_reduces_loss_below_zero_predictor():
    """Labels = true constants: a few epochs must beat predicting 0."""
    from cudasymbolicregression_amd import warmstart

    _, _, ops, _, Ct, Y = _problems(32768, seed=13)
    model = warmstart.WarmStartMLP(_dev(), seed=0)
    hist = warmstart.train(model, ops, Y, Ct, epochs=4, batch=2048, lr=1e-3, lr_final=1e-4)
    zero_loss = float((Ct.double() ** 2).sum()) / Ct.shape[0]
    print(f"training losses {hist}, zero predictor {zero_loss:.4f}", flush=True)
    assert all(math.isfinite(h) for h in hist)
    assert hist[-1] < hist[0]
    assert hist[-1] < 0.9 * zero_loss


def test_warm_start_beats_randn_on_held_out_problems():
    """Train on (Ops, Y, fitted C) from the plain device L-BFGS, then start the
    L-BFGS from the prediction (randn where that is worse) on held-out
    problems: the start loss is never worse than randn's, it is lower for most
    problems, and at least as many problems reach rel-MSE <= 1e-6."""
    from cudasymbolicregression_amd import warmstart
    from cudasymbolicregression_amd.optimization import _fit

    g, X, _, _, _, _ = _problems(1, seed=14)
    ops, ch, Y, label, stats = warmstart.make_training_set(g, X, 65536, seed=14)
    print(f"training set {stats}", flush=True)
    assert stats["labelled"] > 0.3 * stats["sampled"]
    model = warmstart.WarmStartMLP(_dev(), seed=0)
    warmstart.train(model, ops, Y, label, epochs=6, batch=2048, lr=1e-3, lr_final=1e-5)

    tops, tch, _, tY = warmstart.sample_problems(g, X, 4096, seed=15)
    C_rand = warmstart.random_start(tops, 16)
    C0, used, lw, lr = warmstart.warm_start(model, X, tY, tops, tch, C_rand)
    start_loss = torch.where(used.bool(), lw, lr)
    fin = torch.isfinite(lr)
    assert bool((start_loss[fin] <= lr[fin]).all())
    frac_warm = float(used.double().mean())
    print(f"warm start taken for {frac_warm:.3f} of the problems", flush=True)
    assert frac_warm > 0.5
    fits = {}
    for name, C_start in (("randn", C_rand), ("warm", C0)):
        C_fit, *_ = _fit(X, tY, tops, tch, C_start, 100, 10, 1e-7, "strong_wolfe")
        fits[name] = int((warmstart.rel_mse_fp64(X, tY, tops, tch, C_fit) <= 1e-6).sum())
    print(f"fitted: {fits}", flush=True)
    assert fits["warm"] >= fits["randn"]
