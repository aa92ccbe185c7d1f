"""The wide fused evaluator (programs compiled once into plan records,
four consecutive points per lane per tape access) against the
one-wave-per-program kernels it replaces for FP32 calls with at most four
shared variables: y and the error codes equal bit for bit, loss, grad_C and
grad_X equal up to the order of their sums; grad_X the same bits on every
call.  Cases cover every operator, point counts that leave partial quads
and partial tile groups, one to four variables, X too large for the block's
LDS copy (each wave then keeps its variables in tape rows), programs with
more constants than the per-lane FP64 accumulators, and shared nodes."""
import numpy as np
import pytest
import torch

import helpers

pytestmark = pytest.mark.gpu

# the one-wave-per-program kernels of buckets 0 and 1 (variant 2 of each)
OLD = [2, 2, 0, 0, 0]


def _dev():
    return torch.device("cuda:0")


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(_dev())


def _compare(X, ops, ch, C, dy, mode, gx, fwd, label):
    from cudasymbolicregression_amd import fused

    a = fused.fused_root(X, ops, ch, C, dy, mode=mode, want_grad_X=gx, fwd_only=fwd, raise_errors=False)
    b = fused.fused_root(X, ops, ch, C, dy, mode=mode, want_grad_X=gx, fwd_only=fwd, raise_errors=False,
                         config=OLD)
    torch.cuda.synchronize()
    assert torch.equal(a["y"].view(torch.int32), b["y"].view(torch.int32)), label + " y"
    assert torch.equal(a["err"], b["err"]), label + " err"
    for k in ("grad_C", "loss", "grad_X"):
        if a.get(k) is None or b.get(k) is None:
            continue
        x, y = a[k].double(), b[k].double()
        assert torch.equal(torch.isfinite(x), torch.isfinite(y)), f"{label} {k} finite mask"
        fin = torch.isfinite(x)
        # the same terms summed in another order: FP64 sums rounded to FP32
        # (grad_C), FP64 (loss), FP32 partial rows then FP64 (grad_X)
        tol = 1e-12 if k == "loss" else 1e-5
        diff = (x[fin] - y[fin]).abs()
        bound = tol * y[fin].abs() + (1e-30 if k == "loss" else 1e-6)
        bad = int((diff > bound).sum())
        assert bad == 0, f"{label} {k}: {bad} values differ by more than the summation-order bound"


@pytest.mark.parametrize("seed,B,M,N,n_x", [(1, 2000, 32, 301, 3), (2, 1500, 16, 64, 2), (3, 1000, 64, 1024, 4),
                                            (4, 800, 32, 513, 1), (5, 600, 32, 2100, 4)])
def test_wide_equals_per_program_kernels(seed, B, M, N, n_x):
    rng = np.random.default_rng(seed)
    ops, ch = helpers.random_batch(seed, B, M, n_x, depth=5)
    C = helpers.constants_like(ops, rng)
    X = rng.uniform(-3, 3, (N, n_x)).astype(np.float32)
    dy = rng.standard_normal((B, N)).astype(np.float32)
    for mode, gx, fwd in ((0, True, False), (1, False, False), (0, False, True), (1, True, False)):
        label = f"seed={seed} B={B} M={M} N={N} n_x={n_x} mode={mode} gx={gx} fwd={fwd}"
        print("case", label, flush=True)
        _compare(_t(X), _t(ops), _t(ch), _t(C), _t(dy), mode, gx, fwd, label)


def test_wide_many_constants_and_shared_nodes():
    """Programs with up to 16 constants (past the 4 per-lane FP64
    accumulators) and programs whose nodes are shared (constants used twice:
    a wave sum per use; an instruction used twice: an extra adjoint row)."""
    rng = np.random.default_rng(7)
    B, M, N = 512, 40, 200
    rows = []
    for b in range(B):
        k = 2 + b % 15  # c, c, +, c, +, ... with k constants, then * x0
        r = [1, 1, 40]
        for _ in range(k - 2):
            r += [1, 40]
        r += [50, 42]
        rows.append(r + [0] * (M - len(r)))
    ops = np.array(rows, np.int64)
    ch = np.stack([helpers.children_of(r) for r in ops])
    # shared nodes: the second half reuses node 0 (a constant) as the right
    # child of the first addition, and the last addition's left child twice
    for b in range(B // 2, B):
        ch[b, 2, 1] = 0
        last = int((ops[b] != 0).sum()) - 1
        ch[b, last, 1] = ch[b, last, 0]
    C = helpers.constants_like(ops, rng)
    X = rng.uniform(-2, 2, (N, 1)).astype(np.float32)
    dy = rng.standard_normal((B, N)).astype(np.float32)
    for mode, gx in ((0, True), (1, False)):
        _compare(_t(X), _t(ops), _t(ch), _t(C), _t(dy), mode, gx, False, f"constants/shared mode={mode}")


def test_wide_grad_x_same_bits_every_call():
    from cudasymbolicregression_amd import fused

    rng = np.random.default_rng(9)
    B, M, N, n_x = 3000, 32, 777, 3
    ops, ch = helpers.random_batch(9, B, M, n_x, unary=helpers.G2_UNARY, binary=helpers.G2_BINARY, depth=5)
    C = helpers.constants_like(ops, rng)
    X = _t(rng.uniform(-3, 3, (N, n_x)).astype(np.float32))
    dy = _t(rng.standard_normal((B, N)).astype(np.float32))
    outs = [fused.fused_root(X, _t(ops), _t(ch), _t(C), dy, mode=0, want_grad_X=True, raise_errors=False)
            for _ in range(3)]
    torch.cuda.synchronize()
    for o in outs[1:]:
        assert torch.equal(o["grad_X"].view(torch.int32), outs[0]["grad_X"].view(torch.int32))
        assert torch.equal(o["grad_C"].view(torch.int32), outs[0]["grad_C"].view(torch.int32))
