"""GPU parity of the evaluation kernels against the C oracle (FP32 within
|a-b| <= 1e-5*max(|a|,|b|) + 1e-6 with identical NaN/inf masks; FP64 within
1e-12 relative), through the public API and the C-ABI."""
import numpy as np
import pytest
import torch

import helpers
import oracle

pytestmark = pytest.mark.gpu

RTOL32, ATOL32 = 1e-5, 1e-6


def _dev():
    return torch.device("cuda:0")


def _case(seed, B=16, M=24, n_x=3, N=100, dtype=np.float32, **kw):
    rng = np.random.default_rng(seed)
    ops_bm, ch_bm = helpers.random_batch(seed, B, M, n_x, **kw)
    C_bm = helpers.constants_like(ops_bm, rng, dtype)
    X = rng.uniform(-3, 3, (N, n_x)).astype(dtype)
    return X, ops_bm, ch_bm, C_bm


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(_dev())


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("seed", range(4))
def test_backend_forward_matches_oracle(seed, dtype):
    from cudasymbolicregression_amd.evaluation import evaluate_backend

    print(f"case forward seed={seed} dtype={np.dtype(dtype).name}", flush=True)
    X, ops_bm, ch_bm, C_bm = _case(seed, dtype=dtype, N=130)
    ops, ch, C = helpers.to_mb(ops_bm, ch_bm, C_bm)
    got = evaluate_backend(_t(X), _t(ops), _t(ch), _t(C)).cpu().numpy()
    ref = oracle.forward(X, ops, ch, C)
    tol = (1e-12, 1e-13) if dtype == np.float64 else (RTOL32, ATOL32)
    ok, msg = helpers.allclose_nan(got, ref, *tol)
    assert ok, msg


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("seed", range(4))
def test_backend_backward_matches_oracle(seed, dtype):
    from cudasymbolicregression_amd.evaluation import evaluate_backend

    print(f"case backward seed={seed} dtype={np.dtype(dtype).name}", flush=True)
    X, ops_bm, ch_bm, C_bm = _case(seed, dtype=dtype, N=70, unary=helpers.G2_UNARY + [14, 27, 29])
    ops, ch, C = helpers.to_mb(ops_bm, ch_bm, C_bm)
    cache = oracle.forward(X, ops, ch, C)
    rng = np.random.default_rng(seed + 50)
    g = rng.standard_normal(cache.shape).astype(dtype)
    g[np.isnan(cache)] = 0
    gc_ref, gx_ref, e_ref = oracle.backward(g, cache, ops, ch, X.shape[1])
    Xt = _t(X).requires_grad_(True)
    Ct = _t(C).requires_grad_(True)
    out = evaluate_backend(Xt, _t(ops), _t(ch), Ct)
    if e_ref:
        with pytest.raises(RuntimeError, match="backward pass"):
            out.backward(_t(g))
        return
    out.backward(_t(g))
    tol = (1e-10, 1e-12) if dtype == np.float64 else (1e-4, 1e-5)
    ok, msg = helpers.allclose_nan(Ct.grad.cpu().numpy(), gc_ref, *tol)
    assert ok, "grad_C " + msg
    ok, msg = helpers.allclose_nan(Xt.grad.cpu().numpy(), gx_ref, *tol)
    assert ok, "grad_X " + msg


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


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("M", [12, 64, 128])
def test_fused_root_matches_oracle(M, dtype, mode):
    from cudasymbolicregression_amd import fused

    print(f"case fused M={M} dtype={np.dtype(dtype).name} mode={mode}", flush=True)
    seed = M + mode
    X, ops_bm, ch_bm, C_bm = _case(seed, B=300, M=M, N=200, dtype=dtype, unary=helpers.G2_UNARY,
                                   binary=helpers.G2_BINARY, depth=7)
    rng = np.random.default_rng(seed)
    dy = rng.standard_normal((300, 200)).astype(dtype)
    ref = oracle.fused_root(X, ops_bm, ch_bm, C_bm, dy, mode=mode)
    out = fused.fused_root(_t(X), _t(ops_bm), _t(ch_bm), _t(C_bm), _t(dy), mode=mode, want_grad_X=True,
                           raise_errors=False)
    tol = (1e-11, 1e-12) if dtype == np.float64 else (1e-4, 1e-5)
    ok, msg = helpers.allclose_nan(out["y"].cpu().numpy(), ref["y"], *(tol if dtype == np.float64 else (RTOL32,
                                                                                                        ATOL32)))
    assert ok, "y " + msg
    err = out["err"].cpu().numpy()
    assert np.array_equal(err != 0, ref["err"] != 0)
    good = ref["err"] == 0
    ok, msg = helpers.allclose_nan(out["grad_C"].cpu().numpy()[good], ref["grad_C"][good], *tol)
    assert ok, "grad_C " + msg
    if mode == 1:
        ok, msg = helpers.allclose_nan(out["loss"].cpu().numpy(), ref["loss"], 1e-6, 1e-9)
        assert ok, "loss " + msg
    if good.all():
        ok, msg = helpers.allclose_nan(out["grad_X"].cpu().numpy(), ref["grad_X"], 1e-4 if dtype == np.float32
                                       else 1e-10, 1e-4 if dtype == np.float32 else 1e-10)
        assert ok, "grad_X " + msg


def test_public_evaluate_shape_and_view():
    from cudasymbolicregression_amd.evaluation import evaluate

    X, ops_bm, ch_bm, C_bm = _case(3, B=7, M=10, N=33)
    y = evaluate(_t(X), _t(ops_bm), _t(ch_bm), _t(C_bm))
    assert y.shape == (7, 10, 33)
    ops, ch, C = helpers.to_mb(ops_bm, ch_bm, C_bm)
    ref = oracle.forward(X, ops, ch, C).transpose(2, 0, 1)
    ok, msg = helpers.allclose_nan(y.cpu().numpy(), ref, RTOL32, ATOL32)
    assert ok, msg
