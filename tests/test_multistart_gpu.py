"""Multi-start L-BFGS: K fits per expression in one launch.

The multiple-constant evaluation of the reference
(evaluate_multiple_constant_backend, src/symbolic_evaluation.cpp:171-305,
tests/test_evaluation_multiple_constants.py) evaluates K constant sets per
expression; optimize_multistart carries it through to the fit.  Each of the
K fits of an expression must be the fit that optimize computes from that
start: the same constants, loss, iteration and closure counts and error
code, bit for bit, for both line searches, with and without the early stop,
and with programs in every tape-row bucket.  The flattened fits also meet the
oracle bar of the single-start tests, and the chosen start is the one with
the lowest final loss.
"""
import numpy as np
import pytest
import torch

import helpers

pytestmark = pytest.mark.gpu


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def _batch(seed, B, M, K, N=64, n_x=2):
    import lbfgs_oracle

    rng = np.random.default_rng(seed)
    ops, ch = helpers.random_batch(seed, B, M, n_x, unary=[10, 11, 12, 14], binary=[40, 41, 42, 43],
                                   leaf_const_p=0.4, depth=5 if M <= 32 else 7, p_leaf=0.3 if M <= 32 else 0.15)
    X = rng.uniform(-2, 2, (N, n_x)).astype(np.float32)
    C_true = helpers.constants_like(ops, rng)
    Y = lbfgs_oracle.oracle.fused_root(X, ops, ch, C_true, np.zeros((B, N), np.float32), mode=0)["y"]
    with np.errstate(all="ignore"):
        keep = np.isfinite(Y).all(1) & (np.abs(Y).max(1) < 1e4)
    ops, ch, Y = ops[keep], ch[keep], Y[keep]
    C0 = np.stack([helpers.constants_like(ops, rng, scale=1.5) for _ in range(K)], axis=1)  # (B, K, M)
    return X, Y, ops, ch, C0


@pytest.mark.parametrize("line_search,stop,M", [("strong_wolfe", None, 32), ("backtracking", None, 32),
                                                 ("strong_wolfe", 1e-6, 32), ("strong_wolfe", None, 128)])
def test_multistart_equals_separate_fits(line_search, stop, M):
    import cudasymbolicregression_amd as st
    from cudasymbolicregression_amd.optimization import _fit

    K = 3
    print(f"case multistart {line_search} stop={stop} M={M} K={K}", flush=True)
    X, Y, ops, ch, C0 = _batch(3 if M == 32 else 4, 64, M, K)
    B = len(ops)
    Xd, Yd, Od, Cd = _t(X), _t(Y), _t(ops), _t(ch)
    Cb, Yb, info = st.optimize_multistart(Xd, Yd, Od, Cd, _t(C0), max_iter=40, line_search_fn=line_search,
                                          stop_rel_loss=stop, raise_errors=False)
    assert info["C"].shape == (B, K, M) and info["loss"].shape == (B, K)
    var = Yd.double().var(dim=1, unbiased=False)
    stop_t = None if stop is None else (stop * torch.where(var > 0, var, (Yd.double() ** 2).mean(1))).contiguous()
    for k in range(K):
        C, loss, iters, err, flags, ncl = _fit(Xd, Yd, Od, Cd, _t(C0[:, k]), 40, 10, 1e-7, line_search,
                                               stop_loss=stop_t)
        torch.testing.assert_close(info["C"][:, k], C, rtol=0, atol=0, equal_nan=True)
        torch.testing.assert_close(info["loss"][:, k], loss, rtol=0, atol=0, equal_nan=True)
        assert torch.equal(info["iters"][:, k], iters)
        assert torch.equal(info["closures"][:, k], ncl)
        assert torch.equal(info["err"][:, k], err)
    loss = info["loss"].cpu().numpy()
    ok = np.isfinite(loss) & (info["err"].cpu().numpy() == 0)
    key = np.where(ok, loss, np.inf)
    best = info["best"].cpu().numpy()
    np.testing.assert_array_equal(best, key.argmin(1))
    np.testing.assert_array_equal(Cb.cpu().numpy(), info["C"].cpu().numpy()[np.arange(B), best])
    print(f"{B} problems x {K} starts; starts chosen {np.bincount(best, minlength=K).tolist()}; "
          f"mean closures {float(info['closures'].double().mean()):.1f}", flush=True)


def test_multistart_fits_match_oracle():
    import cudasymbolicregression_amd as st

    K = 3
    print(f"case multistart vs oracle K={K}", flush=True)
    X, Y, ops, ch, C0 = _batch(5, 48, 32, K)
    B = len(ops)
    Cb, Yb, info = st.optimize_multistart(_t(X), _t(Y), _t(ops), _t(ch), _t(C0), max_iter=50, raise_errors=False)
    # every (problem, start) as its own problem of the oracle
    rep = lambda a: np.repeat(a, K, axis=0)  # noqa: E731
    opsK, chK, YK = rep(ops), rep(ch), rep(Y)
    C0K = C0.reshape(B * K, -1)
    fits = helpers.oracle_fits(X, YK, opsK, chK, C0K, max_iter=50)
    import lbfgs_oracle

    yg = lbfgs_oracle.oracle.fused_root(X, opsK, chK, info["C"].cpu().numpy().reshape(B * K, -1), YK, mode=1,
                                        want_gx=False)["y"]
    with np.errstate(all="ignore"):
        mse = ((yg.astype(np.float64) - YK) ** 2).mean(1)
    msg = helpers.compare_fit_to_oracle(mse, fits, info["err"].cpu().numpy().reshape(-1), f"multistart K={K}")
    print(msg, flush=True)


def test_multistart_rejects_bad_shapes():
    import cudasymbolicregression_amd as st

    X, Y, ops, ch, C0 = _batch(6, 8, 16, 2)
    with pytest.raises(RuntimeError, match=r"C_init must have shape \(B, K, M\)"):
        st.optimize_multistart(_t(X), _t(Y), _t(ops), _t(ch), _t(C0[:, 0]))
    with pytest.raises(RuntimeError, match=r"C_init must have shape \(B, K, M\)"):
        st.optimize_multistart(_t(X), _t(Y), _t(ops), _t(ch), _t(C0[:, :0]))
