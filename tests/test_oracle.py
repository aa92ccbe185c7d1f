"""Pins the C oracle (CPU): against the reference's own CPU units (oracle/_ref,
bit for bit), against closed forms, and against FP64 finite differences."""
import ctypes

import numpy as np
import pytest

import helpers
import oracle

OPS_ALL = helpers.UNARY + helpers.BINARY

needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="reference CPU units not built (oracle/_ref)")


@pytest.fixture(scope="module", autouse=True)
def _single_thread_reference():
    # One OpenMP thread makes the reference's atomics run in loop order, so its
    # sums are comparable bit for bit with the oracle's sequential order.
    try:
        ctypes.CDLL("libgomp.so.1").omp_set_num_threads(1)
    except OSError:
        pass


def _program(seed, B=12, M=24, n_x=3, dtype=np.float32, **kw):
    rng = np.random.default_rng(seed)
    ops_bm, ch_bm = helpers.random_batch(seed, B, M, n_x, **kw)
    C_bm = helpers.constants_like(ops_bm, rng, dtype)
    X = (rng.standard_normal((37, n_x)) * 1.5).astype(dtype)
    ops, ch, C = helpers.to_mb(ops_bm, ch_bm, C_bm)
    return X, ops, ch, C, ops_bm, ch_bm, C_bm


@needs_ref
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("seed", range(6))
def test_forward_bit_equal_reference(seed, dtype):
    X, ops, ch, C, *_ = _program(seed, dtype=dtype)
    a = oracle.forward(X, ops, ch, C)
    b = oracle.forward(X, ops, ch, C, use_ref=True)
    assert np.array_equal(a, b, equal_nan=True)


@needs_ref
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("seed", range(6))
def test_backward_bit_equal_reference(seed, dtype):
    X, ops, ch, C, *_ = _program(seed, dtype=dtype)
    cache = oracle.forward(X, ops, ch, C)
    rng = np.random.default_rng(100 + seed)
    M, N, B = cache.shape
    g = np.zeros_like(cache)
    # upstream on the root of every expression only (what optimize feeds in)
    roots = np.maximum((ops != 0).sum(0) - 1, 0)
    for b in range(B):
        g[roots[b], :, b] = rng.standard_normal(N).astype(dtype)
    g[np.isnan(cache)] = 0
    gc1, gx1, e1 = oracle.backward(g, cache, ops, ch, X.shape[1])
    gc2, gx2, e2 = oracle.backward(g, cache, ops, ch, X.shape[1], use_ref=True)
    assert (e1 != 0) == (e2 != 0)
    if e1 == 0:
        assert np.array_equal(gc1, gc2, equal_nan=True)
        assert np.array_equal(gx1, gx2, equal_nan=True)


@needs_ref
def test_validate_matches_reference():
    rng = np.random.default_rng(7)
    for trial in range(200):
        M, B = int(rng.integers(1, 6)), int(rng.integers(1, 4))
        ops = rng.choice([0, 1, 10, 11, 40, 42, 50, 51, 9, 33], size=(M, B)).astype(np.int64)
        ch = rng.integers(-1, M + 1, size=(M, B, 2)).astype(np.int64)
        ch[rng.random((M, B, 2)) < 0.5] = -1
        assert oracle.validate(ops, ch) == oracle.validate(ops, ch, use_ref=True), (ops, ch)


@needs_ref
@pytest.mark.parametrize("seed", range(3))
def test_multi_bit_equal_reference(seed):
    X, ops, ch, C, *_ = _program(seed, B=5, M=12)
    rng = np.random.default_rng(seed)
    K = 3
    Ck = np.repeat(C[:, :, None], K, axis=2) + (ops[:, :, None] == 1) * rng.standard_normal(
        (*C.shape, K)).astype(np.float32)
    Ck = Ck.astype(np.float32)
    a = oracle.multi_forward(X, ops, ch, Ck)
    b = oracle.multi_forward(X, ops, ch, Ck, use_ref=True)
    assert np.array_equal(a, b, equal_nan=True)
    g = rng.standard_normal(a.shape).astype(np.float32)
    g[np.isnan(a)] = 0
    r1 = oracle.multi_backward(g, a, ops, ch, X.shape[1])
    r2 = oracle.multi_backward(g, a, ops, ch, X.shape[1], use_ref=True)
    assert (r1[2] != 0) == (r2[2] != 0)
    if r1[2] == 0:
        assert np.array_equal(r1[0], r2[0], equal_nan=True)
        assert np.array_equal(r1[1], r2[1], equal_nan=True)


def test_closed_forms():
    # C + 1 with C = 3.14, and sin(C)
    X = np.zeros((4, 1), np.float32)
    ops = np.array([[1], [2], [40]], np.int64)
    ch = np.array([[[-1, -1]], [[-1, -1]], [[0, 1]]], np.int64)
    C = np.array([[3.14], [0.0], [0.0]], np.float32)
    cache = oracle.forward(X, ops, ch, C)
    assert np.all(cache[0] == np.float32(3.14))
    assert np.all(cache[1] == 1.0)
    assert np.allclose(cache[2], 4.14, atol=1e-6)
    ops = np.array([[1], [10]], np.int64)
    ch = np.array([[[-1, -1]], [[0, -1]]], np.int64)
    C = np.array([[3.14], [0.0]], np.float32)
    cache = oracle.forward(X, ops, ch, C)
    assert np.allclose(cache[1], np.sin(np.float32(3.14)), atol=1e-6)


def test_domain_policy():
    X = np.array([[-1.0], [0.0], [2.0]], np.float32)
    for op, expect in [(13, [np.nan, np.nan, np.log(2.0)]), (15, [np.nan, 0.0, np.sqrt(2.0)]),
                       (28, [-1.0, np.nan, 0.5]), (17, [-np.pi / 2, 0.0, np.nan])]:
        ops = np.array([[50], [op]], np.int64)
        ch = np.array([[[-1, -1]], [[0, -1]]], np.int64)
        cache = oracle.forward(X, ops, ch, np.zeros((2, 1), np.float32))
        ok, msg = helpers.allclose_nan(cache[1, :, 0], np.array(expect, np.float64), atol=1e-6)
        assert ok, (op, msg)
    # x/0 -> NaN, pow(negative, .) -> NaN
    ops = np.array([[50], [51], [43], [44]], np.int64)
    ch = np.array([[[-1, -1]], [[-1, -1]], [[0, 1]], [[0, 1]]], np.int64)
    Xd = np.array([[1.0, 0.0], [-2.0, 3.0]], np.float32)
    cache = oracle.forward(Xd, ops, ch, np.zeros((4, 1), np.float32))
    assert np.isnan(cache[2, 0, 0]) and np.isclose(cache[2, 1, 0], -2.0 / 3.0)
    assert np.isnan(cache[3, 1, 0]) and cache[3, 0, 0] == 1.0


def test_noop_and_error_codes():
    X = np.ones((1, 2), np.float32)
    ops = np.array([[0]], np.int64)
    ch = -np.ones((1, 1, 2), np.int64)
    cache = oracle.forward(X, ops, ch, np.zeros((1, 1), np.float32))
    assert np.isnan(cache).all()
    _, _, e = oracle.backward(np.zeros_like(cache), cache, ops, ch, 2)
    assert e == 0
    _, _, e = oracle.backward(np.ones_like(cache), cache, ops, ch, 2)
    assert e == 201
    # 1 / X0 with X0 = 0 -> 204
    X = np.zeros((1, 2), np.float32)
    ops = np.array([[2], [50], [43]], np.int64)
    ch = np.array([[[-1, -1]], [[-1, -1]], [[0, 1]]], np.int64)
    cache = oracle.forward(X, ops, ch, np.zeros((3, 1), np.float32))
    g = np.zeros_like(cache)
    g[2] = 1.0
    _, _, e = oracle.backward(g, cache, ops, ch, 2)
    assert e == 204


@pytest.mark.parametrize("seed", range(4))
def test_gradients_match_finite_differences(seed):
    """FP64 oracle gradients of sum_n y_root(n) * w(n) against central differences."""
    rng = np.random.default_rng(seed)
    smooth_unary = [10, 11, 12, 14, 16, 19, 20, 21, 22, 27, 29, 30, 31]
    smooth_binary = [40, 41, 42]
    B, M, n_x = 6, 16, 2
    ops_bm, ch_bm = helpers.random_batch(seed, B, M, n_x, unary=smooth_unary, binary=smooth_binary, max_len=12)
    C_bm = helpers.constants_like(ops_bm, rng, np.float64, scale=0.5)
    X = rng.uniform(-1, 1, (9, n_x))
    w = rng.standard_normal((B, 9))
    out = oracle.fused_root(X, ops_bm, ch_bm, C_bm, w, mode=0)
    assert (out["err"] == 0).all()

    def objective(Xv, Cv):
        y = oracle.fused_root(Xv, ops_bm, ch_bm, Cv, w, mode=0, want_gx=False)["y"]
        return float((y * w).sum())

    h = 1e-6
    for b in range(B):
        for k in np.nonzero(ops_bm[b] == 1)[0]:
            Cp, Cm = C_bm.copy(), C_bm.copy()
            Cp[b, k] += h
            Cm[b, k] -= h
            fd = (objective(X, Cp) - objective(X, Cm)) / (2 * h)
            assert abs(fd - out["grad_C"][b, k]) <= 1e-5 * max(1.0, abs(fd)), (b, k, fd, out["grad_C"][b, k])
    for n in range(X.shape[0]):
        for v in range(n_x):
            Xp, Xm = X.copy(), X.copy()
            Xp[n, v] += h
            Xm[n, v] -= h
            fd = (objective(Xp, C_bm) - objective(Xm, C_bm)) / (2 * h)
            assert abs(fd - out["grad_X"][n, v]) <= 1e-5 * max(1.0, abs(fd)), (n, v, fd)


@pytest.mark.parametrize("mode", [0, 1])
def test_fused_equals_full_backward(mode):
    X, ops, ch, C, ops_bm, ch_bm, C_bm = _program(11, B=10, M=20, unary=helpers.G2_UNARY, binary=helpers.G2_BINARY)
    rng = np.random.default_rng(5)
    N = X.shape[0]
    dy = rng.standard_normal((10, N)).astype(np.float32)
    out = oracle.fused_root(X, ops_bm, ch_bm, C_bm, dy, mode=mode)
    cache = oracle.forward(X, ops, ch, C)
    roots = np.maximum((ops != 0).sum(0) - 1, 0)
    y = np.stack([cache[roots[b], :, b] for b in range(10)])
    assert np.array_equal(out["y"], y, equal_nan=True)
    g = np.zeros_like(cache)
    for b in range(10):
        if mode == 1:
            d = y[b] - dy[b]
            g[roots[b], :, b] = (np.float32(2.0) * d) * (np.float32(1.0) / np.float32(N))
            assert np.isclose(out["loss"][b], np.mean(d.astype(np.float64) ** 2), rtol=1e-12, equal_nan=True)
        else:
            g[roots[b], :, b] = dy[b]
    gc, gx, e = oracle.backward(g, cache, ops, ch, X.shape[1])
    ok_rows = out["err"] == 0
    assert (e != 0) == (~ok_rows).any()
    assert np.array_equal(out["grad_C"][ok_rows], gc.T[ok_rows], equal_nan=True)
