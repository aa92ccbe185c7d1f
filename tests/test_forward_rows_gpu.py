"""The full-cache forward with a workspace (row buffer, then a transposing
pass) against the same forward without one, through the C-ABI.

Both compute every node with the same kernel code, so the caches must be
identical bit for bit (NaN padding included), for every workspace size: the
recommended one (every group of 64 expressions through the row buffer), one
that holds only some groups (the others written directly), and the index
arrays alone (no group fits).  Shapes cover a partial last group of
expressions (B = 200), a partial last tile of points (N = 150), programs up to
M = 64 nodes, FP32 and FP64, and K = 3 constant sets.  The values themselves
are pinned against the oracle by test_eval_gpu.py, which runs the same path
through the public API."""
import numpy as np
import pytest
import torch

import helpers

pytestmark = pytest.mark.gpu


def _dev():
    return torch.device("cuda:0")


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(_dev())


def _inputs(seed, B, M, N, K, dtype):
    rng = np.random.default_rng(seed)
    ops_bm, ch_bm = helpers.random_batch(seed, B, M, 3, depth=6)
    X = rng.uniform(-3, 3, (N, 3)).astype(dtype)
    ops, ch = helpers.to_mb(ops_bm, ch_bm)
    C = (rng.standard_normal((M, B, K)) * (ops == 1)[:, :, None]).astype(dtype)
    return _t(X), _t(ops), _t(ch), _t(C)


def _forward(lib, X, ops, ch, C, ws_bytes, ws=None, fill=7.0):
    from cudasymbolicregression_amd import _lib

    M, B, K = C.shape
    N = X.shape[0]
    sfx = "f64" if X.dtype == torch.float64 else "f32"
    cache = torch.full((M, N, B, K), fill, dtype=X.dtype, device=X.device)
    s = _lib.stream_of(X.device)
    if ws_bytes is None:
        rc = getattr(lib, f"sr_multi_forward_mb_{sfx}")(_lib.ptr(X), N, 3, _lib.ptr(ops), _lib.ptr(ch), _lib.ptr(C),
                                                        M, B, K, _lib.ptr(cache), s)
    else:
        if ws is None:
            ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=X.device)
        assert ws.numel() >= ws_bytes
        if K == 1:
            rc = getattr(lib, f"sr_forward_mb_ws_{sfx}")(_lib.ptr(X), N, 3, _lib.ptr(ops), _lib.ptr(ch),
                                                         _lib.ptr(C), M, B, _lib.ptr(cache), _lib.ptr(ws), ws_bytes,
                                                         s)
        else:
            rc = getattr(lib, f"sr_multi_forward_mb_ws_{sfx}")(_lib.ptr(X), N, 3, _lib.ptr(ops), _lib.ptr(ch),
                                                               _lib.ptr(C), M, B, K, _lib.ptr(cache), _lib.ptr(ws),
                                                               ws_bytes, s)
    _lib.check_status(rc, "forward")
    torch.cuda.synchronize()
    return cache.cpu().numpy()


def _written(lib, X, ops, ch, C, ws_bytes, ws=None):
    """The cache from two calls whose outputs start from different fills:
    equal results show that every element was written."""
    a = _forward(lib, X, ops, ch, C, ws_bytes, ws, fill=7.0)
    b = _forward(lib, X, ops, ch, C, ws_bytes, ws, fill=-3.25)
    assert np.array_equal(a, b, equal_nan=True), "some cache elements were not written"
    return a


def _index_bytes(E):
    a = lambda x: (x + 255) // 256 * 256  # noqa: E731
    G = (E + 63) // 64
    return a(8 * (G + 1)) + a(8 * E) + a(E)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("K", [1, 3])
def test_row_buffer_forward_equals_direct(dtype, K):
    from cudasymbolicregression_amd import _lib

    lib = _lib.hip()
    B, M, N = 200, 64, 150
    print(f"case row-buffer forward dtype={np.dtype(dtype).name} K={K} B={B} M={M} N={N}", flush=True)
    X, ops, ch, C = _inputs(5 + K, B, M, N, K, dtype)
    ref = _written(lib, X, ops, ch, C, None)
    E = B * K
    rec = int(lib.sr_forward_mb_workspace_bytes(M, N, B, K, int(dtype == np.float64)))
    assert rec > _index_bytes(E)
    L = (ops.cpu().numpy() != 0) * np.arange(1, M + 1)[:, None]
    rows = np.repeat(L.max(0), K) * N  # row-buffer elements per expression
    some = int(rows[:64 * 2].sum())  # room for about the first two groups
    item = np.dtype(dtype).itemsize
    for ws_bytes in (rec, _index_bytes(E) + 256 + some * item, _index_bytes(E) + 256):
        got = _written(lib, X, ops, ch, C, ws_bytes)
        same = np.array_equal(got, ref, equal_nan=True)
        print(f"  workspace {ws_bytes} B: identical={same}", flush=True)
        assert same


def test_row_buffer_forward_twice_in_a_row():
    """A second call with a smaller batch in the same workspace (stale
    offsets of the first call must not leak)."""
    from cudasymbolicregression_amd import _lib

    lib = _lib.hip()
    print("case row-buffer forward twice", flush=True)
    X, ops, ch, C = _inputs(11, 300, 48, 96, 1, np.float32)
    ws_bytes = int(lib.sr_forward_mb_workspace_bytes(48, 96, 300, 1, 0))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=_dev())
    a = _written(lib, X, ops, ch, C, ws_bytes, ws)
    assert np.array_equal(a, _written(lib, X, ops, ch, C, None), equal_nan=True)
    X2, ops2, ch2, C2 = _inputs(12, 70, 48, 96, 1, np.float32)
    b = _written(lib, X2, ops2, ch2, C2, ws_bytes, ws)
    assert np.array_equal(b, _written(lib, X2, ops2, ch2, C2, None), equal_nan=True)
