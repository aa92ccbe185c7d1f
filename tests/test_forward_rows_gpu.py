"""The full-cache forward with a workspace (row buffer, then a transposing
pass) against the same forward without one, and the full-cache backward with
a row workspace (gathered rows) against the same backward without one,
through the C-ABI.

Both compute every node with the same kernel code, so the caches must be
identical bit for bit (NaN padding included), for every workspace size: the
recommended one (every group of 64 expressions through the row buffer), one
that holds only some groups (the others written directly), the index arrays
alone (no group fits), and none (a NULL workspace, which the API passes when
memory is short).  Shapes cover a partial last group of
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


def _inputs(seed, B, M, N, K, dtype, **kw):
    rng = np.random.default_rng(seed)
    ops_bm, ch_bm = helpers.random_batch(seed, B, M, 3, depth=6, **kw)
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
        if ws is None and ws_bytes > 0:
            ws = torch.empty(ws_bytes, dtype=torch.uint8, device=X.device)
        assert ws_bytes == 0 or ws.numel() >= ws_bytes  # 0: a NULL workspace
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
    for ws_bytes in (rec, _index_bytes(E) + 256 + some * item, _index_bytes(E) + 256, 0):
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


def _backward(lib, ops, ch, cache, g, n_x, K, rows_bytes, is64, recompute=None):
    """(grad_C, grad_X, err) of the multi-constant backward, direct
    (rows_bytes None) or through a row workspace of rows_bytes bytes; with
    recompute = (X, C), the entry point that recomputes the gathered groups'
    forward values from the forward's inputs instead of reading the cache."""
    from cudasymbolicregression_amd import _lib

    M, N, B = cache.shape[0], cache.shape[1], cache.shape[2]
    sfx = "f64" if is64 else "f32"
    dev = cache.device
    gC = torch.full((M, B, K), 5.0, dtype=cache.dtype, device=dev)
    gX = torch.full((N, n_x), 5.0, dtype=cache.dtype, device=dev)
    ws = torch.empty(max(int(lib.sr_multi_backward_mb_workspace_bytes(B, N, n_x, K, int(is64))), 256),
                     dtype=torch.uint8, device=dev)
    err = torch.empty(1, dtype=torch.int32, device=dev)
    gg = g.clone()  # a kernel that does not only read it may use it as scratch
    s = _lib.stream_of(dev)
    if rows_bytes is None:
        rc = getattr(lib, f"sr_multi_backward_mb_{sfx}")(
            _lib.ptr(gg), _lib.ptr(cache), _lib.ptr(ops), _lib.ptr(ch), M, B, N, n_x, K, _lib.ptr(gC), _lib.ptr(gX),
            _lib.ptr(ws), ws.numel(), _lib.ptr(err), s)
    elif recompute is None:
        rws = torch.empty(rows_bytes, dtype=torch.uint8, device=dev) if rows_bytes > 0 else None  # 0: NULL
        rc = getattr(lib, f"sr_multi_backward_mb_rows_{sfx}")(
            _lib.ptr(gg), _lib.ptr(cache), _lib.ptr(ops), _lib.ptr(ch), M, B, N, n_x, K, _lib.ptr(gC), _lib.ptr(gX),
            _lib.ptr(ws), ws.numel(), _lib.ptr(rws), rows_bytes, _lib.ptr(err), s)
    else:
        X, C = recompute
        rws = torch.empty(rows_bytes, dtype=torch.uint8, device=dev) if rows_bytes > 0 else None  # 0: NULL
        rc = getattr(lib, f"sr_multi_backward_mb_rows_recompute_{sfx}")(
            _lib.ptr(gg), _lib.ptr(cache), _lib.ptr(X), _lib.ptr(C), _lib.ptr(ops), _lib.ptr(ch), M, B, N, n_x, K,
            _lib.ptr(gC), _lib.ptr(gX), _lib.ptr(ws), ws.numel(), _lib.ptr(rws), rows_bytes, _lib.ptr(err), s)
    _lib.check_status(rc, "backward")
    torch.cuda.synchronize()
    return gC.cpu().numpy(), gX.cpu().numpy(), int(err.item())


@pytest.mark.parametrize("dtype,M", [(np.float32, 64), (np.float64, 32)])
@pytest.mark.parametrize("K", [1, 3])
def test_row_buffer_backward_equals_direct(dtype, M, K):
    """Gathered rows against the direct backward: grad_C, grad_X and the
    error code identical bit for bit, with the upstream gradient on every
    live row, also when the gathered groups recompute their forward values
    from X and C instead of reading the cache; then with only a nonzero
    gradient on one NO_OP padding row, which all report as 201."""
    from cudasymbolicregression_amd import _lib

    lib = _lib.hip()
    B, N = 200, 150
    is64 = dtype == np.float64
    print(f"case row-buffer backward dtype={np.dtype(dtype).name} K={K} B={B} M={M} N={N}", flush=True)
    X, ops, ch, C = _inputs(31 + K, B, M, N, K, dtype)
    cache = torch.from_numpy(_written(lib, X, ops, ch, C, None)).to(_dev())  # (M, N, B, K)
    rng = np.random.default_rng(40 + K)
    L = ((ops.cpu().numpy() != 0) * np.arange(1, M + 1)[:, None]).max(0)  # per expression b
    live = (np.arange(M)[:, None] < L[None, :])[:, None, :, None]  # (M, 1, B, 1)
    g_np = rng.standard_normal((M, N, B, K)) * live
    g_np[np.isnan(cache.cpu().numpy())] = 0
    g = torch.from_numpy(g_np.astype(dtype)).to(_dev())
    E = B * K
    rec = int(lib.sr_backward_mb_rows_workspace_bytes(M, N, B, K, int(is64)))
    rows = np.repeat(L, K) * N
    some = 2 * int(rows[:64 * 2].sum()) * np.dtype(dtype).itemsize  # about the first two groups
    sizes = (rec, _index_bytes(E) + 512 + some, _index_bytes(E) + 512, 0)
    ref = _backward(lib, ops, ch, cache, g, 3, K, None, is64)
    for rb in sizes:
        for rec_src in (None, (X, C)):
            got = _backward(lib, ops, ch, cache, g, 3, K, rb, is64, rec_src)
            same = all(np.array_equal(a, b, equal_nan=True) for a, b in zip(got[:2], ref[:2])) and got[2] == ref[2]
            print(f"  rows workspace {rb} B, recompute={rec_src is not None}: identical={same} err={got[2]}",
                  flush=True)
            assert same
    # only a gradient on a padding row of expression 3 (in the first group,
    # which fits in the recommended and the partial workspaces), on programs
    # without division: the backward reports the largest code, and a zero
    # denominator raises 204 whatever the upstream gradient
    X, ops, ch, C = _inputs(51 + K, B, M, N, K, dtype, unary=[10, 11, 12, 14], binary=[40, 41, 42])
    cache = torch.from_numpy(_written(lib, X, ops, ch, C, None)).to(_dev())
    L = ((ops.cpu().numpy() != 0) * np.arange(1, M + 1)[:, None]).max(0)
    b3 = 3
    assert L[b3] < M
    g2 = torch.zeros_like(cache)
    g2[L[b3], 7, b3, 0] = 1.0
    assert _backward(lib, ops, ch, cache, torch.zeros_like(cache), 3, K, None, is64)[2] == 0
    ref2 = _backward(lib, ops, ch, cache, g2, 3, K, None, is64)
    assert ref2[2] == 201
    for rb in sizes:
        for rec_src in (None, (X, C)):
            got2 = _backward(lib, ops, ch, cache, g2, 3, K, rb, is64, rec_src)
            print(f"  NO_OP gradient, rows workspace {rb} B, recompute={rec_src is not None}: err={got2[2]}",
                  flush=True)
            assert got2[2] == 201


def test_api_row_workspaces_match_direct(monkeypatch):
    """evaluate_backend forward and backward through autograd at a 64 MB cache
    (where the API uses the row workspaces) equal the same calls with the
    workspaces switched off: cache, grad_X and grad_C bit for bit."""
    from cudasymbolicregression_amd import _C
    from cudasymbolicregression_amd.evaluation import evaluate_backend

    B, M, N = 256, 64, 1024
    print(f"case API row workspaces B={B} M={M} N={N}", flush=True)
    assert M * N * B * 4 >= _C.ROW_WORKSPACE_MIN_CACHE_BYTES
    # no division: a zero denominator would raise 204 in the backward
    X, ops, ch, C = _inputs(61, B, M, N, 1, np.float32, unary=[10, 11, 12, 14], binary=[40, 41, 42])
    C2 = C[:, :, 0].contiguous()
    L = ((ops.cpu().numpy() != 0) * np.arange(1, M + 1)[:, None]).max(0)
    live = torch.from_numpy(np.arange(M)[:, None] < L[None, :]).to(_dev())[:, None, :]  # (M, 1, B)

    def run():
        Xr = X.clone().requires_grad_(True)
        Cr = C2.clone().requires_grad_(True)
        y = evaluate_backend(Xr, ops, ch, Cr)
        gen = torch.Generator(device=_dev()).manual_seed(3)
        g = torch.randn(y.shape, generator=gen, device=_dev()) * live
        g = torch.where(torch.isnan(y.detach()), torch.zeros_like(g), g)
        y.backward(g)
        torch.cuda.synchronize()
        return y.detach().cpu().numpy(), Xr.grad.cpu().numpy(), Cr.grad.cpu().numpy()

    a = run()
    monkeypatch.setattr(_C, "ROW_WORKSPACE_MIN_CACHE_BYTES", 1 << 62)
    b = run()
    for name, u, v in zip(("cache", "grad_X", "grad_C"), a, b):
        same = np.array_equal(u, v, equal_nan=True)
        print(f"  {name}: identical={same}", flush=True)
        assert same


def test_api_multiple_row_workspaces_match_direct(monkeypatch):
    """evaluate_multiple (K constant sets) forward and backward at a 67 MB
    cache, with and without the row workspaces: cache, grad_X and grad_C bit
    for bit."""
    from cudasymbolicregression_amd import _C
    from cudasymbolicregression_amd.evaluation import evaluate_multiple

    B, M, N, K = 64, 64, 1024, 4
    print(f"case API multiple row workspaces B={B} M={M} N={N} K={K}", flush=True)
    assert M * N * B * K * 4 >= _C.ROW_WORKSPACE_MIN_CACHE_BYTES
    rng = np.random.default_rng(71)
    ops_bm, ch_bm = helpers.random_batch(71, B, M, 3, depth=6, unary=[10, 11, 12, 14], binary=[40, 41, 42])
    X = _t(rng.uniform(-3, 3, (N, 3)).astype(np.float32))
    C = _t((rng.standard_normal((B, M, K)) * (ops_bm == 1)[:, :, None]).astype(np.float32))
    ops, ch = _t(ops_bm), _t(ch_bm)
    L = ((ops_bm != 0) * np.arange(1, M + 1)[None, :]).max(1)  # per expression b
    live = torch.from_numpy(np.arange(M)[None, :] < L[:, None]).to(_dev())[:, :, None, None]  # (B, M, 1, 1)

    def run():
        Xr = X.clone().requires_grad_(True)
        Cr = C.clone().requires_grad_(True)
        y = evaluate_multiple(Xr, ops, ch, Cr)  # (B, M, N, K)
        gen = torch.Generator(device=_dev()).manual_seed(5)
        g = torch.randn(y.shape, generator=gen, device=_dev()) * live
        g = torch.where(torch.isnan(y.detach()), torch.zeros_like(g), g)
        y.backward(g)
        torch.cuda.synchronize()
        return y.detach().cpu().numpy(), Xr.grad.cpu().numpy(), Cr.grad.cpu().numpy()

    a = run()
    monkeypatch.setattr(_C, "ROW_WORKSPACE_MIN_CACHE_BYTES", 1 << 62)
    b = run()
    for name, u, v in zip(("cache", "grad_X", "grad_C"), a, b):
        same = np.array_equal(u, v, equal_nan=True)
        print(f"  {name}: identical={same}", flush=True)
        assert same
