"""The allocation-free fused evaluator (sr_fused_root_noalloc_f32) against
the workspace form (sr_fused_root_f32) through the C-ABI: equal bit for bit
on calls in a row, the second with fewer expressions, fewer points and
another mode, the third with programs long enough for global tapes; sizes
past the fixed scratch are refused without a launch."""
import numpy as np
import pytest
import torch

import helpers

pytestmark = pytest.mark.gpu


def _dev():
    return torch.device("cuda:0")


def _bits(t):
    return {torch.float32: lambda: t.view(torch.int32), torch.float64: lambda: t.view(torch.int64)}.get(
        t.dtype, lambda: t)()


def _call(lib, fn_name, X, ops, ch, C, dy, mode, ws=None):
    from cudasymbolicregression_amd import _lib

    B, M = ops.shape
    N, n_x = X.shape
    dev = X.device
    out = dict(y=torch.full((B, N), 7.0, device=dev), loss=torch.full((B,), 7.0, dtype=torch.float64, device=dev),
               grad_C=torch.full((B, M), 7.0, device=dev), err=torch.full((B,), 7, dtype=torch.int32, device=dev),
               flags=torch.full((2,), 7, dtype=torch.int32, device=dev))
    args = [_lib.ptr(X), N, n_x, _lib.ptr(ops), _lib.ptr(ch), _lib.ptr(C), M, B, _lib.ptr(dy), mode, 0,
            _lib.ptr(out["y"]), _lib.ptr(out["loss"]), _lib.ptr(out["grad_C"]), None, _lib.ptr(out["err"]),
            _lib.ptr(out["flags"])]
    if ws is not None:
        args += [_lib.ptr(ws), ws.numel()]
    rc = getattr(lib, fn_name)(*args, _lib.stream_of(dev))
    torch.cuda.synchronize()
    return rc, out


def _case(seed, B, M, N):
    rng = np.random.default_rng(seed)
    ops, ch = helpers.random_batch(seed, B, M, 3, unary=helpers.G2_UNARY, binary=helpers.G2_BINARY, depth=6)
    C = helpers.constants_like(ops, rng)
    X = rng.uniform(-3, 3, (N, 3)).astype(np.float32)
    dy = rng.standard_normal((B, N)).astype(np.float32)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(_dev())  # noqa: E731
    return t(X), t(ops), t(ch), t(C), t(dy)


def test_noalloc_equals_workspace_form_twice():
    from cudasymbolicregression_amd import _lib

    lib = _lib.hip()
    for seed, B, M, N, mode in ((1, 3000, 32, 300, 1), (2, 700, 16, 130, 0), (3, 10, 64, 50, 1)):
        print(f"case noalloc B={B} M={M} N={N} mode={mode}", flush=True)
        X, ops, ch, C, dy = _case(seed, B, M, N)
        ws = torch.empty(lib.sr_fused_workspace_bytes(N, 3, B, M, 0, 0), dtype=torch.uint8, device=_dev())
        rc_a, a = _call(lib, "sr_fused_root_f32", X, ops, ch, C, dy, mode, ws)
        rc_b, b = _call(lib, "sr_fused_root_noalloc_f32", X, ops, ch, C, dy, mode)
        assert rc_a == 0 and rc_b == 0
        for k in ("y", "grad_C", "err", "flags") + (("loss",) if mode == 1 else ()):
            assert torch.equal(_bits(a[k]), _bits(b[k])), k


def test_noalloc_refuses_sizes_past_its_scratch():
    from cudasymbolicregression_amd import _lib

    lib = _lib.hip()
    cap = lib.sr_fused_noalloc_max_bytes()
    assert cap == 128 << 20
    B = 1
    while lib.sr_fused_workspace_bytes(64, 3, B * 2, 32, 0, 0) <= cap:
        B *= 2
    assert B >= 262144
    big = B * 2
    X = torch.zeros(64, 3, device=_dev())
    ops = torch.zeros(big, 32, dtype=torch.int64, device=_dev())
    ch = torch.full((big, 32, 2), -1, dtype=torch.int64, device=_dev())
    C = torch.zeros(big, 32, device=_dev())
    dy = torch.zeros(big, 64, device=_dev())
    rc, out = _call(lib, "sr_fused_root_noalloc_f32", X, ops, ch, C, dy, 0)
    assert rc == -2
    assert (out["flags"] == 7).all()  # nothing was launched
