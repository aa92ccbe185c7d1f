"""Seeded random expression programs for tests (valid postfix trees)."""
from __future__ import annotations

import numpy as np

UNARY = list(range(10, 32))
BINARY = [40, 41, 42, 43, 44]
G2_UNARY = [10, 11, 12, 13, 15]          # sin cos exp log sqrt
G2_BINARY = [40, 41, 42, 43]             # + - * /


def _gen_tree(rng, depth, n_x, unary, binary, p_leaf, leaf_const_p, hard_const_p, out):
    if depth == 0 or rng.random() < p_leaf:
        r = rng.random()
        if r < leaf_const_p:
            out.append(1)
        elif r < leaf_const_p + hard_const_p:
            out.append(int(rng.integers(2, 9)))
        else:
            out.append(50 + int(rng.integers(0, max(n_x, 1))))
        return
    if rng.random() < 0.4:
        _gen_tree(rng, depth - 1, n_x, unary, binary, p_leaf, leaf_const_p, hard_const_p, out)
        out.append(int(unary[int(rng.integers(0, len(unary)))]))
    else:
        _gen_tree(rng, depth - 1, n_x, unary, binary, p_leaf, leaf_const_p, hard_const_p, out)
        _gen_tree(rng, depth - 1, n_x, unary, binary, p_leaf, leaf_const_p, hard_const_p, out)
        out.append(int(binary[int(rng.integers(0, len(binary)))]))


def children_of(ops_row):
    """Postfix children table of one row (NO_OP padded)."""
    M = len(ops_row)
    ch = np.full((M, 2), -1, np.int64)
    st = []
    for k, op in enumerate(ops_row):
        if op == 0:
            break
        a = 1 if 10 <= op <= 31 else (2 if 40 <= op <= 44 else 0)
        if a == 1:
            ch[k, 0] = st.pop()
        elif a == 2:
            r = st.pop(); l = st.pop()
            ch[k, 0], ch[k, 1] = l, r
        st.append(k)
    return ch


def random_postfix(rng: np.random.Generator, M: int, n_x: int, max_len: int | None = None,
                   unary=UNARY, binary=BINARY, leaf_const_p: float = 0.3, hard_const_p: float = 0.1,
                   min_len: int = 1, depth: int = 6, p_leaf: float = 0.3):
    """One random postfix tree of length in [min_len, max_len], NO_OP padded to M.

    Returns (ops (M,), ch (M, 2)).
    """
    max_len = M if max_len is None else min(M, max_len)
    while True:
        out: list = []
        _gen_tree(rng, depth, n_x, unary, binary, p_leaf, leaf_const_p, hard_const_p, out)
        if min_len <= len(out) <= max_len:
            break
    o = np.zeros(M, np.int64)
    o[:len(out)] = out
    return o, children_of(o)


def random_batch(seed: int, B: int, M: int, n_x: int, **kw):
    """Batch-first (B, M) ops and (B, M, 2) children."""
    rng = np.random.default_rng(seed)
    ops = np.zeros((B, M), np.int64)
    ch = np.full((B, M, 2), -1, np.int64)
    for b in range(B):
        ops[b], ch[b] = random_postfix(rng, M, n_x, **kw)
    return ops, ch


def constants_like(ops: np.ndarray, rng: np.random.Generator, dtype=np.float32, scale=1.0):
    C = np.zeros(ops.shape, dtype)
    mask = ops == 1
    C[mask] = (rng.standard_normal(int(mask.sum())) * scale).astype(dtype)
    return C


def to_mb(ops_bm, ch_bm, C_bm=None):
    """Batch-first -> node-first copies."""
    ops = np.ascontiguousarray(ops_bm.T)
    ch = np.ascontiguousarray(ch_bm.transpose(1, 0, 2))
    if C_bm is None:
        return ops, ch
    return ops, ch, np.ascontiguousarray(C_bm.T)


def allclose_nan(a, b, rtol=1e-5, atol=1e-6):
    """|a-b| <= atol + rtol*max(|a|,|b|) with identical NaN and inf masks."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if a.shape != b.shape:
        return False, "shape"
    na, nb = np.isnan(a), np.isnan(b)
    if not np.array_equal(na, nb):
        return False, f"nan mask differs at {np.argwhere(na != nb)[:5].tolist()}"
    ia, ib = np.isinf(a), np.isinf(b)
    if not np.array_equal(ia, ib) or not np.array_equal(a[ia], b[ib]):
        return False, "inf mask differs"
    fin = ~(na | ia)
    d = np.abs(a[fin] - b[fin])
    tol = atol + rtol * np.maximum(np.abs(a[fin]), np.abs(b[fin]))
    bad = d > tol
    if bad.any():
        idx = np.argmax(d - tol)
        return False, f"max excess {float((d - tol)[idx])}: {a[fin][idx]} vs {b[fin][idx]}"
    return True, ""


def _cls(a):
    return np.where(np.isnan(a), 1, np.where(np.isposinf(a), 2, np.where(np.isneginf(a), 3, 0)))


def check_vs_f64(gpu, ref32, ref64, rtol=1e-5, atol=1e-6, factor=2.0):
    """Accuracy of a GPU result judged against the FP64 oracle.

    Where the reference's FP32 result (ref32) and the FP64 oracle (ref64) agree
    on NaN / +inf / -inf, the GPU must agree too; elsewhere (an ill-conditioned
    point where FP32 rounding decides the class) the point is skipped.  For
    finite points the GPU error against ref64 must not exceed
    factor * |ref32 - ref64| + rtol * |ref64| + atol: at least as accurate as
    the reference's FP32 path, within rtol.  Returns (ok, message).
    """
    g = np.asarray(gpu, np.float64)
    r32 = np.asarray(ref32, np.float64)
    r64 = np.asarray(ref64, np.float64)
    if g.shape != r64.shape or r32.shape != r64.shape:
        return False, f"shape {g.shape} {r32.shape} {r64.shape}"
    cg, c32, c64 = _cls(g), _cls(r32), _cls(r64)
    amb = c32 != c64
    badcls = (cg != c32) & ~amb
    if badcls.any():
        idx = np.argwhere(badcls)[:5].tolist()
        return False, f"class mismatch at {idx}: gpu {g[badcls][:5]} ref32 {r32[badcls][:5]}"
    fin = (cg == 0) & (c32 == 0) & ~amb
    eg = np.abs(g[fin] - r64[fin])
    e32 = np.abs(r32[fin] - r64[fin])
    bound = factor * e32 + rtol * np.abs(r64[fin]) + atol
    bad = eg > bound
    if bad.any():
        i = int(np.argmax(eg - bound))
        return False, (f"{int(bad.sum())} points exceed: worst gpu {g[fin][i]!r} ref32 {r32[fin][i]!r} "
                       f"ref64 {r64[fin][i]!r}")
    return True, f"{int(amb.sum())} ill-conditioned points skipped"
