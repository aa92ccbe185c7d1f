"""Seeded random expression programs for tests (valid postfix trees)."""
from __future__ import annotations

import numpy as np

UNARY = list(range(10, 32))
BINARY = [40, 41, 42, 43, 44]
G2_UNARY = [10, 11, 12, 13, 15]          # sin cos exp log sqrt
G2_BINARY = [40, 41, 42, 43]             # + - * /

# Grammar G2 of the benchmark workloads (3 variables, sin/cos/exp/log/sqrt).
G2 = """E -> E + T [0.3]
E -> E - T [0.1]
E -> T [0.6]
T -> T * F [0.3]
T -> T / F [0.1]
T -> F [0.6]
F -> R ( E ) [0.15]
F -> V [0.85]
R -> sin [0.2]
R -> cos [0.2]
R -> exp [0.2]
R -> log [0.2]
R -> sqrt [0.2]
V -> X_0 [0.25]
V -> X_1 [0.25]
V -> X_2 [0.25]
V -> C [0.25]
"""


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


def unstable_points(fn, X, C, base, delta, rtol, atol):
    """Points where the oracle's own result is not stable: re-evaluating with
    X or C scaled by (1 +- delta) changes the NaN/inf class or moves the value
    by more than rtol*max(|a|,|b|) + atol.  fn(X, C) -> array shaped like base."""
    base = np.asarray(base, np.float64)
    sens = np.zeros(base.shape, bool)
    cb = _cls(base)
    for Xp, Cp in ((X * (1 + delta), C), (X * (1 - delta), C), (X, C * (1 + delta)), (X, C * (1 - delta))):
        r = np.asarray(fn(Xp.astype(X.dtype), Cp.astype(C.dtype)), np.float64)
        cr = _cls(r)
        fin = (cr == 0) & (cb == 0)
        d = np.zeros(base.shape)
        d[fin] = np.abs(r[fin] - base[fin]) - (rtol * np.maximum(np.abs(r[fin]), np.abs(base[fin])) + atol)
        sens |= (cr != cb) | (d > 0)
    return sens


def check_stable(gpu, ref, unstable, rtol=1e-5, atol=1e-6, max_unstable=0.05, extra_tol=None):
    """Strict parity (same NaN/inf class, |a-b| <= rtol*max(|a|,|b|) + atol
    [+ extra_tol]) at every point where the reference is stable; unstable
    points are skipped, and they may be at most max_unstable of all points."""
    g = np.asarray(gpu, np.float64)
    r = np.asarray(ref, np.float64)
    frac = float(unstable.mean()) if unstable.size else 0.0
    if frac > max_unstable:
        return False, f"{frac:.4f} of the points are unstable"
    keep = ~unstable
    cg, cr = _cls(g), _cls(r)
    bad = (cg != cr) & keep
    if bad.any():
        return False, f"class mismatch at stable points {np.argwhere(bad)[:5].tolist()}: {g[bad][:3]} vs {r[bad][:3]}"
    fin = keep & (cg == 0)
    d = np.abs(g[fin] - r[fin]) - (rtol * np.maximum(np.abs(g[fin]), np.abs(r[fin])) + atol)
    if extra_tol is not None:
        d = d - np.broadcast_to(extra_tol, g.shape)[fin]
    if (d > 0).any():
        i = int(np.argmax(d))
        return False, f"{int((d > 0).sum())} stable points exceed: {g[fin][i]!r} vs {r[fin][i]!r}"
    return True, f"{frac:.4%} unstable points skipped"


def unstable_tape_points(X, ops_bm, ch_bm, C_bm, delta, rtol, atol):
    """(B, N) mask of points where ANY node of the program is unstable in the
    sense of unstable_points.  A root can look stable (NaN under every
    perturbation) while an inner node is chaotic (cos of a huge argument
    feeding a log), and then a last-ulp difference of the device libm flips
    the root's class; such points are only visible on the whole tape."""
    import oracle

    om = np.ascontiguousarray(ops_bm.T)
    cm = np.ascontiguousarray(ch_bm.transpose(1, 0, 2))
    Cm = np.ascontiguousarray(C_bm.T)
    live = (om != 0)[:, None, :]  # (M, 1, B)
    base = oracle.forward(X, om, cm, Cm)  # (M, N, B)
    sens = unstable_points(lambda x, c: oracle.forward(x, om, cm, np.ascontiguousarray(c.T)), X, C_bm, base, delta,
                           rtol, atol)
    return (sens & live).any(axis=0).T  # (B, N)


def rounding_bound(X, ops_bm, ch_bm, C_bm, ulps=8, eps=2.0 ** -24):
    """(B, N) first-order bound on how far two evaluations of the root
    (unit roundoff eps: 2^-24 for FP32, 2^-53 for FP64) may drift apart when
    every node's libm result may differ by `ulps` units in the last place:
    ulps * eps * sum_k |dy/dv_k| * |v_k|, with values
    and adjoints from the FP64 oracle (non-finite terms give inf).  This is
    the classical forward error bound; it depends on the program and the
    point only, never on the implementation under test."""
    import oracle

    om = np.ascontiguousarray(ops_bm.T)
    cm = np.ascontiguousarray(ch_bm.transpose(1, 0, 2))
    Cm = np.ascontiguousarray(C_bm.T, np.float64)
    M, B = om.shape
    N, n_x = X.shape
    v = oracle.forward(X.astype(np.float64), om, cm, Cm)
    root = np.maximum((om != 0).sum(axis=0) - 1, 0)
    g = np.zeros_like(v)
    g[root, :, np.arange(B)] = 1.0
    with np.errstate(all="ignore"):
        _, _, _, adj = oracle.backward(g, v, om, cm, n_x, return_adjoints=True)
        t = np.abs(adj) * np.abs(v)
        t[~np.isfinite(t) & (om != 0)[:, None, :]] = np.inf
        t[(om == 0)[:, None, :].repeat(N, axis=1)] = 0.0
        s = t.sum(axis=0).T  # (B, N)
    return ulps * eps * s


def fit_mse_tolerance(mse_ref):
    """Per-problem parity tolerance of a fit's final MSE: the reference runs in
    FP32, the device sums in FP64, so trajectories agree to a relative 1e-3."""
    return 1e-3 * np.abs(mse_ref) + 1e-7


def oracle_fits(X, Y, ops, ch, C0, rel_x=2.0 ** -21, **kw):
    """The oracle's fit with the reference's FP32 arithmetic
    (lbfgs_oracle.lbfgs_run, pinned bit for bit to the reference's
    trajectories), with the device's FP64 sums, inner products and Wolfe
    tests (lbfgs_oracle.lbfgs_run64), and three more times with the device's
    arithmetic on perturbed inputs: X scaled by 1 +- rel_x, and every
    closure value scaled by 1 +- 2^-19 (random signs).  The perturbed runs
    tell which problems the algorithm determines: where one ends elsewhere,
    the outcome hinges on the last ulp of the node arithmetic (say cos of a
    1e6 argument, which moves 0.25 rad per FP32 ulp) or of the sums, where
    the device's transcendentals and reduction order and the host's
    legitimately differ.  Returns dict(C32, mse32, C64, mse64, C64x,
    mse64x), the x entries being the perturbed run farthest from the
    unperturbed one; mse is the FP64 MSE of the final prediction at the
    unperturbed X."""
    import lbfgs_oracle

    Y = np.asarray(Y, np.float32)
    X = np.asarray(X, np.float32)
    out = {}

    def mse_at(C):
        y = lbfgs_oracle.oracle.fused_root(X, ops, ch, C, Y, mode=1, want_gx=False)["y"]
        with np.errstate(all="ignore"):
            return ((y.astype(np.float64) - Y) ** 2).mean(1)

    def fit(Xr, prec, noise=None):
        C, _, _, _ = lbfgs_oracle.optimize(Xr.astype(np.float32), Y, ops, ch, C0, precision=prec, noise=noise, **kw)
        return C, mse_at(C)

    out["C32"], out["mse32"] = fit(X, "fp32")
    out["C64"], out["mse64"] = fit(X, "fp64")
    # the perturbed runs; C64x / mse64x hold, per problem, the one farthest
    # from the unperturbed device-precision result
    runs = [fit(X * np.float32(1 + rel_x), "fp64"), fit(X * np.float32(1 - rel_x), "fp64"),
            fit(X, "fp64", noise=(2.0 ** -19, 77))]
    with np.errstate(invalid="ignore"):
        dm = np.stack([np.nan_to_num(np.abs(m - out["mse64"]) / fit_mse_tolerance(out["mse64"]), nan=np.inf)
                       for _, m in runs])
        dc = np.stack([np.nan_to_num((np.abs(C - out["C64"]) / (1 + np.abs(out["C64"]))).max(1), nan=np.inf)
                       for C, _ in runs])
    wm, wc = dm.argmax(0), dc.argmax(0)
    idx = np.arange(len(Y))
    out["mse64x"] = np.stack([m for _, m in runs])[wm, idx]
    out["C64x"] = np.stack([C for C, _ in runs])[wc, idx]
    return out


def compare_fit_to_oracle(mse_gpu, fits, err_gpu, label, max_undetermined=0.15, max_chaotic=0.02, extra_tol=None):
    """The device's final MSE against the oracle (oracle_fits), per problem
    without an error code, with tolerance fit_mse_tolerance (+ extra_tol):
    * against the device-precision oracle on every finite problem that the
      algorithm determines (every perturbed run ends within the tolerance),
      except at most max_chaotic of them (at least one), which are listed:
      the perturbations sample a trajectory's sensitivity without bounding
      it, and a chaotic one still parts (a jump to a point where every trial
      is rejected gives a zero step, after which the reference's rules give
      zero directions for good);
    * against the reference-precision oracle on EVERY determined problem
      where the two oracle precisions agree and the device matches the
      device-precision run.  Where the precisions disagree, the reference's
      FP32 sums and the device's FP64 ones take the problem to different
      ends; those are counted.
    Undetermined problems are counted (at most max_undetermined of the
    finite ones).  Returns a printable summary; raises AssertionError."""
    m32, m64, m64x = fits["mse32"], fits["mse64"], fits["mse64x"]
    extra = 0.0 if extra_tol is None else extra_tol
    ok = (np.asarray(err_gpu) == 0) & np.isfinite(m64)
    with np.errstate(invalid="ignore"):
        det = ok & (np.abs(m64x - m64) <= fit_mse_tolerance(m64))
        c64 = np.abs(mse_gpu - m64) <= fit_mse_tolerance(m64) + extra
        agree = np.isfinite(m32) & (np.abs(m32 - m64) <= fit_mse_tolerance(m32))
        c32 = np.abs(mse_gpu - m32) <= fit_mse_tolerance(m32) + extra
    und = np.nonzero(ok & ~det)[0]
    out64 = np.nonzero(det & ~c64)[0]
    out32 = np.nonzero(det & agree & c64 & ~c32)[0]
    msg = (f"{label}: {ok.sum()} finite problems, {det.sum()} determined; device vs device-precision oracle: "
           f"{len(out64)} outside {[(int(i), float(mse_gpu[i]), float(m64[i])) for i in out64[:6]]}; "
           f"precisions agree on {int((det & agree).sum())}, device vs reference-precision oracle there: "
           f"{len(out32)} outside {[(int(i), float(mse_gpu[i]), float(m32[i])) for i in out32[:6]]}; "
           f"undetermined {len(und)} (device off the device-precision oracle on "
           f"{int((~c64[und]).sum()) if len(und) else 0} of them)")
    assert len(out64) <= max(1, int(max_chaotic * det.sum())), msg
    assert len(out32) == 0, msg
    assert len(und) <= max_undetermined * max(ok.sum(), 1), msg
    return msg


def compare_constants_to_oracle(Cg, fits, label, max_undetermined=0.15):
    """Short fits: constants within 1e-3 (1 + |C|) of the device-precision
    oracle's on every finite problem that the algorithm determines (as in
    compare_fit_to_oracle, judged on the constants), and of the
    reference-precision oracle's wherever the two precisions agree as well.
    Returns a printable summary; raises AssertionError."""
    C32, C64, C64x = fits["C32"], fits["C64"], fits["C64x"]

    def close(a, b):
        with np.errstate(invalid="ignore"):
            return (np.abs(a - b) <= 1e-3 * (1 + np.abs(b))).all(1)

    fin = np.isfinite(C64).all(1) & np.isfinite(Cg).all(1)
    det = fin & close(C64x, C64)
    out64 = np.nonzero(det & ~close(Cg, C64))[0]
    agree = np.isfinite(C32).all(1) & close(C32, C64)
    out32 = np.nonzero(det & agree & ~close(Cg, C32))[0]
    und = np.nonzero(fin & ~det)[0]
    msg = (f"{label}: {fin.sum()} finite, {det.sum()} determined; vs device-precision oracle {len(out64)} outside "
           f"{out64[:8].tolist()}; precisions agree on {int((det & agree).sum())}, vs reference-precision oracle "
           f"there {len(out32)} outside {out32[:8].tolist()}; undetermined {und[:8].tolist()}")
    assert len(out64) == 0, msg
    assert len(out32) == 0, msg
    assert len(und) <= max_undetermined * max(fin.sum(), 1), msg
    return msg
