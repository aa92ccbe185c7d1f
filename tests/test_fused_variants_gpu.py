"""Launch variants of the fused evaluator (sr_fused_root_opts_*) and its
autotuner (autotune.tune_fused).

* A batch whose programs fill every tape-row bucket (lengths 1-4, 5-8,
  9-16, 17-32 and 33-127, padded to 128) is evaluated once per (bucket,
  variant) with the other buckets at their defaults, with and without
  grad_X, in both modes, and at grad_X sizes that put the partial rows in
  LDS (N*n_x*4*waves <= 8 KiB), in the workspace, and past the register
  variables (n_x > 4).  A variant changes how many waves share a block and
  how many points a lane takes: each point's arithmetic is the same, so y
  and the error codes equal the default's bit for bit; only the order of the
  sums over points (FP64: loss, grad_C) and over expressions (FP32 partial
  rows, then FP64: grad_X) moves.  loss and grad_C must equal the default's
  within that FP64 rounding, and grad_C meets the oracle bar of
  test_eval_gpu.py (helpers.check_vs_f64: at least as accurate as the
  reference's FP32 algorithm against the FP64 oracle) wherever the
  reference's own grad_C is stable under a 2^-21 move of X or C (at most
  5% of the entries are not).  grad_X is the sum of the
  expressions' own FP32 contributions g_b (each taken from a call on that
  expression alone); any grouping of B such values into FP32 partial sums
  with FP64 sums on top lies within (B - 1) 2^-24 sum_b |g_b| (+ the FP64
  term) of their exact sum, and every variant must.  (The oracle parity of
  the default launch's grad_X is tested in test_eval_gpu.py; on this batch
  of deep sin/cos programs a third of the grad_X entries move by more than
  1e-5 relative when X moves by 2^-21, so the reference's own FP32 error is
  no bar there.)
* The FP64 evaluator has one variant per bucket; out-of-range choices are
  rejected.
* Without options the evaluator picks per N (buckets 0 and 1 take one
  point per lane when 64-point tiles leave at most 80% of the padded points
  of 128-point ones: N <= 64 or 129 <= N <= 192): the result equals the
  explicit choice bit for bit.
* tune_fused times the variants, logs every trial, stores the fastest
  configuration under the workload key, and fused_root finds it there.
"""
import numpy as np
import pytest
import torch

import helpers
import oracle
from cudasymbolicregression_amd import autotune, fused

pytestmark = pytest.mark.gpu


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def _batch(seed, per_range, n_x, N):
    rng = np.random.default_rng(seed)
    M = 128
    ops, ch = [], []
    for lo, hi in ((1, 4), (5, 8), (9, 16), (17, 32), (33, 127)):
        o, c = helpers.random_batch(int(rng.integers(1 << 30)), per_range, M, n_x, unary=[10, 11],
                                    binary=[40, 41, 42], min_len=lo, max_len=hi, depth=9, p_leaf=0.15)
        ops.append(o)
        ch.append(c)
    ops, ch = np.concatenate(ops), np.concatenate(ch)
    C = helpers.constants_like(ops, rng)
    X = rng.uniform(-2, 2, (N, n_x)).astype(np.float32)
    dy = rng.standard_normal((len(ops), N)).astype(np.float32)
    return X, ops, ch, C, dy


def _same_bits(a, b):
    return np.array_equal(a.view(np.int32), b.view(np.int32))


@pytest.mark.parametrize("n_x,N", [(3, 150), (3, 700), (6, 90)])
def test_every_variant_matches_default(n_x, N):
    nvar = [len(v) for v in autotune.fused_variants(False)]
    print(f"case fused variants n_x={n_x} N={N} variants/bucket {nvar}", flush=True)
    X, ops, ch, C, dy = _batch(10 * n_x + N, 40, n_x, N)
    rows = np.array([int((o != 0).sum()) for o in ops])
    assert rows.min() <= 4 and rows.max() > 32
    ref = {m: (oracle.fused_root(X, ops, ch, C, dy, mode=m),
               oracle.fused_root(X.astype(np.float64), ops, ch, C.astype(np.float64), dy.astype(np.float64), mode=m))
           for m in (0, 1)}
    stable_gc = {}
    for m in (0, 1):
        pert = iter([oracle.fused_root(x, ops, ch, c, dy, mode=m)["grad_C"]
                     for x, c in ((X * (1 + 2.0 ** -21), C), (X * (1 - 2.0 ** -21), C),
                                  (X, C * (1 + 2.0 ** -21)), (X, C * (1 - 2.0 ** -21)))])
        ug = helpers.unstable_points(lambda x, c: next(pert), X, C, ref[m][0]["grad_C"], 0.0, 1e-5, 1e-5)
        assert ug.mean() <= 0.05
        stable_gc[m] = ~ug
    tX, tops, tch, tC, tdy = _t(X), _t(ops), _t(ch), _t(C), _t(dy)
    for mode in (0, 1):
        for gx in (False, True):
            base = fused.fused_root(tX, tops, tch, tC, tdy, mode=mode, want_grad_X=gx, raise_errors=False,
                                    config=[0] * 5)
            base = {k: v.cpu().numpy() for k, v in base.items()}
            r32, r64 = ref[mode]
            if gx:
                g = np.stack([fused.fused_root(tX, tops[b:b + 1], tch[b:b + 1], tC[b:b + 1], tdy[b:b + 1], mode=mode,
                                               want_grad_X=True, config=[0] * 5)["grad_X"].cpu().numpy()
                              for b in range(len(ops))]).astype(np.float64)
                exact, mag = g.sum(0), np.abs(g).sum(0)
                tol_gx = (len(ops) - 1) * 2.0 ** -24 * mag + 2.0 ** -40 * mag + 1e-30
            assert np.array_equal(base["err"], r32["err"]) and (base["err"] == 0).all()
            for b in range(5):
                for v in range(nvar[b]):
                    cfg = [0] * 5
                    cfg[b] = v
                    print(f"  mode={mode} grad_X={gx} config {cfg}", flush=True)
                    out = fused.fused_root(tX, tops, tch, tC, tdy, mode=mode, want_grad_X=gx, raise_errors=False,
                                           config=cfg)
                    out = {k: x.cpu().numpy() for k, x in out.items()}
                    assert _same_bits(out["y"], base["y"]), cfg
                    assert np.array_equal(out["err"], base["err"]), cfg
                    if mode == 1:
                        np.testing.assert_allclose(out["loss"], base["loss"], rtol=1e-12, atol=0)
                    d = np.abs(out["grad_C"].astype(np.float64) - base["grad_C"])
                    assert (d <= 2.0 ** -23 * np.abs(base["grad_C"]) + 1e-9).all(), (cfg, d.max())
                    st = stable_gc[mode]
                    ok, msg = helpers.check_vs_f64(out["grad_C"][st], r32["grad_C"][st], r64["grad_C"][st], atol=1e-5)
                    assert ok, f"{cfg} grad_C {msg}"
                    if gx:
                        e = np.abs(out["grad_X"] - exact)
                        assert (e <= tol_gx).all(), (cfg, float((e / tol_gx).max()))


def test_fp64_single_variant_and_bad_config():
    assert [len(v) for v in autotune.fused_variants(True)] == [1] * 5
    X, ops, ch, C, dy = _batch(3, 8, 2, 64)
    t64 = lambda a: _t(a.astype(np.float64))  # noqa: E731
    a = fused.fused_root(t64(X), _t(ops), _t(ch), t64(C), t64(dy), mode=1, config=[0] * 5)
    b = fused.fused_root(t64(X), _t(ops), _t(ch), t64(C), t64(dy), mode=1)
    assert torch.equal(a["y"], b["y"]) and torch.equal(a["loss"], b["loss"])
    with pytest.raises(ValueError):
        fused.fused_root(t64(X), _t(ops), _t(ch), t64(C), t64(dy), mode=1, config=[1, 0, 0, 0, 0])
    with pytest.raises(ValueError):
        fused.fused_root(_t(X), _t(ops), _t(ch), _t(C), _t(dy), mode=1, config=[0, 0, 0, 0, 9])


def test_tune_fused_stores_and_is_used():
    X, ops, ch, C, dy = _batch(5, 200, 3, 256)
    tX, tops, tch, tC, tdy = _t(X), _t(ops), _t(ch), _t(C), _t(dy)
    autotune.cache_clear()
    log = []
    try:
        r = autotune.tune_fused(tX, tops, tch, tC, tdy, mode=1, want_grad_X=True, reps=3, log=log)
        nvar = [len(v) for v in autotune.fused_variants(False)]
        # the automatic choice (-1), then every variant of each bucket that has several
        assert len(r["trials"]) == 1 + sum(n for n in nvar if n > 1) == len(log)
        assert r["ms"] == pytest.approx(min(t for _, t in r["trials"]) * 1e3)
        print("\n".join(log), flush=True)
        got = autotune.fused_lookup(tX.device, torch.float32, 128, 256, 3, True, False)
        assert got == r["config"]
        # fused_root takes the stored configuration when none is given
        a = fused.fused_root(tX, tops, tch, tC, tdy, mode=1, want_grad_X=True)
        b = fused.fused_root(tX, tops, tch, tC, tdy, mode=1, want_grad_X=True, config=r["config"])
        for k in ("y", "loss", "grad_C", "grad_X"):
            assert torch.equal(a[k], b[k]), k
    finally:
        autotune.cache_clear()


@pytest.mark.parametrize("N,auto", [(64, [1, 1, 0, 0, 0]), (150, [1, 1, 0, 0, 0]), (200, [0] * 5), (1000, [0] * 5)])
def test_automatic_choice_by_point_count(N, auto):
    print(f"case automatic variants N={N}", flush=True)
    X, ops, ch, C, dy = _batch(7, 20, 3, N)
    tX, tops, tch, tC, tdy = _t(X), _t(ops), _t(ch), _t(C), _t(dy)
    a = fused.fused_root(tX, tops, tch, tC, tdy, mode=1, want_grad_X=True)
    b = fused.fused_root(tX, tops, tch, tC, tdy, mode=1, want_grad_X=True, config=auto)
    c = fused.fused_root(tX, tops, tch, tC, tdy, mode=1, want_grad_X=True, config=[-1] * 5)
    for k in ("y", "loss", "grad_C", "grad_X", "err"):
        assert torch.equal(a[k], b[k]) and torch.equal(a[k], c[k]), k
