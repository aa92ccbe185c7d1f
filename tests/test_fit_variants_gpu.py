"""The device L-BFGS with the backtracking line search, every launch variant
of every bucket, and the autotuner, each against the oracle's L-BFGS.

* Backtracking (symbolic_torch/lbfgs_batch.py:105-136 for the Armijo
  search, :266-326 for the step): random programs fitted from random
  starts; the final MSE matches the oracle's backtracking fits
  (helpers.compare_fit_to_oracle, the same per-problem bar as the
  strong-Wolfe tests).
* Launch variants: a batch whose programs fill every tape-row bucket
  (lengths 1-8, 9-16, 17-32, 33-64, 65-127, padded 128) is fitted once per
  (bucket, variant) with the other buckets at their defaults.  A variant
  changes how many waves share a problem and how many points a lane takes,
  so only the order of the FP64 sums moves; every variant's fits meet the
  oracle bar.  Long random programs with many constants are often
  ill-conditioned: of the 33-127-node ones, about 40% end elsewhere when X
  moves by 2^-21 (the oracle's own perturbed runs), so up to 25% of this
  batch may be undetermined (counted per length range and printed); the
  determined ones are held to the full bar.
* tune_fit times the variants on that batch, logs every trial, stores the
  fastest configuration under the workload's key; optimize then runs with
  it (the cache lookup) and still meets the oracle bar.
"""
import numpy as np
import pytest
import torch

import helpers
import lbfgs_oracle

pytestmark = pytest.mark.gpu


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def _problems(seed, B, M, n_x, N, lengths, **kw):
    """Random programs with lengths in each of `lengths` ranges (B per
    range), targets from random true constants, finite on the points."""
    rng = np.random.default_rng(seed)
    ops_l, ch_l = [], []
    for i, (lo, hi) in enumerate(lengths):
        depth = max(3, int(np.ceil(np.log2(hi + 1))) + 1)
        o, c = helpers.random_batch(seed * 31 + i, B, M, n_x, min_len=lo, max_len=hi, depth=depth,
                                    p_leaf=0.3 if hi <= 32 else 0.1, **kw)
        ops_l.append(o)
        ch_l.append(c)
    ops, ch = np.concatenate(ops_l), np.concatenate(ch_l)
    X = rng.uniform(-2, 2, (N, n_x)).astype(np.float32)
    C_true = helpers.constants_like(ops, rng, scale=1.0)
    Y = lbfgs_oracle.oracle.fused_root(X, ops, ch, C_true, np.zeros((len(ops), N), np.float32), mode=0)["y"]
    with np.errstate(all="ignore"):
        keep = np.isfinite(Y).all(1) & (np.abs(Y).max(1) < 1e4)
    ops, ch, Y = ops[keep], ch[keep], Y[keep]
    C0 = helpers.constants_like(ops, rng, scale=1.0)
    return X, Y, ops, ch, C0


def _mse(Yg, Y):
    with np.errstate(all="ignore"):
        return ((Yg.cpu().numpy().astype(np.float64) - Y) ** 2).mean(1)


@pytest.mark.parametrize("seed", [0, 1])
def test_optimize_backtracking_matches_oracle(seed):
    import cudasymbolicregression_amd as st

    print(f"case backtracking seed={seed}", flush=True)
    X, Y, ops, ch, C0 = _problems(seed, 48, 32, 2, 64, [(1, 8), (9, 32)], unary=[10, 11, 12, 14],
                                  binary=[40, 41, 42], leaf_const_p=0.4)
    fits = helpers.oracle_fits(X, Y, ops, ch, C0, max_iter=50, line_search_fn="backtracking")
    Cg, Yg, info = st.optimize(_t(X), _t(Y), _t(ops), _t(ch), _t(C0), max_iter=50, line_search_fn="backtracking",
                               return_info=True, raise_errors=False)
    msg = helpers.compare_fit_to_oracle(_mse(Yg, Y), fits, info["err"].cpu().numpy(), f"backtracking seed {seed}")
    print(msg, flush=True)
    # non-constant slots keep C_init, or turn NaN where the reference's do
    # (x += s * (~done) with a NaN step, symbolic_torch/lbfgs_batch.py:307)
    nonc = ops != 1
    Cn = Cg.cpu().numpy()
    assert ((Cn == C0) | np.isnan(Cn))[nonc].all()


LENGTHS = [(1, 8), (9, 16), (17, 32), (33, 64), (65, 127)]


@pytest.fixture(scope="module")
def bucket_batch():
    X, Y, ops, ch, C0 = _problems(7, 32, 128, 2, 64, LENGTHS, unary=[10, 11], binary=[40, 41, 42],
                                  leaf_const_p=0.25)
    L = (ops != 0).sum(1)
    counts = [int(((L >= lo) & (L <= hi)).sum()) for lo, hi in LENGTHS]
    print(f"problems per length range {LENGTHS}: {counts}", flush=True)
    assert min(counts) >= 8
    fits = helpers.oracle_fits(X, Y, ops, ch, C0, max_iter=30)
    return X, Y, ops, ch, C0, fits


def test_every_fit_variant_matches_oracle(bucket_batch):
    from cudasymbolicregression_amd import autotune
    from cudasymbolicregression_amd.optimization import _fit

    X, Y, ops, ch, C0, fits = bucket_batch
    Xd, Yd, Od, Cd, C0d = _t(X), _t(Y), _t(ops), _t(ch), _t(C0)
    base = autotune.default_config()
    var = autotune.variants()
    print(f"defaults {base}; variants per bucket {[len(v) for v in var]}", flush=True)
    ran = 0
    for b, row in enumerate(var):
        for v, info in enumerate(row):
            cfg = list(base)
            cfg[b] = v
            print(f"case variant bucket={b} variant={v} {info} config={cfg}", flush=True)
            C, loss, iters, err, flags, ncl = _fit(Xd, Yd, Od, Cd, C0d, 30, 10, 1e-7, "strong_wolfe", config=cfg)
            torch.cuda.synchronize()
            yg = lbfgs_oracle.oracle.fused_root(X, ops, ch, C.cpu().numpy(), Y, mode=1, want_gx=False)["y"]
            with np.errstate(all="ignore"):
                mse = ((yg.astype(np.float64) - Y) ** 2).mean(1)
            msg = helpers.compare_fit_to_oracle(mse, fits, err.cpu().numpy(), f"bucket {b} variant {v}",
                                                max_undetermined=0.25)
            print(msg, flush=True)
            ran += 1
    assert ran == sum(len(r) for r in var)


def test_tune_fit_config_is_stored_and_used(bucket_batch):
    import cudasymbolicregression_amd as st
    from cudasymbolicregression_amd import autotune

    X, Y, ops, ch, C0, fits = bucket_batch
    Xd, Yd, Od, Cd, C0d = _t(X), _t(Y), _t(ops), _t(ch), _t(C0)
    autotune.cache_clear()
    try:
        log = []
        res = autotune.tune_fit(Xd, Yd, Od, Cd, C0d, max_iter=30, reps=2, log=log)
        print("\n".join(log), flush=True)
        print(f"chosen {res['config']} ({res['ms']:.3f} ms) for {res['key']}", flush=True)
        n_trials = 1 + sum(len(r) - 1 for r in autotune.variants())
        assert len(log) == len(res["trials"]) == n_trials
        assert autotune.lookup(Xd.device, ops.shape[1], X.shape[0], X.shape[1]) == res["config"]
        Cg, Yg, info = st.optimize(Xd, Yd, Od, Cd, C0d, max_iter=30, return_info=True, raise_errors=False)
        msg = helpers.compare_fit_to_oracle(_mse(Yg, Y), fits, info["err"].cpu().numpy(), "tuned configuration",
                                            max_undetermined=0.25)
        print(msg, flush=True)
    finally:
        autotune.cache_clear()
