"""Pins of the oracle's L-BFGS restatement (oracle/lbfgs_oracle.py) against
the reference's LBFGS_batch (symbolic_torch/lbfgs_batch.py:61-326):

* golden trajectories (tests/golden/lbfgs_trajectories.npz, written by
  tests/golden/make_lbfgs_golden.py from the reference's own optimizer): the
  oracle is handed the reference's recorded losses and gradients, one
  closure call at a time, and must ask for the same points bit for bit and
  end at the same parameters.  A mistake in the direction, the line
  searches, the zoom, the step or the history shows up as the first point
  that differs;
* hand-computed first strong-Wolfe searches on f = w (x - 10)^2 from x = 0,
  where p = -g = 20 w and phi'(alpha) / phi'(0) = 1 - 2 w alpha (from the
  rule of lbfgs_batch.py:138-202: c1 = 1e-4, c2 = 0.9, alpha0 = 1,
  alpha <- alpha + 0.8 (20 - alpha), curvature tested before the bracket):
    w = 1     alpha 1 overshoots to phi(1) = phi(0) (sufficient decrease
              fails): zoom on [0, 1], which recomputes phi(0) and phi(lo),
              then accepts the midpoint 0.5, the exact minimum;
    w = 0.03  alpha 1 gives |phi'| = 0.94 |phi'(0)| (not accepted, no
              bracket), alpha 16.2 gives 0.028 |phi'(0)|: accepted;
    w = 0.04  alpha 16.2 is past the minimum (phi' > 0) but |phi'| =
              0.296 |phi'(0)| <= 0.9: accepted without a zoom, because the
              curvature test comes first;
    w = 0.002 no trial is accepted in 10: the search returns the step
              after the last trial, which was never evaluated;
    w = 0.6   alpha 1 already satisfies the curvature test (|1 - 1.2| =
              0.2): accepted at once.
"""
import os

import numpy as np
import pytest

import lbfgs_oracle

F32 = np.float32
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "lbfgs_trajectories.npz")


def _bit_diff(a, b):
    """True where a and b differ bit for bit; any two NaNs are equal (their
    sign and payload bits depend on the library that made them)."""
    a, b = np.asarray(a, F32), np.asarray(b, F32)
    both_nan = np.isnan(a) & np.isnan(b)
    return (a.view(np.int32) != b.view(np.int32)) & ~both_nan


class Replay:
    """A closure that returns the reference's recorded (f, g) for its k-th
    call after checking that it is asked for the reference's k-th point."""

    def __init__(self, xs, fs, gs, name):
        self.xs, self.fs, self.gs, self.name = xs, fs, gs, name
        self.k = 0

    def __call__(self, x):
        k = self.k
        assert k < len(self.xs), f"{self.name}: more closure calls than the reference's {len(self.xs)}"
        want = self.xs[k]
        diff = _bit_diff(np.asarray(x, F32), want)
        if diff.any():
            bad = np.argwhere(diff)
            d, b = bad[0]
            raise AssertionError(f"{self.name}: call {k} asks for x[{d},{b}] = {x[d, b]!r}, "
                                 f"the reference evaluated {want[d, b]!r} ({len(bad)} values differ)")
        self.k += 1
        return self.fs[k].copy(), self.gs[k].copy()


def _golden():
    return np.load(GOLDEN, allow_pickle=False)


def _names():
    return [str(n) for n in _golden()["names"]]


@pytest.mark.parametrize("name", _names())
def test_oracle_reproduces_reference_trajectory(name):
    z = _golden()
    xs, fs, gs = z[f"{name}/xs"], z[f"{name}/fs"], z[f"{name}/gs"]
    max_iter, hist, ls, ls_max = (int(v) for v in z[f"{name}/opts"])
    cl = Replay(xs, fs, gs, name)
    x, f = lbfgs_oracle.lbfgs_run(cl, z[f"{name}/x0"], max_iter=max_iter, history_size=hist,
                                  line_search_fn="strong_wolfe" if ls == 0 else "backtracking",
                                  line_search_max_iter=ls_max)
    assert cl.k == len(xs), f"{name}: {cl.k} closure calls, the reference made {len(xs)}"
    assert not _bit_diff(x, z[f"{name}/x_final"]).any()
    assert not _bit_diff(f, z[f"{name}/f_return"]).any()


def test_golden_cases_cover_the_quirks():
    """The fixture exercises what it is meant to: zooms, exhausted
    searches, non-finite losses and problems finishing at different
    iterations (otherwise the trajectory test would pin less)."""
    z = _golden()
    fs = z["expfit_sw/fs"]
    assert (~np.isfinite(fs)).any() and np.isfinite(fs).any()
    # rosen: some problems finish (gradient below tolerance) while others run on
    g = z["rosen_sw/gs"][-1]
    gn = np.sqrt((g.astype(np.float64) ** 2).sum(0))
    assert (gn < 1e-7).any() or np.isfinite(gn).all()
    assert len(z["quad1d_sw/xs"]) > 10


def _quad(w):
    w = np.asarray(w, F32)

    def cl(x):
        d = (x - F32(10)).astype(F32)
        return (w * d * d).sum(0).astype(F32), (F32(2) * w * d).astype(F32)

    cl.calls = []

    def rec(x):
        cl.calls.append(np.asarray(x, F32).copy())
        return cl(x)

    rec.calls = cl.calls
    return rec


def _next(a):
    """The bracketing update alpha + (20 - alpha) * 0.8 in float32."""
    return F32(a + (F32(20) - a) * F32(0.8))


A1 = F32(1)
A2 = _next(A1)  # 16.2 in float32


@pytest.mark.parametrize("w,trials,alpha", [
    (1.0, [A1, F32(0), F32(0), F32(0.5)], F32(0.5)),  # zoom: phi0 and phi(lo) recomputed, then the midpoint
    (0.03, [A1, A2], A2),
    (0.04, [A1, A2], A2),  # past the minimum but accepted: curvature first
    (0.6, [A1], A1),
])
def test_first_strong_wolfe_search_by_hand(w, trials, alpha):
    cl = _quad([w])
    x0 = np.zeros((1, 1), F32)
    x, _ = lbfgs_oracle.lbfgs_run(cl, x0, max_iter=1)
    p = F32(20) * F32(w)
    # calls: f(x0) of step, phi0 of the search, the trials (zoom: phi0, phi(lo), midpoints), f(x1)
    got = [float(c[0, 0]) for c in cl.calls]
    want = [0.0, 0.0] + [float(F32(a * p)) for a in trials] + [float(F32(alpha * p))]
    assert got == want, (got, want)
    assert x[0, 0] == F32(alpha * p)


def test_exhausted_search_returns_the_untried_step():
    """w = 0.002: phi'(alpha)/phi'(0) = 1 - 0.004 alpha stays above 0.9 for
    every trial alpha <= 20, so none is accepted and none brackets; the
    search returns alpha after one more update, never evaluated."""
    cl = _quad([0.002])
    x, _ = lbfgs_oracle.lbfgs_run(cl, np.zeros((1, 1), F32), max_iter=1)
    seq = [A1]
    for _ in range(9):
        seq.append(_next(seq[-1]))
    untried = _next(seq[-1])
    p = F32(20) * F32(0.002)
    got = [float(c[0, 0]) for c in cl.calls[2:-1]]
    assert got == [float(F32(t * p)) for t in seq]
    assert x[0, 0] == F32(untried * p)  # the 11th step length
    assert untried not in seq


def test_backtracking_first_step_by_hand():
    """Armijo backtracking (lbfgs_batch.py:105-136) on w = 1: alpha 1 gives
    phi(1) = phi(0) > phi(0) + 1e-4 * phi'(0), alpha 0.5 lands on the
    minimum: two trials."""
    cl = _quad([1.0])
    x, _ = lbfgs_oracle.lbfgs_run(cl, np.zeros((1, 1), F32), max_iter=1, line_search_fn="backtracking")
    got = [float(c[0, 0]) for c in cl.calls]
    assert got == [0.0, 20.0, 10.0, 10.0]
    assert x[0, 0] == F32(10)


def _bracket_first_strong_wolfe(cl, x, p, max_ls, c1=1e-4, c2=0.9):
    """A plausible mistake: the bracket test (sufficient decrease failed or
    the minimum passed) before the curvature test."""
    phi0, g0 = cl(x)
    dphi0 = lbfgs_oracle._bdot(g0, p)[0]
    a_prev = np.zeros_like(phi0)
    phi_prev = phi0.copy()
    a = np.ones_like(phi0)
    final = np.zeros_like(phi0)
    done = np.zeros(phi0.shape, bool)
    for i in range(max_ls):
        if np.all(done):
            break
        phi, ga = cl(x + a[None, :] * p)
        dphi = lbfgs_oracle._bdot(ga, p)[0]
        bad = (phi > phi0 + F32(c1) * a * dphi0) | ((i > 0) & (phi >= phi_prev))
        good = np.abs(dphi) <= F32(c2) * np.abs(dphi0)
        nz = ~done & (bad | (dphi >= 0))
        if np.any(nz):
            lo = np.where(bad, a_prev, a)
            hi = np.where(bad, a, a_prev)
            final = np.where(nz, lbfgs_oracle._zoom(cl, x, p, lo, hi, nz, c1, c2), final)
            done |= nz
        ng = ~done & good
        final = np.where(ng, a, final)
        done |= ng
        cont = ~done
        a_prev = np.where(cont, a, a_prev)
        phi_prev = np.where(cont, phi, phi_prev)
        a = np.where(cont, (a + (F32(20) - a) * F32(0.8)).astype(F32), a)
    return np.where(done, final, a).astype(F32)[None, :]


def _reversed_history(g, S, Yh, eps):
    """A plausible mistake: the history pairs taken in reverse order (oldest
    first in the second loop, gamma from the oldest pair)."""
    if len(S) > 1:
        S, Yh = S[::-1], Yh[::-1]
    return _ORIG_DIRECTION(g, S, Yh, eps)


_ORIG_DIRECTION = lbfgs_oracle._direction


@pytest.mark.parametrize("mutant,case", [("bracket_first", "quad1d_sw"), ("bracket_first", "rosen_sw"),
                                         ("reversed_history", "rosen_sw"), ("reversed_history", "illquad_bt")])
def test_golden_replay_catches_planted_mistakes(monkeypatch, mutant, case):
    if mutant == "bracket_first":
        monkeypatch.setattr(lbfgs_oracle, "_strong_wolfe", _bracket_first_strong_wolfe)
    else:
        monkeypatch.setattr(lbfgs_oracle, "_direction", _reversed_history)
    with pytest.raises(AssertionError):
        test_oracle_reproduces_reference_trajectory(case)
