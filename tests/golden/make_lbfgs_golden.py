"""Writes tests/golden/lbfgs_trajectories.npz: golden L-BFGS trajectories of
the reference's LBFGS_batch (symbolic_torch/lbfgs_batch.py:7-326).

Run in the build container only (it reads the reference's source from
/root/reference, which the GPU box does not have):

    python tests/golden/make_lbfgs_golden.py

The reference's optimizer file is loaded by path (its package __init__
needs the compiled extension, which is not built here) with bytecode
writing off, so nothing is written under /root/reference.  Every case runs
LBFGS_batch.step on a closure written here in torch (float32, CPU) and
records each closure call: the point x it was called at (D, B), the
returned loss f (B,) and gradient g (D, B), in call order, plus the final
point and the returned loss.  The sequence of evaluated points is the whole
trajectory: the trial step lengths of the line searches, the zoom's
bisections and its fallbacks, the accepted steps and the history updates all
show in it.  tests/test_lbfgs_golden.py replays the recorded f and g to the
oracle's restatement (oracle/lbfgs_oracle.py), which must ask for exactly the
same points, bit for bit.

Closures (all problems independent, parameters of shape (D, B)):
  quad1d   f = w (x - 10)^2, one variable, w per problem chosen so that the
           first strong-Wolfe search accepts at alpha 1, at 16.2 before or
           after the minimum, zooms, or runs out of trials;
  rosen    2-D Rosenbrock from randn starts (the reference's own demo);
  illquad  4-D quadratic with curvatures 1e-2 .. 1e2 (zooms, history wrap);
  sinfit   optimize-style: 6 slots, constants in slots 1 and 3,
           y = c1 * sin(c3 * x) fitted to 24 points, gradient zeroed off
           the constant slots (symbolic_torch/optimization.py:95-97);
  expfit   y = exp(c0 * x) * c1 on x in [0, 30]: overflow to inf and NaN
           losses, NaN steps, and problems that finish early beside them.
Every batched dot product has at most 4 non-zero terms per problem: torch's
float32 column sums of more rows take an order that depends on the column
count and the CPU's vector width (an implementation detail of torch, not
of the algorithm); up to 4 terms every path adds left to right.
"""
from __future__ import annotations

import importlib.util
import os
import sys

import numpy as np
import torch

REF = "/root/reference/symbolic_torch/lbfgs_batch.py"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lbfgs_trajectories.npz")


def load_reference():
    sys.dont_write_bytecode = True
    spec = importlib.util.spec_from_file_location("reference_lbfgs_batch", REF)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.LBFGS_batch


# ------------------------------------------------------------- closures
def quad1d(x):
    w = torch.tensor([1.0, 0.03, 0.04, 0.002, 0.6], dtype=torch.float32)[None, :]
    return (w * (x - 10.0) ** 2).sum(0)


def rosen(x):
    return (1.0 - x[0]) ** 2 + 100.0 * (x[1] - x[0] ** 2) ** 2


_IQ_W = torch.logspace(-2, 2, 4, dtype=torch.float32)[:, None] * torch.tensor(
    [1.0, 0.7, 1.3, 0.5, 2.0], dtype=torch.float32)[None, :]
_IQ_T = torch.arange(4, dtype=torch.float32)[:, None] - 2.0


def illquad(x):
    return (_IQ_W * (x - _IQ_T) ** 2).sum(0)


_SIN_X = torch.linspace(-2.0, 2.0, 24, dtype=torch.float32)
_SIN_Y = torch.stack([a * torch.sin(b * _SIN_X) for a, b in
                      [(1.5, 0.8), (-2.0, 1.3), (0.5, 2.5), (3.0, -0.6), (1.0, 4.0)]])  # (B, N)
_SIN_MASK = torch.zeros(6, 1, dtype=torch.bool)
_SIN_MASK[1] = True
_SIN_MASK[3] = True


def sinfit(x):
    y = x[1][:, None] * torch.sin(x[3][:, None] * _SIN_X[None, :])
    return ((y - _SIN_Y) ** 2).mean(1)


_EXP_X = torch.linspace(0.0, 30.0, 16, dtype=torch.float32)
_EXP_Y = torch.stack([c1 * torch.exp(c0 * _EXP_X) for c0, c1 in
                      [(0.05, 1.0), (-0.2, 2.0), (0.5, 0.1), (1.5, 1.0), (0.01, -1.0), (2.0, 3.0)]])


def expfit(x):
    y = torch.exp(x[0][:, None] * _EXP_X[None, :]) * x[1][:, None]
    return ((y - _EXP_Y) ** 2).mean(1)


def cases():
    g = torch.Generator().manual_seed(0)
    out = []
    out.append(dict(name="quad1d_sw", fn=quad1d, mask=None, x0=torch.zeros(1, 5), max_iter=3, hist=10,
                    ls="strong_wolfe", ls_max=10))
    out.append(dict(name="quad1d_bt", fn=quad1d, mask=None, x0=torch.zeros(1, 5), max_iter=3, hist=10,
                    ls="backtracking", ls_max=10))
    x0 = torch.randn(2, 6, generator=g)
    out.append(dict(name="rosen_sw", fn=rosen, mask=None, x0=x0, max_iter=60, hist=5, ls="strong_wolfe", ls_max=10))
    out.append(dict(name="rosen_bt", fn=rosen, mask=None, x0=x0, max_iter=60, hist=5, ls="backtracking", ls_max=10))
    x0 = torch.randn(4, 5, generator=g) * 3
    out.append(dict(name="illquad_sw", fn=illquad, mask=None, x0=x0, max_iter=25, hist=3, ls="strong_wolfe",
                    ls_max=10))
    out.append(dict(name="illquad_bt", fn=illquad, mask=None, x0=x0, max_iter=25, hist=3, ls="backtracking",
                    ls_max=6))
    x0 = torch.randn(6, 5, generator=g)
    out.append(dict(name="sinfit_sw", fn=sinfit, mask=_SIN_MASK, x0=x0, max_iter=40, hist=10,
                    ls="strong_wolfe", ls_max=10))
    out.append(dict(name="sinfit_bt", fn=sinfit, mask=_SIN_MASK, x0=x0, max_iter=40, hist=10,
                    ls="backtracking", ls_max=10))
    x0 = torch.tensor([[0.1, -0.1, 0.4, 1.0, 0.0, 0.5], [1.0, 1.0, 1.0, 1.0, 1.0, 1.0]])
    out.append(dict(name="expfit_sw", fn=expfit, mask=None, x0=x0, max_iter=30, hist=10, ls="strong_wolfe",
                    ls_max=10))
    out.append(dict(name="expfit_bt", fn=expfit, mask=None, x0=x0, max_iter=30, hist=10, ls="backtracking",
                    ls_max=10))
    return out


def run_case(LBFGS_batch, c):
    x = c["x0"].clone().float().requires_grad_(True)
    opt = LBFGS_batch([x], max_iter=c["max_iter"], history_size=c["hist"], tolerance_grad=1e-7,
                      line_search_fn=c["ls"], line_search_max_iter=c["ls_max"])
    xs, fs, gs = [], [], []

    def closure():
        opt.zero_grad(set_to_none=True)
        xs.append(x.detach().clone())
        f = c["fn"](x)
        f.sum().backward()
        if c["mask"] is not None:
            x.grad[~c["mask"].expand_as(x.grad)] = 0.0
        fs.append(f.detach().clone())
        gs.append(x.grad.detach().clone())
        return f

    with np.errstate(all="ignore"):
        f_ret = opt.step(closure)
    return (torch.stack(xs).numpy(), torch.stack(fs).numpy(), torch.stack(gs).numpy(), x.detach().numpy().copy(),
            f_ret.detach().numpy().copy())


def main():
    torch.set_num_threads(1)
    LBFGS_batch = load_reference()
    arrays = {}
    names = []
    for c in cases():
        xs, fs, gs, xf, fret = run_case(LBFGS_batch, c)
        n = c["name"]
        names.append(n)
        arrays[f"{n}/x0"] = c["x0"].float().numpy()
        arrays[f"{n}/xs"] = xs
        arrays[f"{n}/fs"] = fs
        arrays[f"{n}/gs"] = gs
        arrays[f"{n}/x_final"] = xf
        arrays[f"{n}/f_return"] = fret
        arrays[f"{n}/opts"] = np.array([c["max_iter"], c["hist"], 0 if c["ls"] == "strong_wolfe" else 1, c["ls_max"]],
                                       np.int64)
        print(f"{n}: {len(xs)} closure calls, non-finite losses {int((~np.isfinite(fs)).sum())}", flush=True)
    arrays["names"] = np.array(names)
    np.savez_compressed(OUT, **arrays)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)")


if __name__ == "__main__":
    main()
