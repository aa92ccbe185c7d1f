"""Batch-first evaluation wrappers (reference symbolic_torch/evaluation.py:4-65).

`evaluate` permutes the batch-first (B, M) inputs to the node-first (M, B)
layout of the backend and returns the backend's (M, N, B) cache as a
(B, M, N) view, exactly like the reference wrapper.
"""
from __future__ import annotations

import torch

from ._C import evaluate_backend, evaluate_multiple_constant_backend

__all__ = ["evaluate", "evaluate_multiple", "evaluate_backend", "evaluate_multiple_constant_backend"]


def evaluate(
    X: torch.Tensor,  # (N, n_x)
    Ops: torch.Tensor,  # (B, M)
    Ch: torch.Tensor,  # (B, M, MAX_ARITY)
    C: torch.Tensor,  # (B, M)
) -> torch.Tensor:
    """Evaluates a batch of postfix expressions at every point of X.

    Returns the value of every node at every point, shape (B, M, N); it is a
    view of an (M, N, B) tensor.  Differentiable with respect to X and C.
    """
    result = evaluate_backend(X, Ops.permute(1, 0), Ch.permute(1, 0, 2), C.permute(1, 0))
    return result.permute(2, 0, 1)


def evaluate_multiple(
    X: torch.Tensor,  # (N, n_x)
    Ops: torch.Tensor,  # (B, M)
    Ch: torch.Tensor,  # (B, M, 2)
    C: torch.Tensor,  # (B, M, K)
) -> torch.Tensor:
    """Evaluates every expression for K constant sets; returns (B, M, N, K)."""
    result = evaluate_multiple_constant_backend(
        X, Ops.permute(1, 0).contiguous(), Ch.permute(1, 0, 2).contiguous(), C.permute(1, 0, 2).contiguous()
    )
    return result.permute(2, 0, 1, 3)
