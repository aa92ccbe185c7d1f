"""create_constants (reference symbolic_torch/utils.py:7-42).

Makes the same torch RNG calls in the same order as the reference, so a given
torch seed yields bit-identical constants.
"""
from __future__ import annotations

from typing import Any

import torch

from ._C import Operator


def create_constants(ops_tensor: torch.Tensor, method: str = "randn", **kwargs: Any) -> torch.Tensor:
    """Float32 tensor shaped like `ops_tensor`: zeros, with fresh constants at
    the LEARNABLE_CONSTANT positions (row-major mask order).

    method 'randn': N(mean, std) with kwargs mean (0.0) and std (1.0).
    method 'range': U(lower, upper) with kwargs lower (0.0) and upper (1.0).
    """
    const_mask = ops_tensor == int(Operator.LEARNABLE_CONSTANT)
    num_constants = int(const_mask.sum().item())
    device = ops_tensor.device
    dtype = torch.float32
    if method == "randn":
        mean = kwargs.get("mean", 0.0)
        std = kwargs.get("std", 1.0)
        constants_flat = torch.randn(num_constants, dtype=dtype, device=device) * std + mean
    elif method == "range":
        lower = kwargs.get("lower", 0.0)
        upper = kwargs.get("upper", 1.0)
        constants_flat = torch.rand(num_constants, dtype=dtype, device=device) * (upper - lower) + lower
    else:
        raise ValueError(f"Unknown method for constant creation: {method}")
    output = torch.zeros_like(ops_tensor, dtype=dtype, device=device)
    if num_constants > 0:
        output[const_mask] = constants_flat
    return output
