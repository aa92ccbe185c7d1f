"""MI355X-native implementation of the `symbolic_torch` API.

Drop-in surface of JobPetrovcic/CUDASymbolicRegression's package
(reference symbolic_torch/__init__.py:4-16): `Operator`,
`ProbabilisticContextFreeGrammar`, `get_arity`, `create_constants`,
`evaluation` (evaluate / evaluate_multiple), `optimize`, plus
`lbfgs_batch.LBFGS_batch`.  Expression evaluation and constant fitting run in
hand-written gfx950 HIP kernels (libsr_hip.so); grammar sampling and parsing
run in a host C++ engine (libsr_host.so).
"""
import torch  # noqa: F401

from ._C import Operator, ProbabilisticContextFreeGrammar, get_arity
from .utils import create_constants
from . import evaluation
from .optimization import optimize, optimize_multistart, optimize_ragged

__all__ = [
    "evaluation",
    "Operator",
    "ProbabilisticContextFreeGrammar",
    "create_constants",
    "optimize",
    "optimize_multistart",
    "optimize_ragged",
    "get_arity",
]
