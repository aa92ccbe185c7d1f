"""`symbolic_torch`, the reference package's import path, served by this
implementation (cudasymbolicregression_amd).

Code written for the reference (symbolic_torch/__init__.py:4-16) imports its
modules by these names: `symbolic_torch._C`, `.evaluation`, `.optimization`,
`.lbfgs_batch`, `.utils`.  Each is a module object of its own here that
re-exports the names of the matching cudasymbolicregression_amd module, so
`from symbolic_torch._C import evaluate_multiple_constant_backend` and
`from symbolic_torch.lbfgs_batch import LBFGS_batch` resolve to the gfx950
implementation.
"""
from cudasymbolicregression_amd import (Operator, ProbabilisticContextFreeGrammar, create_constants, get_arity,
                                        optimize, optimize_ragged)

from . import _C, evaluation, lbfgs_batch, optimization, utils  # noqa: E402

__all__ = ["evaluation", "Operator", "ProbabilisticContextFreeGrammar", "create_constants", "optimize",
           "optimize_ragged", "get_arity"]
