"""`symbolic_torch.evaluation` (reference symbolic_torch/evaluation.py)."""
from cudasymbolicregression_amd.evaluation import (evaluate, evaluate_backend, evaluate_multiple,  # noqa: F401
                                                   evaluate_multiple_constant_backend)
