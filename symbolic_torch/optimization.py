"""`symbolic_torch.optimization` (reference symbolic_torch/optimization.py)."""
from cudasymbolicregression_amd.optimization import optimize, optimize_ragged  # noqa: F401
