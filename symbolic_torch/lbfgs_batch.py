"""`symbolic_torch.lbfgs_batch` (reference symbolic_torch/lbfgs_batch.py)."""
from cudasymbolicregression_amd.lbfgs_batch import LBFGS_batch  # noqa: F401
