"""`symbolic_torch.utils` (reference symbolic_torch/utils.py)."""
from cudasymbolicregression_amd.utils import create_constants  # noqa: F401
