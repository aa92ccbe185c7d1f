"""`symbolic_torch._C` (reference pybind11 module, src/main.cpp:8-13): the
names of cudasymbolicregression_amd._C."""
from cudasymbolicregression_amd._C import *  # noqa: F401,F403
from cudasymbolicregression_amd._C import (MAX_ARITY, NULL_CHILD, Operator, ProbabilisticContextFreeGrammar,  # noqa: F401
                                           evaluate_backend, evaluate_multiple_constant_backend, get_arity,
                                           get_error_message, is_valid_op)
