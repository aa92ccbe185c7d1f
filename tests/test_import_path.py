"""The reference's own import lines resolve to this implementation
(tests/test_evaluation_multiple_constants.py:5-9, tests/test_lbfgs_batch.py:10,
tests/test_symbolic_evaluation.py:8-12, tests/test_optimization.py:6-7,
tests/test_utils.py:2, symbolic_torch/__init__.py:4-16)."""


def test_reference_import_lines_resolve():
    import symbolic_torch
    from symbolic_torch import Operator, ProbabilisticContextFreeGrammar, create_constants, get_arity, optimize
    from symbolic_torch._C import evaluate_multiple_constant_backend
    from symbolic_torch.evaluation import evaluate, evaluate_backend
    from symbolic_torch.lbfgs_batch import LBFGS_batch

    import cudasymbolicregression_amd as impl
    from cudasymbolicregression_amd import _C, evaluation, lbfgs_batch

    assert Operator is impl.Operator and get_arity is impl.get_arity
    assert ProbabilisticContextFreeGrammar is impl.ProbabilisticContextFreeGrammar
    assert create_constants is impl.create_constants and optimize is impl.optimize
    assert evaluate_multiple_constant_backend is _C.evaluate_multiple_constant_backend
    assert evaluate is evaluation.evaluate and evaluate_backend is evaluation.evaluate_backend
    assert LBFGS_batch is lbfgs_batch.LBFGS_batch
    assert symbolic_torch.evaluation.evaluate_multiple is evaluation.evaluate_multiple
    assert int(symbolic_torch._C.ADD) == 40 and symbolic_torch._C.is_valid_op(50, 1)
