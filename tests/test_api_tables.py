"""The operator tables and the evaluate_backend input checks (CPU).

* get_arity, is_valid_op, the Operator enum values, and the PCFG's
  get_arities() / valid_ops() / n_operators over every id 0 .. 54 with five
  variables, gaps included (reference tests/test_pcfg_arities_validity.py:
  32-238, tests/test_utils.py:6-61, include/operators.h:10-125).
* evaluate_backend raises the reference's RuntimeError texts for wrong
  dtypes, shapes and devices before any device work
  (src/symbolic_evaluation.cpp:306-334, tests/test_symbolic_evaluation.py:
  327-340); these checks run on CPU tensors.
"""
import pytest
import torch

import cudasymbolicregression_amd as st
from cudasymbolicregression_amd import Operator, ProbabilisticContextFreeGrammar, _C, get_arity
from cudasymbolicregression_amd.evaluation import evaluate_backend

# expected tables, id -> value, written out as the reference's tests list them
ARITY = {**{i: 0 for i in range(0, 10)}, **{i: 1 for i in range(10, 32)}, **{i: 0 for i in range(32, 40)},
         **{i: 2 for i in range(40, 45)}, **{i: 0 for i in range(45, 55)}}
VALID = {0: False, **{i: True for i in range(1, 9)}, 9: False, **{i: True for i in range(10, 32)},
         **{i: False for i in range(32, 40)}, **{i: True for i in range(40, 45)},
         **{i: False for i in range(45, 50)}, **{i: True for i in range(50, 55)}}
ENUM = dict(NO_OP=0, LEARNABLE_CONSTANT=1, CONST_1=2, CONST_2=3, CONST_3=4, CONST_4=5, CONST_5=6, PI=7, E=8,
            SIN=10, COS=11, EXP=12, LOG=13, SQUARE=14, SQRT=15, TAN=16, ARCSIN=17, ARCCOS=18, ARCTAN=19,
            SINH=20, COSH=21, TANH=22, FLOOR=23, CEIL=24, LN=25, LOG10=26, NEG=27, INV=28, CUBE=29,
            FOURTH=30, FIFTH=31, ADD=40, SUB=41, MUL=42, DIV=43, POW=44, VAR_START_ID=50)


@pytest.fixture(scope="module")
def pcfg():
    return ProbabilisticContextFreeGrammar("S -> 1 [1.0]", "S", 32, 5, torch.device("cpu"))


def test_enum_values_and_module_attributes():
    for name, v in ENUM.items():
        assert int(getattr(Operator, name)) == v
        assert int(getattr(_C, name)) == v  # pybind11 export_values()
    assert {o.name for o in Operator} == set(ENUM)


def test_get_arity_table():
    for op, a in ARITY.items():
        assert get_arity(op) == a, op
    assert get_arity(int(Operator.VAR_START_ID) + 10) == 0


def test_is_valid_op_table():
    for op, v in VALID.items():
        assert _C.is_valid_op(op, 5) == v, op
    assert not _C.is_valid_op(55, 5) and _C.is_valid_op(55, 6)


def test_pcfg_arities_and_validity(pcfg):
    ar, va = pcfg.get_arities(), pcfg.valid_ops()
    assert pcfg.n_operators == 55 == len(ar) == len(va)
    for i in range(55):
        assert int(ar[i]) == ARITY[i], i
        assert bool(va[i]) == VALID[i], i


# ---------------------------------------------------------------- input checks
def _ok_inputs():
    X = torch.randn(10, 2, dtype=torch.float32)
    ops = torch.tensor([[int(Operator.VAR_START_ID)]], dtype=torch.int64)
    ch = -torch.ones(1, 1, 2, dtype=torch.int64)
    C = torch.zeros_like(ops, dtype=torch.float32)
    return X, ops, ch, C


def test_wrong_datatypes_raise_reference_texts():
    """tests/test_symbolic_evaluation.py:327-340."""
    X, ops, ch, C = _ok_inputs()
    with pytest.raises(RuntimeError, match="X must be a floating-point tensor"):
        evaluate_backend(X.to(torch.int32), ops, ch, C)
    with pytest.raises(RuntimeError, match="must be a Long tensor"):
        evaluate_backend(X, ops.to(torch.float32), ch, C)
    with pytest.raises(RuntimeError, match="must be a Long tensor"):
        evaluate_backend(X, ops, ch.to(torch.float32), C)
    with pytest.raises(RuntimeError, match="C must be a floating-point tensor"):
        evaluate_backend(X, ops, ch, C.to(torch.int32))


def test_shape_checks_raise_reference_texts():
    """src/symbolic_evaluation.cpp:314-334, in the reference's order."""
    X, ops, ch, C = _ok_inputs()
    with pytest.raises(RuntimeError, match="Ops must be a 2D tensor"):
        evaluate_backend(X, ops[0], ch, C)
    with pytest.raises(RuntimeError, match="Ch must be a 3D tensor"):
        evaluate_backend(X, ops, ch[0], C)
    with pytest.raises(RuntimeError, match="C must be a 2D tensor"):
        evaluate_backend(X, ops, ch, C[0])
    with pytest.raises(RuntimeError, match=r"Ops and Ch must have the same max length \(dim 0\)"):
        evaluate_backend(X, ops, torch.cat([ch, ch]), C)
    with pytest.raises(RuntimeError, match=r"Ops and Ch must have the same batch size \(dim 1\)"):
        evaluate_backend(X, ops, torch.cat([ch, ch], 1), C)
    with pytest.raises(RuntimeError, match="Ops and C must have the same shape"):
        evaluate_backend(X, ops, ch, torch.cat([C, C]))
    with pytest.raises(RuntimeError, match="Ch must have MaxArity as its last dimension"):
        evaluate_backend(X, ops, torch.cat([ch, ch], 2), C)


def test_device_mismatch_raises_reference_text():
    X, ops, ch, C = _ok_inputs()
    with pytest.raises(RuntimeError, match="All input tensors must be on the same device."):
        evaluate_backend(X, ops.to("meta"), ch, C)


def test_cpu_tensors_are_refused_after_the_checks():
    """Valid CPU inputs pass the reference's checks and are then refused: this
    implementation evaluates only with its gfx950 kernels (no CPU fallback)."""
    X, ops, ch, C = _ok_inputs()
    with pytest.raises(RuntimeError, match="ROCm GPU device"):
        evaluate_backend(X, ops, ch, C)


def test_create_constants_errors_and_layout():
    """symbolic_torch/utils.py:7-42: constants only in the constant slots, in
    row-major order, and a ValueError for an unknown method."""
    ops = torch.tensor([[1, 50, 42, 1, 40], [50, 1, 40, 0, 0]])
    torch.manual_seed(0)
    C = st.create_constants(ops, method="randn")
    torch.manual_seed(0)
    ref = torch.randn(3)
    assert C.dtype == torch.float32 and C.shape == ops.shape
    assert torch.equal(C[ops == 1], ref) and (C[ops != 1] == 0).all()
    with pytest.raises(ValueError):
        st.create_constants(ops, method="nope")
