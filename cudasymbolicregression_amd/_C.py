"""Python mirror of the reference's compiled ``symbolic_torch._C`` module.

Same names, argument meaning, error types and message texts as the pybind11
module bound in the reference's src/main.cpp:8-13, src/utils.cpp:3-9,
src/symbolic_evaluation.cpp:357-401 and src/pcfg.cpp:1006-1061.  The compute
goes through the C-ABI of libsr_hip.so (gfx950 kernels) and libsr_host.so
(host PCFG engine); there is no CPU evaluation path.
"""
from __future__ import annotations

import enum

import torch

from . import _lib


class Operator(enum.IntEnum):
    """Opcode table (reference include/operators.h:10-59)."""

    NO_OP = 0
    LEARNABLE_CONSTANT = 1
    CONST_1 = 2
    CONST_2 = 3
    CONST_3 = 4
    CONST_4 = 5
    CONST_5 = 6
    PI = 7
    E = 8
    SIN = 10
    COS = 11
    EXP = 12
    LOG = 13
    SQUARE = 14
    SQRT = 15
    TAN = 16
    ARCSIN = 17
    ARCCOS = 18
    ARCTAN = 19
    SINH = 20
    COSH = 21
    TANH = 22
    FLOOR = 23
    CEIL = 24
    LN = 25
    LOG10 = 26
    NEG = 27
    INV = 28
    CUBE = 29
    FOURTH = 30
    FIFTH = 31
    ADD = 40
    SUB = 41
    MUL = 42
    DIV = 43
    POW = 44
    VAR_START_ID = 50


# pybind11's export_values(): the enumerators are also module attributes.
globals().update({op.name: op for op in Operator})

MAX_ARITY = 2
NULL_CHILD = -1

_VALID_FIXED = {int(o) for o in Operator if o not in (Operator.NO_OP, Operator.VAR_START_ID)}


def get_arity(op: int) -> int:
    """Arity of an opcode (reference include/operators.h:118-125)."""
    op = int(op)
    if int(Operator.SIN) <= op <= int(Operator.FIFTH):
        return 1
    if int(Operator.ADD) <= op <= int(Operator.POW):
        return 2
    return 0


def is_valid_op(op: int, n_variables: int) -> bool:
    """Validity of an opcode for n_variables inputs (reference include/operators.h:66-116)."""
    op = int(op)
    if op in _VALID_FIXED:
        return True
    return int(Operator.VAR_START_ID) <= op < int(Operator.VAR_START_ID) + int(n_variables)


# ---------------------------------------------------------------------------
# error texts (reference include/error_codes.h:72-195)
# ---------------------------------------------------------------------------
_MESSAGES = {
    0: "No error.",
    101: "Validation failed: A child's index is greater than or equal to its parent's.",
    102: "Validation failed: Operator arity does not match the number of children.",
    201: "Backward pass error: A non-zero gradient was passed to a NO_OP node.",
    204: "Backward pass error: Gradient for a/b where b = 0.",
    301: "Generation failed: Internal stack exceeded capacity.",
    302: "Generation failed: Produced expression exceeded maximum length.",
    303: "Generation failed: Could not produce a valid expression within the max_tries limit.",
    304: "Generation failed: No valid production rule found during sampling.",
    401: "Parsing failed: Output queue overflowed (expression likely too complex).",
    402: "Parsing failed: Operator stack overflowed (expression likely too complex).",
    403: "Parsing failed: Mismatched or unbalanced parentheses.",
    404: "Parsing failed: Unary operator is missing an operand during tree construction.",
    405: "Parsing failed: Binary operator is missing one or both operands during tree construction.",
    406: "Parsing failed: Child stack overflowed during tree construction.",
    407: "Parsing failed: Malformed expression (e.g., too many operands).",
    408: "Parsing failed: Unexpected token encountered during parsing.",
    501: "Conversion failed: Resulting infix expression is longer than the maximum allowed length.",
    502: "Conversion failed: Resulting postfix expression is longer than the maximum allowed length.",
    503: "Conversion failed: Resulting prefix expression is longer than the maximum allowed length.",
    505: "Conversion failed: Malformed input, unary operator is missing an operand.",
    506: "Conversion failed: Malformed input, binary operator is missing one or both operands.",
    507: "Conversion failed: Malformed input (e.g., too many operands).",
    508: "Conversion failed: Resulting infix expression has multiple roots, which is not allowed.",
    509: "Conversion failed: Unexpected token encountered (e.g., an operator without operands).",
    1000: "An internal error occurred that does not fit into the above categories.",
}


def get_error_message(code: int, context: str = "") -> str:
    code = int(code)
    base = _MESSAGES.get(code, f"Unknown error code: {code}")
    if context:
        return f"Error during {context}: {base}"
    return base


# ---------------------------------------------------------------------------
# evaluation backends
# ---------------------------------------------------------------------------
def _require_gpu(t: torch.Tensor, what: str) -> None:
    if t.device.type != "cuda":
        raise RuntimeError(
            f"{what} must be on a ROCm GPU device: this implementation evaluates expressions "
            f"only with its gfx950 kernels (got device '{t.device}')"
        )


def validate_inputs(Ops: torch.Tensor, Ch: torch.Tensor, n_x: int) -> None:
    """Checks child indices and arities (reference src/symbolic_evaluation.cpp:7-37)."""
    if Ops.device != Ch.device:
        raise RuntimeError("Ops and Ch must be on the same device")
    if not Ops.is_contiguous():
        raise RuntimeError("Ops tensor must be contiguous")
    if not Ch.is_contiguous():
        raise RuntimeError("Ch tensor must be contiguous")
    _require_gpu(Ops, "Ops")
    M, B = Ops.shape
    first = torch.empty(1, dtype=torch.int64, device=Ops.device)
    lib = _lib.hip()
    _lib.check_status(
        lib.sr_validate_mb(_lib.ptr(Ops), _lib.ptr(Ch), M, B, _lib.ptr(first), _lib.stream_of(Ops.device)),
        "validate_inputs",
    )
    v = int(first.item())
    if v != -1:  # all ones == no failing node
        raise RuntimeError(get_error_message(v & 1023, "input validation"))


def _workspace(nbytes: int, device) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device)


# Caches smaller than this are evaluated without row workspaces: their extra
# launches would cost more than the narrow cache accesses they avoid.
ROW_WORKSPACE_MIN_CACHE_BYTES = 64 << 20


def _speed_workspace(nbytes: int, device, cache_bytes: int):
    """An optional workspace (row buffers of the full-cache evaluation): any
    size is valid there, so for small caches, or when memory is short, the
    call runs without it (the direct path, same results) instead of failing.
    Returns (tensor or None, its size in bytes)."""
    if nbytes <= 0 or cache_bytes < ROW_WORKSPACE_MIN_CACHE_BYTES:
        return None, 0
    try:
        return torch.empty(int(nbytes), dtype=torch.uint8, device=device), int(nbytes)
    except torch.cuda.OutOfMemoryError:
        return None, 0


class _SymbolicEvaluation(torch.autograd.Function):
    """Autograd function of the full-cache evaluation
    (reference src/symbolic_evaluation.cpp:39-168)."""

    @staticmethod
    def forward(ctx, X, Ops, Ch, C):  # noqa: D401
        M, B = Ops.shape
        N, n_x = X.shape
        cache = torch.empty((M, N, B), dtype=X.dtype, device=X.device)
        is64 = X.dtype == torch.float64
        sfx = "f64" if is64 else "f32"
        lib = _lib.hip()
        ws, ws_n = _speed_workspace(lib.sr_forward_mb_workspace_bytes(M, N, B, 1, int(is64)), X.device,
                                    cache.numel() * cache.element_size())
        _lib.check_status(
            getattr(lib, f"sr_forward_mb_ws_{sfx}")(
                _lib.ptr(X), N, n_x, _lib.ptr(Ops), _lib.ptr(Ch), _lib.ptr(C), M, B, _lib.ptr(cache),
                _lib.ptr(ws), ws_n, _lib.stream_of(X.device),
            ),
            "symbolic forward",
        )
        del ws
        ctx.save_for_backward(X, Ops, Ch, C, cache)
        return cache

    @staticmethod
    def backward(ctx, grad_output):
        X, Ops, Ch, C, cache = ctx.saved_tensors
        if grad_output.shape != cache.shape:
            raise RuntimeError(
                f"Cache and grad_output must have the same shape. Cache shape: {list(cache.shape)}, "
                f"grad_output shape: {list(grad_output.shape)}"
            )
        M, B = Ops.shape
        N, n_x = X.shape
        # grad_cache: the upstream gradient, copied when the kernel of this
        # size uses it as scratch (the grouped kernel only reads it).
        lib0 = _lib.hip()
        g = grad_output.to(dtype=cache.dtype).contiguous()
        grouped = bool(lib0.sr_backward_mb_reads_only(Ops.shape[0], int(X.dtype == torch.float64)))
        if not grouped:
            g = g.clone()
        grad_C = torch.empty_like(C)
        want_x = ctx.needs_input_grad[0]
        grad_X = torch.empty_like(X) if want_x else None
        is64 = X.dtype == torch.float64
        lib = _lib.hip()
        ws = _workspace(lib.sr_backward_mb_workspace_bytes(B, N, n_x, int(is64)) if want_x else 0, X.device)
        # gathered gradient and value rows (read by the kernel that only reads g)
        rws, rws_n = _speed_workspace(lib.sr_backward_mb_rows_workspace_bytes(M, N, B, 1, int(is64)) if grouped else 0,
                                      X.device, cache.numel() * cache.element_size())
        err = torch.empty(1, dtype=torch.int32, device=X.device)
        # the cache is this function's forward output (saved, unmodified):
        # the backward recomputes its values from X and C where it gathers rows
        _lib.check_status(
            getattr(lib, f"sr_backward_mb_rows_recompute_{'f64' if is64 else 'f32'}")(
                _lib.ptr(g), _lib.ptr(cache), _lib.ptr(X), _lib.ptr(C), _lib.ptr(Ops), _lib.ptr(Ch), M, B, N, n_x,
                _lib.ptr(grad_C), _lib.ptr(grad_X), _lib.ptr(ws), ws.numel(), _lib.ptr(rws), rws_n,
                _lib.ptr(err), _lib.stream_of(X.device),
            ),
            "symbolic backward",
        )
        del rws
        code = int(err.item())
        if code != 0:
            raise RuntimeError(get_error_message(code, "symbolic evaluation backward pass"))
        return grad_X, None, None, grad_C


def _check_eval_inputs(X, Ops, Ch, C, multiple: bool) -> None:
    dev = X.device
    if not (Ops.device == dev and Ch.device == dev and C.device == dev):
        raise RuntimeError("All input tensors must be on the same device.")
    if Ops.dim() != 2:
        raise RuntimeError("Ops must be a 2D tensor")
    if Ch.dim() != 3:
        raise RuntimeError("Ch must be a 3D tensor")
    if multiple:
        if C.dim() != 3:
            raise RuntimeError("C must be a 3D tensor for multiple constant evaluation")
        if not (Ops.shape[0] == Ch.shape[0] and Ops.shape[0] == C.shape[0]):
            raise RuntimeError("Ops, Ch, and C must have the same max length (dim 0)")
        if not (Ops.shape[1] == Ch.shape[1] and Ops.shape[1] == C.shape[1]):
            raise RuntimeError("Ops, Ch, and C must have the same batch size (dim 1)")
    else:
        if C.dim() != 2:
            raise RuntimeError("C must be a 2D tensor")
        if Ops.shape[0] != Ch.shape[0]:
            raise RuntimeError("Ops and Ch must have the same max length (dim 0)")
        if Ops.shape[1] != Ch.shape[1]:
            raise RuntimeError("Ops and Ch must have the same batch size (dim 1)")
        if Ops.shape != C.shape:
            raise RuntimeError("Ops and C must have the same shape")
    if Ch.shape[2] != MAX_ARITY:
        raise RuntimeError("Ch must have MaxArity as its last dimension")
    if X.dtype not in (torch.float32, torch.float64):
        raise RuntimeError("X must be a floating-point tensor (float or double)")
    if C.dtype not in (torch.float32, torch.float64):
        raise RuntimeError("C must be a floating-point tensor (float or double)")
    if Ops.dtype != torch.int64:
        raise RuntimeError("Ops must be a Long tensor")
    if Ch.dtype != torch.int64:
        raise RuntimeError("Ch must be a Long tensor")
    if C.dtype != X.dtype:
        raise RuntimeError(f"expected scalar type {X.dtype} for C but found {C.dtype}")
    if X.dim() != 2:
        raise RuntimeError("X must be a 2D tensor (N, n_x)")
    _require_gpu(X, "X")
    if Ops.shape[0] > 128:
        raise RuntimeError("Ops: the gfx950 evaluator supports at most 128 nodes per expression (dim 0)")


def evaluate_backend(X: torch.Tensor, Ops: torch.Tensor, Ch: torch.Tensor, C: torch.Tensor) -> torch.Tensor:
    """Evaluates node-first expressions; returns the (M, N, B) cache
    (reference src/symbolic_evaluation.cpp:306-355)."""
    _check_eval_inputs(X, Ops, Ch, C, multiple=False)
    Ops = Ops.contiguous()
    Ch = Ch.contiguous()
    C = C.contiguous()
    X_c = X.contiguous()
    validate_inputs(Ops, Ch, X.shape[1])
    return _SymbolicEvaluation.apply(X_c, Ops, Ch, C)


class _SymbolicEvaluationMultiple(torch.autograd.Function):
    """Autograd function of the multiple-constant evaluation
    (reference src/symbolic_evaluation.cpp:206-305)."""

    @staticmethod
    def forward(ctx, X, Ops, Ch, C):
        M, B, K = C.shape
        N, n_x = X.shape
        cache = torch.empty((M, N, B, K), dtype=X.dtype, device=X.device)
        is64 = X.dtype == torch.float64
        sfx = "f64" if is64 else "f32"
        lib = _lib.hip()
        ws, ws_n = _speed_workspace(lib.sr_forward_mb_workspace_bytes(M, N, B, K, int(is64)), X.device,
                                    cache.numel() * cache.element_size())
        _lib.check_status(
            getattr(lib, f"sr_multi_forward_mb_ws_{sfx}")(
                _lib.ptr(X), N, n_x, _lib.ptr(Ops), _lib.ptr(Ch), _lib.ptr(C), M, B, K, _lib.ptr(cache),
                _lib.ptr(ws), ws_n, _lib.stream_of(X.device),
            ),
            "symbolic multiple forward",
        )
        del ws
        ctx.save_for_backward(X, Ops, Ch, C, cache)
        return cache

    @staticmethod
    def backward(ctx, grad_output):
        X, Ops, Ch, C, cache = ctx.saved_tensors
        if grad_output.shape != cache.shape:
            raise RuntimeError("Cache and grad_output must have the same shape.")
        M, B, K = C.shape
        N, n_x = X.shape
        g = grad_output.to(dtype=cache.dtype).contiguous()
        grouped = bool(_lib.hip().sr_backward_mb_reads_only(M, int(X.dtype == torch.float64)))
        if not grouped:
            g = g.clone()  # the kernel of this size uses it as scratch
        grad_C = torch.empty_like(C)
        want_x = ctx.needs_input_grad[0]
        grad_X = torch.empty_like(X) if want_x else None
        is64 = X.dtype == torch.float64
        lib = _lib.hip()
        ws = _workspace(lib.sr_multi_backward_mb_workspace_bytes(B, N, n_x, K, int(is64)) if want_x else 0,
                        X.device)
        rws, rws_n = _speed_workspace(lib.sr_backward_mb_rows_workspace_bytes(M, N, B, K, int(is64)) if grouped else 0,
                                      X.device, cache.numel() * cache.element_size())
        err = torch.empty(1, dtype=torch.int32, device=X.device)
        _lib.check_status(
            getattr(lib, f"sr_multi_backward_mb_rows_recompute_{'f64' if is64 else 'f32'}")(
                _lib.ptr(g), _lib.ptr(cache), _lib.ptr(X), _lib.ptr(C), _lib.ptr(Ops), _lib.ptr(Ch), M, B, N, n_x,
                K, _lib.ptr(grad_C), _lib.ptr(grad_X), _lib.ptr(ws), ws.numel(), _lib.ptr(rws), rws_n,
                _lib.ptr(err), _lib.stream_of(X.device),
            ),
            "symbolic multiple backward",
        )
        del rws
        code = int(err.item())
        if code != 0:
            raise RuntimeError(get_error_message(code, "symbolic evaluation (multiple constants) backward pass"))
        return grad_X, None, None, grad_C


def evaluate_multiple_constant_backend(X: torch.Tensor, Ops: torch.Tensor, Ch: torch.Tensor,
                                       C: torch.Tensor) -> torch.Tensor:
    """Evaluates node-first expressions for K constant sets; returns (M, N, B, K)
    (reference src/symbolic_evaluation.cpp:171-204)."""
    _check_eval_inputs(X, Ops, Ch, C, multiple=True)
    Ops = Ops.contiguous()
    Ch = Ch.contiguous()
    C = C.contiguous()
    validate_inputs(Ops, Ch, X.shape[1])
    return _SymbolicEvaluationMultiple.apply(X.contiguous(), Ops, Ch, C)


from .pcfg import ProbabilisticContextFreeGrammar  # noqa: E402  (re-export, as the reference module does)
