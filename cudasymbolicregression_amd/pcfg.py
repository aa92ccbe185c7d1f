"""Probabilistic context-free grammar over the operator vocabulary.

Mirrors the reference's ProbabilisticContextFreeGrammar class
(src/pcfg.cpp:44-1061, include/pcfg.h:26-120): the same symbol table and
ids, grammar-text parser and its error texts, CSR rule tables, and the same
methods, plus `sample(B) -> (Ops, Ch)` from the README contract
(README.md:64-69).  As the reference dispatches on the device
(src/pcfg.cpp:572-986), rows on a GPU go through the device kernels of
libsr_hip.so (sampling on a grammar whose device is a GPU, infix ->
postfix, the children table and postfix -> infix; one lane per row,
include/sr_pcfg_device.h), and the rest through the host engine
libsr_host.so (one OpenMP iteration per row), whose results are copied to
the input's device.  Both give the same tokens and error codes; sampling
on a GPU draws its own stream (Philox-4x32-10), as the reference's CUDA
sampler does (curand there), and `sample_infix(B, engine="host")` keeps
the reference's CPU stream on any device.
"""
from __future__ import annotations

import math
import re
from typing import Dict, List, Tuple

import torch

from . import _lib
from ._C import Operator, get_arity, get_error_message, is_valid_op

HARD_MAX_LENGTH = 128
DEFAULT_TOLERANCE = 9.999999974752427e-07  # the float32 value of 1e-6
INT64_MAX = 9223372036854775807

# (spelling, opcode) pairs in the registration order of src/pcfg.cpp:44-170.
# A later spelling of an id does not replace the first one as the id's
# printable name; a later registration of a spelling re-targets it.
_SYMBOLS: List[Tuple[str, int]] = [
    ("NO_OP", 0), ("C", 1),
    ("1", 2), ("2", 3), ("3", 4), ("4", 5), ("5", 6), ("pi", 7), ("PI", 7), ("e", 8), ("E", 8),
    ("sin", 10), ("SIN", 10), ("Sin", 10),
    ("cos", 11), ("COS", 11), ("Cos", 11),
    ("exp", 12), ("EXP", 12), ("Exp", 12),
    ("log", 13), ("LOG", 13), ("Log", 13), ("ln", 13), ("LN", 13), ("Ln", 13),
    ("^2", 14), ("square", 14), ("SQUARE", 14), ("Square", 14),
    ("sqrt", 15), ("SQRT", 15), ("Sqrt", 15),
    ("tan", 16), ("TAN", 16), ("Tan", 16),
    ("arcsin", 17), ("ARCSIN", 17), ("asin", 17), ("ASIN", 17),
    ("arccos", 18), ("ARCCOS", 18), ("acos", 18), ("ACOS", 18),
    ("arctan", 19), ("ARCTAN", 19), ("atan", 19), ("ATAN", 19),
    ("sinh", 20), ("SINH", 20), ("Sinh", 20),
    ("cosh", 21), ("COSH", 21), ("Cosh", 21),
    ("tanh", 22), ("TANH", 22), ("Tanh", 22),
    ("floor", 23), ("FLOOR", 23), ("Floor", 23),
    ("ceil", 24), ("CEIL", 24), ("Ceil", 24),
    ("ln", 25), ("LN", 25),
    ("log10", 26), ("LOG10", 26),
    ("u-", 27), ("neg", 27), ("NEG", 27), ("minus", 27),
    ("^-1", 28), ("inv", 28), ("INV", 28),
    ("^3", 29), ("cube", 29), ("CUBE", 29), ("Cube", 29),
    ("^4", 30), ("fourth", 30), ("FOURTH", 30), ("Fourth", 30),
    ("^5", 31), ("fifth", 31), ("FIFTH", 31), ("Fifth", 31),
    ("+", 40), ("-", 41), ("*", 42), ("/", 43), ("^", 44),
]

# Shunting-yard precedence (src/pcfg.cpp:177-236): operands 0, + - 1, * / 2,
# ^ 3, unary functions and parentheses 4.
_PREC_BINARY = {40: 1, 41: 1, 42: 2, 43: 2, 44: 3}


def _trim(s: str) -> str:
    return s.strip(" \t\n\v\f\r")


_FLOAT_PREFIX = re.compile(
    r"[ \t\n\v\f\r]*([+-]?(?:(?:\d+\.?\d*(?:[eE][+-]?\d+)?)|(?:\.\d+(?:[eE][+-]?\d+)?)|inf(?:inity)?|nan))",
    re.IGNORECASE,
)


def _stof(s: str) -> float:
    """std::stof: parses the longest leading float, as float32."""
    m = _FLOAT_PREFIX.match(s)
    if not m:
        raise ValueError("stof")
    v = float(m.group(1))
    return float(torch.tensor(v, dtype=torch.float32).item())


def _cpp_double_str(v: float) -> str:
    """std::to_string(double) (printf "%f")."""
    return "%f" % v


class ProbabilisticContextFreeGrammar:
    """Grammar compiler, sampler and expression-format converter."""

    def __init__(
        self,
        grammar: str,
        start_symbol: str,
        padded_maximum_length: int,
        n_variables: int,
        device: torch.device,
        max_tries: int = 64,
        tolerance: float = DEFAULT_TOLERANCE,
        verbose: bool = False,
    ) -> None:
        self._start_symbol = str(start_symbol)
        self._padded_maximum_length = int(padded_maximum_length)
        self._n_variables = int(n_variables)
        self._device = torch.device(device)
        self._max_tries = int(max_tries)
        self._tolerance = float(torch.tensor(float(tolerance), dtype=torch.float32).item())
        self._verbose = bool(verbose)
        if self._padded_maximum_length > HARD_MAX_LENGTH:
            raise ValueError(
                f"padded_maximum_length must be less than or equal to HARD_MAX_LENGTH ({HARD_MAX_LENGTH})"
            )
        self._build_symbol_table()
        self._parse_grammar(grammar)
        self._device_tables: Dict[str, Dict[str, torch.Tensor]] = {}

    # ------------------------------------------------------------ symbols
    def _add_symbol(self, s: str, i: int) -> None:
        self._symbol_to_id[s] = i
        if i not in self._id_to_symbol:
            self._id_to_symbol[i] = s

    def _build_symbol_table(self) -> None:
        self._symbol_to_id: Dict[str, int] = {}
        self._id_to_symbol: Dict[int, str] = {}
        for s, i in _SYMBOLS:
            self._add_symbol(s, i)
        nv = self._n_variables
        for i in range(nv):
            self._add_symbol(f"X_{i}", int(Operator.VAR_START_ID) + i)
        self._lparen = int(Operator.VAR_START_ID) + nv
        self._rparen = self._lparen + 1
        self._add_symbol("(", self._lparen)
        self._add_symbol(")", self._rparen)
        self._next_id = int(Operator.VAR_START_ID) + nv + 2
        self._terminal_limit = self._next_id
        self._n_operators = int(Operator.VAR_START_ID) + nv
        prec = [0] * self._next_id
        for op, p in _PREC_BINARY.items():
            prec[op] = p
        for op in range(int(Operator.SIN), int(Operator.FIFTH) + 1):
            if op != 32:
                prec[op] = 4
        prec[self._lparen] = 4
        prec[self._rparen] = 4
        self._precedence_list = prec

    def _token_id(self, s: str) -> int:
        if not s:
            raise ValueError("Symbol cannot be empty.")
        if s in self._symbol_to_id:
            return self._symbol_to_id[s]
        i = self._next_id
        self._next_id += 1
        self._add_symbol(s, i)
        return i

    def _parse_sides(self, s: str) -> List[int]:
        return [self._token_id(tok) for tok in s.split()]

    # ------------------------------------------------------------ grammar
    def _parse_grammar(self, grammar: str) -> None:
        lhs_ids: List[int] = []
        rhs_list: List[List[int]] = []
        probs: List[float] = []
        for raw in grammar.split("\n"):
            line = _trim(raw)
            if not line:
                continue
            arrow = line.find("->")
            if arrow < 0:
                raise ValueError("Invalid grammar rule: " + line)
            lhs = _trim(line[:arrow])
            if not lhs:
                raise ValueError("LHS cannot be empty in rule: " + line)
            rhs_and_prob = _trim(line[arrow + 2:])
            if "->" in rhs_and_prob:
                raise ValueError("Invalid grammar rule with multiple '->': " + line)
            ppos = rhs_and_prob.find("[")
            if ppos < 0:
                raise ValueError("Invalid grammar rule without probability: " + line)
            rhs = _trim(rhs_and_prob[:ppos])
            prob_str = rhs_and_prob[ppos + 1:]
            if not prob_str.endswith("]"):
                raise ValueError("Invalid grammar rule with missing ']': " + line)
            if prob_str.count("]") > 1:
                raise ValueError("Invalid grammar rule with multiple ']' characters: " + line)
            p = _stof(_trim(prob_str[:-1]))
            lhs_tokens = self._parse_sides(lhs)
            if len(lhs_tokens) != 1:
                raise ValueError("LHS must contain exactly one non-terminal: " + lhs)
            rhs_tokens = self._parse_sides(rhs)
            if not rhs_tokens:
                raise ValueError("RHS cannot be empty: " + rhs)
            lhs_ids.append(lhs_tokens[0])
            rhs_list.append(rhs_tokens)
            probs.append(p)

        n = len(lhs_ids)
        perm = torch.zeros(n, dtype=torch.int64)
        if n:
            lhs_t = torch.tensor(lhs_ids, dtype=torch.int64)
            _lib.host().sr_pcfg_sort_rules(_lib.ptr(lhs_t), n, _lib.ptr(perm))
        order = perm.tolist()
        s_lhs = [lhs_ids[i] for i in order]
        s_rhs = [rhs_list[i] for i in order]
        s_prob = [probs[i] for i in order]
        concat: List[int] = []
        rptr = [0]
        for r in s_rhs:
            concat.extend(r)
            rptr.append(len(concat))
        if not s_lhs:
            raise ValueError("No rules found in the grammar.")
        nsym = self._next_id
        ntp = [0] * (nsym + 1)
        ntp[s_lhs[0]] = 0
        for i in range(1, n):
            for j in range(s_lhs[i - 1] + 1, s_lhs[i] + 1):
                ntp[j] = i
        for j in range(s_lhs[-1] + 1, nsym + 1):
            ntp[j] = n
        cum = [0.0] * n
        for i in range(nsym):
            a, b = ntp[i], ntp[i + 1]
            if a == b:
                continue
            total = 0.0
            for j in range(a, b):
                total += s_prob[j]
            if abs(total - 1.0) > self._tolerance:
                name = self._id_to_symbol.get(i, "")
                raise ValueError(
                    f"Non-terminal '{name}' does not sum to 1.0. Total probability: {_cpp_double_str(total)}."
                )
            c = 0.0
            for j in range(a, b):
                c += s_prob[j]
                cum[j] = c
            cum[b - 1] = 1.0
        for i in range(self._terminal_limit, nsym):
            if ntp[i] == ntp[i + 1]:
                raise ValueError(
                    f"Symbol {self._id_to_symbol[i]} with ID {i} has no rules and is not a terminal."
                )
        self._rule_lhs = torch.tensor(s_lhs, dtype=torch.int64)
        self._rhs_ptr = torch.tensor(rptr, dtype=torch.int64)
        self._rhs_concat = torch.tensor(concat, dtype=torch.int64)
        self._nt_rule_ptr = torch.tensor(ntp, dtype=torch.int64)
        self._cum_probs = torch.tensor(cum, dtype=torch.float32)
        self._precedence = torch.tensor(self._precedence_list + [0] * (nsym - len(self._precedence_list)),
                                        dtype=torch.int64)
        if self._verbose:
            print(f"ProbabilisticContextFreeGrammar initialized with {n} rules.")
            print(f"\tStart symbol: {self._start_symbol}, ID: {self._token_id(self._start_symbol)}")
            print(f"\tMaximum length: {self._padded_maximum_length}")
            print(f"\tNumber of variables: {self._n_variables}")
            print(f"\tDevice: {self._device}")
            print(f"\tMax tries: {self._max_tries}")
            print(f"\tTolerance: {self._tolerance:g}")
            for s, i in self._symbol_to_id.items():
                print(f"\tSymbol: {s}, ID: {i}")
            print("Processed grammar rules:")
            for i in range(n):
                rhs_s = " ".join(self._id_to_symbol[t] for t in s_rhs[i])
                print(f"\tRule {i + 1}: {self._id_to_symbol[s_lhs[i]]} -> {rhs_s} [{s_prob[i]:g}]")

    # ------------------------------------------------------- attributes
    device = property(lambda self: self._device)
    start_symbol = property(lambda self: self._start_symbol)
    padded_maximum_length = property(lambda self: self._padded_maximum_length)
    n_variables = property(lambda self: self._n_variables)
    max_tries = property(lambda self: self._max_tries)
    tolerance = property(lambda self: self._tolerance)
    verbose = property(lambda self: self._verbose)
    n_operators = property(lambda self: self._n_operators)

    def get_symbol_id(self, symbol: str) -> int:
        if symbol in self._symbol_to_id:
            return self._symbol_to_id[symbol]
        raise RuntimeError("Symbol not found: " + symbol)

    def get_arities(self) -> torch.Tensor:
        return torch.tensor([get_arity(i) for i in range(self._n_operators)], dtype=torch.int64,
                            device=self._device)

    def valid_ops(self) -> torch.Tensor:
        return torch.tensor([is_valid_op(i, self._n_variables) for i in range(self._n_operators)],
                            dtype=torch.bool, device=self._device)

    def available_operators(self) -> List[str]:
        return [s for s, i in self._symbol_to_id.items() if i < self._n_operators]

    # ------------------------------------------------------- helpers
    def _tables_on(self, dev: torch.device) -> Dict[str, torch.Tensor]:
        """The CSR grammar and precedence table on a GPU (copied once)."""
        key = str(dev)
        t = self._device_tables.get(key)
        if t is None:
            t = dict(rhs_ptr=self._rhs_ptr.to(dev), rhs_concat=self._rhs_concat.to(dev),
                     nt_rule_ptr=self._nt_rule_ptr.to(dev), cum=self._cum_probs.to(dev),
                     prec=self._precedence.to(dev))
            self._device_tables[key] = t
        return t

    @staticmethod
    def _on_gpu(t: torch.Tensor) -> bool:
        return t.device.type == "cuda"

    def _host_i64(self, t: torch.Tensor) -> torch.Tensor:
        return t.detach().to("cpu", torch.int64).contiguous()

    def _check_range(self, t: torch.Tensor, limit: int, what: str, msg_limit: str) -> None:
        if t.numel() and not bool((t >= 0).all()):
            raise RuntimeError(f"All symbols in the {what} must be non-negative integers.")
        if t.numel() and not bool((t < limit).all()):
            raise RuntimeError(f"All symbols in the {what} must be {msg_limit}.")

    def _raise_errors(self, errors: torch.Tensor, expressions: torch.Tensor, context: str, verbosity: int) -> None:
        errs = errors.tolist()
        if not any(errs):
            return
        counts: Dict[int, int] = {}
        for e in errs:
            if e:
                counts[e] = counts.get(e, 0) + 1
        lines = [f"Failed to parse expressions during {context} for some batch items.\n", "Error Summary:\n"]
        for code in sorted(counts):
            lines.append(f"  {get_error_message(code)}: {counts[code]} occurrences.\n")
        if verbosity > 0:
            bad = [i for i, e in enumerate(errs) if e]
            limit = min(5, len(bad)) if verbosity == 1 else len(bad)
            lines.append(f"\nDisplaying {limit} out of {len(bad)} problematic expressions:\n")
            exprs = expressions.to("cpu")
            for i in bad[:limit]:
                row = exprs[i].unsqueeze(0)
                lines.append("----------------------------------------\n")
                lines.append(f"Problem at Index: {i}\n")
                lines.append(f"  Error: {get_error_message(errs[i])}\n")
                lines.append(f"  Input Expression: '{self.to_string(row)[0]}'\n")
                body = "  ".join(str(int(v)) for v in row[0].tolist())
                lines.append(f"  Expression Tensor:  {body}\n[ CPULongType{{1,{row.shape[1]}}} ]\n")
            if len(bad) > limit:
                lines.append("----------------------------------------\n")
                lines.append(f"... and {len(bad) - limit} more.\n")
        raise RuntimeError("".join(lines))

    # ------------------------------------------------------- sampling
    def _sample_errors(self, errs: List[int]) -> None:
        if any(errs):
            counts: Dict[int, int] = {}
            for e in errs:
                if e:
                    counts[e] = counts.get(e, 0) + 1
            msg = "Failed to generate expressions for some batch items. Error counts:\n"
            for code in sorted(counts):
                msg += f"  {get_error_message(code)}: {counts[code]} occurrences.\n"
            raise RuntimeError(msg)

    def sample_infix(self, B: int, engine: str = "auto") -> torch.Tensor:
        """B expressions in infix token form, NO_OP padded to padded_maximum_length.

        engine "auto": on the device kernels when the grammar's device is a
        GPU (Philox-4x32-10 rows from one seed drawn from torch's generator
        of that device), else the host engine; "host": the host engine
        (std::mt19937 per row, seeds from torch's generator, the
        reference's CPU stream) on any device."""
        B = int(B)
        L = self._padded_maximum_length
        start_id = self._token_id(self._start_symbol)
        if engine not in ("auto", "host"):
            raise ValueError(f"engine must be 'auto' or 'host', got {engine!r}")
        if engine == "auto" and self._device.type == "cuda":
            if start_id >= self._nt_rule_ptr.numel() - 1:
                raise RuntimeError("Start symbol has no rules: " + self._start_symbol)
            seed = int(torch.randint(0, INT64_MAX, (1,), dtype=torch.int64, device=self._device).item())
            t = self._tables_on(self._device)
            out = torch.empty((B, L), dtype=torch.int64, device=self._device)
            errors = torch.empty(B, dtype=torch.int64, device=self._device)
            _lib.check_status(_lib.hip().sr_pcfg_sample_device(
                _lib.ptr(t["rhs_ptr"]), _lib.ptr(t["rhs_concat"]), _lib.ptr(t["nt_rule_ptr"]), _lib.ptr(t["cum"]),
                self._nt_rule_ptr.numel() - 1, start_id, L, self._max_tries, _lib.ptr(out), _lib.ptr(errors),
                seed, B, _lib.stream_of(self._device)), "sr_pcfg_sample_device")
            if B and bool(errors.any()):
                self._sample_errors(errors.tolist())
            return out
        seeds = torch.randint(0, INT64_MAX, (B,), dtype=torch.int64, device=self._device).to("cpu")
        out = torch.zeros((B, L), dtype=torch.int64)
        errors = torch.zeros(B, dtype=torch.int64)
        if start_id >= self._nt_rule_ptr.numel() - 1:
            raise RuntimeError("Start symbol has no rules: " + self._start_symbol)
        _lib.host().sr_pcfg_sample(
            _lib.ptr(self._rule_lhs), _lib.ptr(self._rhs_ptr), _lib.ptr(self._rhs_concat),
            _lib.ptr(self._nt_rule_ptr), _lib.ptr(self._cum_probs), start_id, L, self._max_tries,
            _lib.ptr(out), _lib.ptr(errors), _lib.ptr(seeds), B,
        )
        self._sample_errors(errors.tolist())
        return out.to(self._device)

    def sample(self, B: int) -> Tuple[torch.Tensor, torch.Tensor]:
        """B postfix expressions and their children table: (Ops (B, L), Ch (B, L, 2))."""
        ops = self.infix_to_postfix(self.sample_infix(B))
        return ops, self.get_postfix_children(ops)

    def to_string(self, expressions: torch.Tensor) -> List[str]:
        rows = self._host_i64(expressions).tolist()
        out = []
        for row in rows:
            toks = []
            for t in row:
                if t == 0:
                    break
                if t in self._id_to_symbol:
                    toks.append(self._id_to_symbol[t])
            out.append(" ".join(toks))
        return out

    # ------------------------------------------------------- parsing
    def _parse_infix(self, expressions: torch.Tensor, fn: str, context: str, verbosity: int) -> torch.Tensor:
        e = self._host_i64(expressions)
        self._check_range(e, self._terminal_limit, "infix_ops", "terminal symbols (less than terminal_limit)")
        B, M = e.shape
        if M > HARD_MAX_LENGTH:
            raise ValueError(
                f"M must be less than or equal to HARD_MAX_LENGTH, but got {M} > {HARD_MAX_LENGTH}."
            )
        out = torch.zeros((B, M), dtype=torch.int64)
        errors = torch.zeros(B, dtype=torch.int64)
        getattr(_lib.host(), fn)(_lib.ptr(self._precedence), _lib.ptr(e), _lib.ptr(out), _lib.ptr(errors),
                                 self._lparen, self._rparen, B, M)
        self._raise_errors(errors, e, context, verbosity)
        return out.to(expressions.device)

    def infix_to_postfix(self, expressions: torch.Tensor, verbosity: int = 0) -> torch.Tensor:
        if self._on_gpu(expressions):
            e = expressions.detach().to(torch.int64).contiguous()
            self._check_range(e, self._terminal_limit, "infix_ops", "terminal symbols (less than terminal_limit)")
            B, M = e.shape
            if M > HARD_MAX_LENGTH:
                raise ValueError(
                    f"M must be less than or equal to HARD_MAX_LENGTH, but got {M} > {HARD_MAX_LENGTH}."
                )
            prec = self._tables_on(e.device)["prec"]
            out = torch.empty((B, M), dtype=torch.int64, device=e.device)
            errors = torch.empty(B, dtype=torch.int64, device=e.device)
            if B:
                _lib.check_status(_lib.hip().sr_pcfg_infix_to_postfix_device(
                    _lib.ptr(prec), prec.numel(), _lib.ptr(e), _lib.ptr(out), _lib.ptr(errors), self._lparen,
                    self._rparen, B, M, _lib.stream_of(e.device)), "sr_pcfg_infix_to_postfix_device")
                if bool(errors.any()):
                    self._raise_errors(errors.cpu(), e.cpu(), "postfix parsing", verbosity)
            return out
        return self._parse_infix(expressions, "sr_pcfg_infix_to_postfix", "postfix parsing", verbosity)

    def infix_to_prefix(self, expressions: torch.Tensor, verbosity: int = 0) -> torch.Tensor:
        return self._parse_infix(expressions, "sr_pcfg_infix_to_prefix", "prefix parsing", verbosity)

    def _to_infix(self, expressions: torch.Tensor, max_infix_len: int, fn: str, context: str,
                  verbosity: int, check_device: bool) -> torch.Tensor:
        if check_device and expressions.device != self._device:
            raise RuntimeError(
                f"Input tensor must be on the same device as the PCFG, but got input on {expressions.device} "
                f"and PCFG on {self._device}."
            )
        Mi = int(max_infix_len)
        if self._on_gpu(expressions) and fn == "sr_pcfg_postfix_to_infix":
            e = expressions.detach().to(torch.int64).contiguous()
            self._check_range(e, self._n_operators, "prefix_ops", "valid operator IDs")
            B, M = e.shape
            out = torch.empty((B, Mi), dtype=torch.int64, device=e.device)
            errors = torch.empty(B, dtype=torch.int64, device=e.device)
            if B and M and Mi > 0:
                if M > HARD_MAX_LENGTH:
                    raise ValueError(
                        f"M must be less than or equal to HARD_MAX_LENGTH, but got {M} > {HARD_MAX_LENGTH}."
                    )
                _lib.check_status(_lib.hip().sr_pcfg_postfix_to_infix_device(
                    _lib.ptr(e), _lib.ptr(out), _lib.ptr(errors), self._lparen, self._rparen, B, M, Mi,
                    _lib.stream_of(e.device)), "sr_pcfg_postfix_to_infix_device")
                if bool(errors.any()):
                    self._raise_errors(errors.cpu(), e.cpu(), context, verbosity)
                return out
        e = self._host_i64(expressions)
        self._check_range(e, self._n_operators, "prefix_ops", "valid operator IDs")
        B, M = e.shape
        out = torch.zeros((B, Mi), dtype=torch.int64)
        errors = torch.zeros(B, dtype=torch.int64)
        getattr(_lib.host(), fn)(_lib.ptr(e), _lib.ptr(out), _lib.ptr(errors), self._lparen, self._rparen, B, M, Mi)
        self._raise_errors(errors, e, context, verbosity)
        return out.to(expressions.device)

    def postfix_to_infix(self, expressions: torch.Tensor, max_infix_len: int, verbosity: int = 0) -> torch.Tensor:
        return self._to_infix(expressions, max_infix_len, "sr_pcfg_postfix_to_infix",
                              "postfix to infix conversion", verbosity, check_device=True)

    def prefix_to_infix(self, expressions: torch.Tensor, max_infix_len: int, verbosity: int = 0) -> torch.Tensor:
        return self._to_infix(expressions, max_infix_len, "sr_pcfg_prefix_to_infix",
                              "prefix to infix conversion", verbosity, check_device=False)

    def _reorder(self, expressions: torch.Tensor, fn: str, context: str, what: str, verbosity: int) -> torch.Tensor:
        e = self._host_i64(expressions)
        self._check_range(e, self._n_operators, what, "valid operator IDs")
        B, M = e.shape
        out = torch.zeros((B, M), dtype=torch.int64)
        errors = torch.zeros(B, dtype=torch.int64)
        getattr(_lib.host(), fn)(_lib.ptr(e), _lib.ptr(out), _lib.ptr(errors), B, M)
        self._raise_errors(errors, e, context, verbosity)
        return out.to(expressions.device)

    def prefix_to_postfix(self, expressions: torch.Tensor, verbosity: int = 0) -> torch.Tensor:
        return self._reorder(expressions, "sr_pcfg_prefix_to_postfix", "prefix to postfix parent conversion",
                             "prefix_ops", verbosity)

    def postfix_to_prefix(self, expressions: torch.Tensor, verbosity: int = 0) -> torch.Tensor:
        return self._reorder(expressions, "sr_pcfg_postfix_to_prefix", "postfix to prefix parent conversion",
                             "postfix_ops", verbosity)

    def get_prefix_parent(self, prefix_ops: torch.Tensor, verbosity: int = 0) -> torch.Tensor:
        e = self._host_i64(prefix_ops)
        self._check_range(e, self._n_operators, "prefix_ops", "valid operator IDs")
        B, M = e.shape
        out = torch.full((B, M), -1, dtype=torch.int64)
        errors = torch.zeros(B, dtype=torch.int64)
        _lib.host().sr_pcfg_get_prefix_parent(_lib.ptr(e), _lib.ptr(out), _lib.ptr(errors), B, M)
        self._raise_errors(errors, e, "get prefix parent", verbosity)
        return out.to(prefix_ops.device)

    def get_postfix_children(self, postfix_ops: torch.Tensor, verbosity: int = 0) -> torch.Tensor:
        if self._on_gpu(postfix_ops) and postfix_ops.dim() == 2 and 0 < postfix_ops.shape[1] <= HARD_MAX_LENGTH:
            e = postfix_ops.detach().to(torch.int64).contiguous()
            self._check_range(e, self._n_operators, "postfix_ops", "valid operator IDs")
            B, M = e.shape
            out = torch.empty((B, M, 2), dtype=torch.int64, device=e.device)
            errors = torch.empty(B, dtype=torch.int64, device=e.device)
            if B:
                _lib.check_status(_lib.hip().sr_pcfg_get_postfix_children_device(
                    _lib.ptr(e), _lib.ptr(out), _lib.ptr(errors), B, M, _lib.stream_of(e.device)),
                    "sr_pcfg_get_postfix_children_device")
                if bool(errors.any()):
                    self._raise_errors(errors.cpu(), e.cpu(), "get postfix children", verbosity)
            return out
        e = self._host_i64(postfix_ops)
        self._check_range(e, self._n_operators, "postfix_ops", "valid operator IDs")
        B, M = e.shape
        out = torch.full((B, M, 2), -1, dtype=torch.int64)
        errors = torch.zeros(B, dtype=torch.int64)
        _lib.host().sr_pcfg_get_postfix_children(_lib.ptr(e), _lib.ptr(out), _lib.ptr(errors), B, M)
        self._raise_errors(errors, e, "get postfix children", verbosity)
        return out.to(postfix_ops.device)
