"""Pins of the grammar compiler and the sampler's distribution
(reference src/pcfg.cpp:274-500, tests/test_sampling.py).

* The CSR tables are rebuilt here from the rule text alone, by the rules of
  src/pcfg.cpp:362-458: rules ordered by left-hand-side id with std::sort
  (no stable-sort guarantee: for more than 16 rules libstdc++'s introsort
  may reorder rules of one non-terminal, so the expected order comes from a
  small C++ program that calls std::sort with the same comparison);
  nt_rule_ptr[j] = number of rules whose left side is below j; cumulative
  probabilities summed in double over the float probabilities, stored as
  float, and the last of each non-terminal set to exactly 1.0.  Non-terminal
  ids are assigned in order of first appearance, from 52 + n_variables
  (tests/test_sampling.py:161-166 expects id 54 for n_variables = 1).
* The distribution of sampled expressions matches a plain Python sampler of
  the same grammar within 20% on the reference's target expressions, over
  100k samples (tests/test_sampling.py:169-287).
* The trivial grammars of tests/test_sampling.py:50-123 and the max_tries
  failure (:208-230).
"""
import os
import random
import subprocess
from collections import defaultdict

import numpy as np
import pytest
import torch

import helpers
from cudasymbolicregression_amd import Operator, ProbabilisticContextFreeGrammar

CPU = torch.device("cpu")

README_GRAMMAR = """E -> E + T [0.4]
E -> T [0.6]
T -> F * T [0.4]
T -> F [0.6]
F -> R ( V ) [0.3]
F -> V [0.7]
R -> sin [0.5]
R -> cos [0.5]
V -> X_0 [0.5]
V -> C [0.5]
"""

# the reference's tests/test_sampling.py test grammar (17 rules, out-of-order
# non-terminals, probabilities that are not exactly representable)
TEST_GRAMMAR = """E -> E + F [0.2]
E -> E - F [0.2]
E -> F [0.6]
F -> F * T [0.2]
F -> F / T [0.2]
F -> T [0.6]
T -> ( E ) P [0.2]
T -> R ( E ) [0.2]
T -> V [0.4000000000000001]
T -> ( E ) [0.20000000000000004]
V -> X_0 [1.0]
R -> sin [0.2]
R -> cos [0.2]
R -> exp [0.2]
R -> sqrt [0.2]
R -> log [0.2]
P -> ^2 [1.0]
"""


def _interleaved_grammar():
    """40 rules whose non-terminals are listed out of order and interleaved,
    so that the sort moves rules of one non-terminal past each other."""
    rng = np.random.default_rng(0)
    nts = ["A", "B", "D", "G", "H"]
    rules = []
    for nt in nts:
        k = 8
        w = rng.uniform(0.5, 1.5, k)
        w = w / w.sum()
        for i in range(k):
            rhs = rng.choice(["X_0", "1", "C", "2", "X_1"]) if i % 2 else f"{nt} + X_0"
            rules.append(f"{nt} -> {rhs} [{float(w[i])!r}]")
    order = rng.permutation(len(rules))
    return "S -> A [0.2]\nS -> B [0.2]\nS -> D [0.2]\nS -> G [0.2]\nS -> H [0.2]\n" + "\n".join(
        rules[i] for i in order)


GRAMMARS = [("readme", README_GRAMMAR, 1), ("g2", helpers.G2, 3), ("test", TEST_GRAMMAR, 1),
            ("interleaved", _interleaved_grammar(), 2)]


@pytest.fixture(scope="module")
def std_sort(tmp_path_factory):
    """std::sort over rule indices by left-hand-side id (libstdc++), as the
    reference orders its rules."""
    d = tmp_path_factory.mktemp("stdsort")
    src = d / "s.cpp"
    src.write_text("""#include <algorithm>
#include <cstdio>
#include <numeric>
#include <vector>
int main() { long n; if (std::scanf("%ld", &n) != 1) return 1; std::vector<long> v(n), idx(n);
  for (auto &x : v) if (std::scanf("%ld", &x) != 1) return 1;
  std::iota(idx.begin(), idx.end(), 0L);
  std::sort(idx.begin(), idx.end(), [&](long a, long b) { return v[a] < v[b]; });
  for (long i : idx) std::printf("%ld ", i); return 0; }
""")
    exe = d / "s"
    subprocess.run(["g++", "-O2", "-std=c++17", str(src), "-o", str(exe)], check=True)

    def run(lhs):
        out = subprocess.run([str(exe)], input=f"{len(lhs)} " + " ".join(map(str, lhs)), capture_output=True,
                             text=True, check=True).stdout
        return [int(t) for t in out.split()]

    return run


def _expected_tables(grammar, n_x, pcfg, sort):
    """The CSR tables from the rule text (see the module docstring)."""
    term_ids = {}
    next_id = 52 + n_x
    lhs, rhs, probs = [], [], []

    def tid(tok):
        nonlocal next_id
        try:
            return pcfg.get_symbol_id(tok)  # terminals and symbols seen so far
        except Exception:
            pass
        if tok not in term_ids:
            term_ids[tok] = next_id
            next_id += 1
        return term_ids[tok]

    for line in grammar.strip().split("\n"):
        l_s, r_s = line.split("->")
        r_s, p_s = r_s.rsplit("[", 1)
        lhs.append(tid(l_s.strip()))
        rhs.append([tid(t) for t in r_s.split()])
        probs.append(float(np.float32(p_s.strip().rstrip("]"))))
    order = sort(lhs)
    s_lhs = [lhs[i] for i in order]
    s_rhs = [rhs[i] for i in order]
    s_p = [probs[i] for i in order]
    concat = [t for r in s_rhs for t in r]
    ptr = np.cumsum([0] + [len(r) for r in s_rhs])
    nsym = max(next_id, max(max(r) for r in rhs) + 1, max(lhs) + 1)
    ntp = [sum(1 for x in s_lhs if x < j) for j in range(nsym + 1)]
    cum = np.zeros(len(s_lhs), np.float32)
    for j in range(nsym):
        a, b = ntp[j], ntp[j + 1]
        c = 0.0
        for i in range(a, b):
            c += s_p[i]
            cum[i] = np.float32(c)
        if b > a:
            cum[b - 1] = np.float32(1.0)
    return np.array(s_lhs), ptr, np.array(concat), np.array(ntp), cum


@pytest.mark.parametrize("name,grammar,n_x", GRAMMARS, ids=[g[0] for g in GRAMMARS])
def test_csr_tables_match_the_rule_text(std_sort, name, grammar, n_x):
    g = ProbabilisticContextFreeGrammar(grammar, "S" if name == "interleaved" else "E", 64, n_x, CPU)
    e_lhs, e_ptr, e_cat, e_ntp, e_cum = _expected_tables(grammar, n_x, g, std_sort)
    assert np.array_equal(g._rule_lhs.numpy(), e_lhs)
    assert np.array_equal(g._rhs_ptr.numpy(), e_ptr)
    assert np.array_equal(g._rhs_concat.numpy(), e_cat)
    assert np.array_equal(g._nt_rule_ptr.numpy()[: len(e_ntp)], e_ntp)
    assert np.array_equal(g._cum_probs.numpy().view(np.int32), e_cum.view(np.int32))
    if name in ("test", "interleaved"):
        # more than 16 rules: the order of equal keys is std::sort's own
        assert len(e_lhs) > 16


def test_nonterminal_ids_follow_first_appearance():
    """tests/test_sampling.py:161-166: with n_variables = 1 the first
    non-terminal gets id 53 and an undefined symbol A id 54."""
    with pytest.raises(ValueError, match="Symbol A with ID 54 has no rules and is not a terminal."):
        ProbabilisticContextFreeGrammar("S -> A [1.0]", "S", 10, 1, CPU)
    g = ProbabilisticContextFreeGrammar("S -> T [1.0]\nT -> X_0 [1.0]", "S", 10, 1, CPU)
    assert g.get_symbol_id("(") == 51 and g.get_symbol_id(")") == 52
    assert g.get_symbol_id("S") == 53 and g.get_symbol_id("T") == 54


def test_cumulative_probabilities_last_is_one():
    # 0.1 + 0.2 + 0.7 summed in double is 0.9999999... in float: stored 1.0
    g = ProbabilisticContextFreeGrammar("S -> 1 [0.1]\nS -> 2 [0.2]\nS -> 3 [0.7]", "S", 4, 0, CPU)
    cum = g._cum_probs.numpy()
    assert cum[-1] == np.float32(1.0)
    assert cum[0] == np.float32(np.float32(0.1))
    assert cum[1] == np.float32(float(np.float32(0.1)) + float(np.float32(0.2)))


# ------------------------------------------------------------ distribution
def _parse(grammar):
    rules = defaultdict(list)
    for line in grammar.strip().split("\n"):
        lhs, rp = line.split("->")
        rhs, p = rp.rsplit("[", 1)
        rules[lhs.strip()].append((rhs.split(), float(p.strip().rstrip("]"))))
    return rules


def _py_sample(rules, start, limit, rnd):
    """Leftmost expansion with an explicit stack (the reference test's
    recursive sampler, tests/test_sampling.py:183-206, without recursion).
    A draw is abandoned and drawn again once it must exceed `limit` tokens:
    every right-hand side is non-empty, so each pending symbol still adds at
    least one token.  Abandoning early or late gives the same distribution
    of accepted draws."""
    import bisect

    table = {}
    for k, v in rules.items():
        cum, c = [], 0.0
        for _, p in v:
            c += p
            cum.append(c)
        table[k] = ([prod for prod, _ in v], cum)
    while True:
        out, stack = [], [start]
        while stack and len(out) + len(stack) <= limit:
            sym = stack.pop()
            if sym not in table:
                out.append(sym)
                continue
            prods, cum = table[sym]
            i = min(bisect.bisect_right(cum, rnd.random() * cum[-1]), len(prods) - 1)
            stack.extend(reversed(prods[i]))
        if not stack:
            return out


TARGETS = ["X_0", "X_0 * X_0", "sin ( X_0 )", "cos ( X_0 )", "exp ( X_0 )", "( X_0 ) ^2", "X_0 + X_0",
           "X_0 - X_0", "X_0 / X_0"]
N_DIST = 100000


@pytest.fixture(scope="module")
def python_counts():
    rules = _parse(TEST_GRAMMAR)
    rnd = random.Random(0)
    py = defaultdict(int)
    for _ in range(N_DIST):
        s = " ".join(_py_sample(rules, "E", 128, rnd))
        if s in TARGETS:
            py[s] += 1
    return py


def check_distribution(strings, py, label):
    """The target expressions' frequencies within 20% (relative) of the
    Python sampler's (tests/test_sampling.py:233-287)."""
    got = defaultdict(int)
    for s in strings:
        if s in TARGETS:
            got[s] += 1
    for t in TARGETS:
        a, b = py[t] / N_DIST, got[t] / len(strings)
        print(f"{t}: python {a:.5f} {label} {b:.5f}")
        assert abs(a - b) / max(a, b) < 0.2, (t, a, b)


def test_sample_distribution_matches_a_python_sampler(python_counts):
    """tests/test_sampling.py:233-287 on the host sampler (the reference's
    std::mt19937 stream): 100k samples."""
    torch.manual_seed(0)
    g = ProbabilisticContextFreeGrammar(TEST_GRAMMAR, "E", 128, 1, CPU)
    check_distribution(g.to_string(g.sample_infix(N_DIST)), python_counts, "host")


def sample_philox_host(g, B, seed):
    """The device sampler's rows drawn by the host engine
    (sr_pcfg_sample_philox)."""
    from cudasymbolicregression_amd import _lib

    L = g.padded_maximum_length
    out = torch.zeros((B, L), dtype=torch.int64)
    errors = torch.zeros(B, dtype=torch.int64)
    _lib.host().sr_pcfg_sample_philox(
        _lib.ptr(g._rhs_ptr), _lib.ptr(g._rhs_concat), _lib.ptr(g._nt_rule_ptr), _lib.ptr(g._cum_probs),
        g.get_symbol_id(g._start_symbol), L, g._max_tries, _lib.ptr(out), _lib.ptr(errors), seed, B)
    return out, errors


def test_philox_sample_distribution_matches_a_python_sampler(python_counts):
    """The same check on the Philox-4x32-10 stream of the device sampler
    (drawn here by the host engine; the device kernel draws these rows bit
    for bit, tests/test_pcfg_device_gpu.py)."""
    g = ProbabilisticContextFreeGrammar(TEST_GRAMMAR, "E", 128, 1, CPU)
    out, errors = sample_philox_host(g, N_DIST, 0x5EED)
    assert int(errors.abs().sum()) == 0
    check_distribution(g.to_string(out), python_counts, "philox")


# ------------------------------------------------------------ trivial grammars
def test_trivial_grammars():
    g = ProbabilisticContextFreeGrammar("S -> 1 [1.0]", "S", 1, 0, CPU)
    s = g.sample_infix(10)
    assert s.shape == (10, 1) and (s == int(Operator.CONST_1)).all()
    assert all(x == "1" for x in g.to_string(s))
    g = ProbabilisticContextFreeGrammar("S -> C [1.0]", "S", 1, 0, CPU)
    assert (g.sample_infix(10) == int(Operator.LEARNABLE_CONSTANT)).all()
    for i in range(10):
        g = ProbabilisticContextFreeGrammar(f"S -> X_{i} [1.0]", "S", 1, 10, CPU)
        s = g.sample_infix(10)
        assert s.shape == (10, 1) and (s == int(Operator.VAR_START_ID) + i).all()
        assert all(x == f"X_{i}" for x in g.to_string(s))


@pytest.mark.parametrize("op,op_id", [("exp", int(Operator.EXP)), ("sin", int(Operator.SIN)),
                                      ("cos", int(Operator.COS))])
def test_unary_operator_grammar(op, op_id):
    torch.manual_seed(1)
    g = ProbabilisticContextFreeGrammar(f"S -> {op} ( 1 ) [0.5]\nS -> 1 [0.5]", "S", 5, 0, CPU)
    s = g.sample_infix(1000)
    assert s.shape == (1000, 5)
    assert (s == op_id).sum() > 400 and (s == int(Operator.CONST_1)).sum() > 400
    assert all(op in x or x == "1" for x in g.to_string(s))


@pytest.mark.parametrize("sym,op_id", [("+", int(Operator.ADD)), ("-", int(Operator.SUB)), ("*", int(Operator.MUL)),
                                       ("/", int(Operator.DIV)), ("^", int(Operator.POW))])
def test_binary_operator_grammar(sym, op_id):
    torch.manual_seed(2)
    g = ProbabilisticContextFreeGrammar(f"S -> 1 {sym} 1 [0.5]\nS -> 1 [0.5]", "S", 3, 0, CPU)
    s = g.sample_infix(1000)
    assert s.shape == (1000, 3)
    assert torch.any(s == op_id, dim=1).sum() > 400
    assert (s == int(Operator.CONST_1)).sum() > 1400
    assert all(sym in x or x == "1" for x in g.to_string(s))


def test_generation_max_tries_exceeded():
    g = ProbabilisticContextFreeGrammar("S -> 1 S [1.0]", "S", 5, 0, CPU, max_tries=10)
    with pytest.raises(RuntimeError, match="Could not produce a valid expression within the max_tries limit"):
        g.sample_infix(1)
