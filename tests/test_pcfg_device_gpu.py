"""The grammar kernels on the GPU (include/sr_pcfg_device.h) against the
host engine (pcfg_host.cpp, bit-exact with the reference's
src/pcfg_cpu.cpp on these functions, tests/test_pcfg.py).

* Sampling: the device rows equal, bit for bit, the rows the host engine
  draws from the same Philox-4x32-10 streams (sr_pcfg_sample_philox, whose
  generator and distribution tests/test_philox.py and
  tests/test_pcfg_grammar.py pin), on three grammars and at the padded
  lengths 16, 64 and 128; the max_tries failure gives code 303 on every row.
* infix -> postfix, the children table and postfix -> infix: the device
  tokens and error codes equal the host engine's on sampled rows and on
  fuzzed (mostly malformed) rows, as the host engine's are checked against
  the reference's (tests/test_pcfg.py:310-380), including rows whose
  rendering does not fit max_infix_len.
* The Python methods dispatch on the input's device: a GPU grammar samples
  on the device, and errors raise the host path's texts.
"""
import numpy as np
import pytest
import torch

from cudasymbolicregression_amd import ProbabilisticContextFreeGrammar, _lib
from test_pcfg import G2, GRAMMAR_A, GRAMMAR_INFIX, _fuzz_rows
from test_pcfg_grammar import README_GRAMMAR, sample_philox_host

pytestmark = pytest.mark.gpu
CPU = torch.device("cpu")
DEV = torch.device("cuda:0")


def _dev_sample(g, B, seed):
    t = g._tables_on(DEV)
    L = g.padded_maximum_length
    out = torch.empty((B, L), dtype=torch.int64, device=DEV)
    errors = torch.empty(B, dtype=torch.int64, device=DEV)
    _lib.check_status(_lib.hip().sr_pcfg_sample_device(
        _lib.ptr(t["rhs_ptr"]), _lib.ptr(t["rhs_concat"]), _lib.ptr(t["nt_rule_ptr"]), _lib.ptr(t["cum"]),
        g._nt_rule_ptr.numel() - 1, g.get_symbol_id(g._start_symbol), L, g._max_tries, _lib.ptr(out),
        _lib.ptr(errors), seed, B, _lib.stream_of(DEV)), "sample")
    torch.cuda.synchronize()
    return out.cpu(), errors.cpu()


@pytest.mark.parametrize("name,grammar,start,n_x,L", [("readme", README_GRAMMAR, "E", 1, 32),
                                                     ("test", GRAMMAR_A, "E", 1, 128),
                                                     ("g2", G2, "E", 3, 64), ("infix", GRAMMAR_INFIX, "E", 2, 16)])
def test_device_sample_equals_host_philox(name, grammar, start, n_x, L):
    print(f"case sample {name} L={L}", flush=True)
    g = ProbabilisticContextFreeGrammar(grammar, start, L, n_x, CPU)
    B = 20000
    for seed in (1, 0xDEADBEEFCAFE1234):
        d_out, d_err = _dev_sample(g, B, seed)
        h_out, h_err = sample_philox_host(g, B, seed)
        assert torch.equal(d_err, h_err), name
        ok = h_err == 0
        assert torch.equal(d_out[ok], h_out[ok]), name
        assert (d_out[~ok] == 0).all()
        print(f"  seed {seed:#x}: {int(ok.sum())} rows, {int((h_err != 0).sum())} failed "
              f"{sorted(set(h_err[h_err != 0].tolist()))}", flush=True)


def test_device_sample_max_tries_error():
    print("case sample max_tries", flush=True)
    g = ProbabilisticContextFreeGrammar("S -> 1 S [1.0]", "S", 5, 0, CPU, max_tries=10)
    d_out, d_err = _dev_sample(g, 100, 7)
    assert (d_err == 303).all() and (d_out == 0).all()
    gd = ProbabilisticContextFreeGrammar("S -> 1 S [1.0]", "S", 5, 0, DEV, max_tries=10)
    with pytest.raises(RuntimeError) as e_dev:
        gd.sample_infix(10)
    with pytest.raises(RuntimeError) as e_host:
        g.sample_infix(10)
    assert str(e_dev.value) == str(e_host.value)


def _host_call(fn, *args):
    getattr(_lib.host(), fn)(*args)


def _p(a):
    return a.ctypes.data


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def test_device_parsing_matches_host_engine():
    gr = ProbabilisticContextFreeGrammar(GRAMMAR_INFIX, "E", 64, 2, CPU)
    rng = np.random.default_rng(5)
    lp, rp = gr.get_symbol_id("("), gr.get_symbol_id(")")
    vocab = np.array([1, 2, 6, 7, 10, 13, 14, 16, 27, 28, 29, 40, 41, 42, 43, 44, 50, 51, lp, rp, lp, rp, 50, 1])
    prec = gr._precedence.numpy()
    prec_d = _dev(prec)
    torch.manual_seed(3)
    good = gr.sample_infix(3000).numpy()
    for M, rows in ((5, _fuzz_rows(rng, 4000, 5, vocab)), (12, _fuzz_rows(rng, 4000, 12, vocab)),
                    (40, _fuzz_rows(rng, 4000, 40, vocab)), (64, np.ascontiguousarray(good)),
                    (128, _fuzz_rows(rng, 2000, 128, vocab))):
        print(f"case infix_to_postfix M={M}", flush=True)
        B = len(rows)
        h_out = np.zeros_like(rows)
        h_err = np.zeros(B, np.int64)
        _host_call("sr_pcfg_infix_to_postfix", _p(prec), _p(rows), _p(h_out), _p(h_err), lp, rp, B, M)
        d_in = _dev(rows)
        d_out = torch.empty_like(d_in)
        d_err = torch.empty(B, dtype=torch.int64, device=DEV)
        _lib.check_status(_lib.hip().sr_pcfg_infix_to_postfix_device(
            _lib.ptr(prec_d), prec_d.numel(), _lib.ptr(d_in), _lib.ptr(d_out), _lib.ptr(d_err), lp, rp, B, M,
            _lib.stream_of(DEV)), "infix_to_postfix")
        d_out, d_err = d_out.cpu().numpy(), d_err.cpu().numpy()
        np.testing.assert_array_equal(d_err, h_err)
        ok = h_err == 0
        np.testing.assert_array_equal(d_out[ok], h_out[ok])
        assert (d_out[~ok] == 0).all()
        print(f"  {int(ok.sum())} of {B} parse; codes {sorted(set(h_err[~ok].tolist()))}", flush=True)


def test_device_children_and_infix_match_host_engine():
    gr = ProbabilisticContextFreeGrammar(GRAMMAR_INFIX, "E", 64, 2, CPU)
    rng = np.random.default_rng(9)
    lp, rp = gr.get_symbol_id("("), gr.get_symbol_id(")")
    torch.manual_seed(2)
    good = gr.sample(3000)[0].numpy()
    vocab = np.array([1, 2, 5, 10, 14, 15, 28, 29, 40, 41, 42, 44, 50, 51])
    sets = [np.ascontiguousarray(np.concatenate([good, _fuzz_rows(rng, 3000, good.shape[1], vocab)])),
            _fuzz_rows(rng, 2000, 8, vocab), _fuzz_rows(rng, 2000, 128, vocab)]
    for rows in sets:
        B, M = rows.shape
        print(f"case children M={M}", flush=True)
        h_out = np.full((B, M, 2), -1, np.int64)
        h_err = np.zeros(B, np.int64)
        _host_call("sr_pcfg_get_postfix_children", _p(rows), _p(h_out), _p(h_err), B, M)
        d_in = _dev(rows)
        d_out = torch.empty((B, M, 2), dtype=torch.int64, device=DEV)
        d_err = torch.empty(B, dtype=torch.int64, device=DEV)
        _lib.check_status(_lib.hip().sr_pcfg_get_postfix_children_device(
            _lib.ptr(d_in), _lib.ptr(d_out), _lib.ptr(d_err), B, M, _lib.stream_of(DEV)), "children")
        d_o, d_e = d_out.cpu().numpy(), d_err.cpu().numpy()
        np.testing.assert_array_equal(d_e, h_err)
        ok = h_err == 0
        np.testing.assert_array_equal(d_o[ok], h_out[ok])
        print(f"  {int(ok.sum())} of {B} valid", flush=True)
        for mi in (M, 3 * M // 2, 200, 3 * M + 5):
            print(f"case postfix_to_infix M={M} Mi={mi}", flush=True)
            h_out = np.zeros((B, mi), np.int64)
            h_err = np.zeros(B, np.int64)
            _host_call("sr_pcfg_postfix_to_infix", _p(rows), _p(h_out), _p(h_err), lp, rp, B, M, mi)
            d_out = torch.empty((B, mi), dtype=torch.int64, device=DEV)
            d_err = torch.empty(B, dtype=torch.int64, device=DEV)
            _lib.check_status(_lib.hip().sr_pcfg_postfix_to_infix_device(
                _lib.ptr(d_in), _lib.ptr(d_out), _lib.ptr(d_err), lp, rp, B, M, mi, _lib.stream_of(DEV)), "infix")
            d_o, d_e = d_out.cpu().numpy(), d_err.cpu().numpy()
            np.testing.assert_array_equal(d_e, h_err)
            ok = h_err == 0
            np.testing.assert_array_equal(d_o[ok], h_out[ok])
            assert (d_o[~ok] == 0).all()
            print(f"  {int(ok.sum())} of {B} render; codes {sorted(set(h_err[~ok].tolist()))}", flush=True)


def test_gpu_grammar_methods_dispatch_to_the_device():
    print("case gpu grammar methods", flush=True)
    gd = ProbabilisticContextFreeGrammar(G2, "E", 64, 3, DEV)
    gh = ProbabilisticContextFreeGrammar(G2, "E", 64, 3, CPU)
    torch.manual_seed(11)
    ops, ch = gd.sample(5000)
    assert ops.device.type == "cuda" and ch.device.type == "cuda"
    # the host engine on the same rows
    ops_h = ops.cpu()
    assert torch.equal(gh.get_postfix_children(ops_h), ch.cpu())
    assert torch.equal(gh.postfix_to_infix(ops_h, 150), gd.postfix_to_infix(ops, 150).cpu())
    inf = gd.sample_infix(3000)
    assert inf.device.type == "cuda"
    assert torch.equal(gd.infix_to_postfix(inf).cpu(), gh.infix_to_postfix(inf.cpu()))
    # error texts of the host path
    bad = torch.zeros((3, 6), dtype=torch.int64)
    bad[:, 0] = 40  # a lone '+'
    for verbosity in (0, 1):
        with pytest.raises(RuntimeError) as e_dev:
            gd.get_postfix_children(bad.to(DEV), verbosity=verbosity)
        with pytest.raises(RuntimeError) as e_host:
            gh.get_postfix_children(bad, verbosity=verbosity)
        assert str(e_dev.value) == str(e_host.value)
        with pytest.raises(RuntimeError) as e_dev:
            gd.postfix_to_infix(bad.to(DEV), 20, verbosity=verbosity)
        with pytest.raises(RuntimeError) as e_host:
            gh.postfix_to_infix(bad, 20, verbosity=verbosity)
        assert str(e_dev.value) == str(e_host.value)
    # the reference's CPU stream (std::mt19937 per row, seeds from the
    # grammar device's generator) stays available on a GPU grammar
    torch.manual_seed(4)
    a = gd.sample_infix(100, engine="host")
    torch.manual_seed(4)
    seeds = torch.randint(0, 9223372036854775807, (100,), dtype=torch.int64, device=DEV).cpu()
    b = torch.zeros((100, 64), dtype=torch.int64)
    err = torch.zeros(100, dtype=torch.int64)
    _lib.host().sr_pcfg_sample(_lib.ptr(gh._rule_lhs), _lib.ptr(gh._rhs_ptr), _lib.ptr(gh._rhs_concat),
                               _lib.ptr(gh._nt_rule_ptr), _lib.ptr(gh._cum_probs), gh.get_symbol_id("E"), 64,
                               gh._max_tries, _lib.ptr(b), _lib.ptr(err), _lib.ptr(seeds), 100)
    assert a.device.type == "cuda" and torch.equal(a.cpu(), b) and int(err.abs().sum()) == 0
