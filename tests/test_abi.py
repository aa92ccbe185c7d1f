"""The C-ABI boundary: every function declared in include/*.h is exported by
the built libraries, the headers compile as C and as C++, and the Python
loader declared every symbol it binds (no compute calls: runs without a GPU)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = {"sr_eval.h": "libsr_hip.so", "sr_mlp.h": "libsr_hip.so", "sr_pcfg.h": "libsr_host.so",
           "sr_pcfg_device.h": "libsr_hip.so"}


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sr_[a-z0-9_]+)\s*\(", text)))


def _exported(lib):
    path = os.path.join(ROOT, "cudasymbolicregression_amd", lib)
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


@pytest.mark.parametrize("header", sorted(HEADERS))
def test_header_symbols_are_exported(header):
    names = _declared(header)
    assert names, header
    missing = [n for n in names if n not in _exported(HEADERS[header])]
    assert not missing, f"{header}: declared but not exported: {missing}"


@pytest.mark.parametrize("lang", ["c", "c++"])
@pytest.mark.parametrize("header", sorted(HEADERS))
def test_header_compiles(header, lang):
    src = f'#include "{header}"\nint main(void) {{ return 0; }}\n'
    r = subprocess.run(["gcc" if lang == "c" else "g++", "-x", lang, "-fsyntax-only", "-Wall", "-Werror",
                        "-I", os.path.join(ROOT, "include"), "-"], input=src, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_python_loader_binds_every_symbol():
    from cudasymbolicregression_amd import _lib

    _lib.hip()
    _lib.host()
    assert _lib.MISSING_SYMBOLS == []
    exported = _exported("libsr_hip.so") | _exported("libsr_host.so")
    declared = set().union(*(_declared(h) for h in HEADERS))
    # every sr_* function the libraries export is part of the documented boundary
    assert {n for n in exported if n.startswith("sr_")} <= declared


def test_fit_variant_table_is_consistent():
    from cudasymbolicregression_amd import autotune

    v = autotune.variants()
    cfg = autotune.default_config()
    assert len(v) == len(cfg) == autotune.N_BUCKETS
    for b, row in enumerate(v):
        assert 0 <= cfg[b] < len(row)
        for d in row:
            assert d["waves"] >= 1 and d["ppt"] >= 1
    assert autotune.validate_config(cfg) == cfg
    with pytest.raises(ValueError):
        autotune.validate_config([len(row) for row in v])


def test_autotune_cache_is_keyed_and_persists(tmp_path):
    """The workload cache: entries are found by their key only, survive a
    save/load round trip, and bad configurations are refused."""
    import torch

    from cudasymbolicregression_amd import autotune

    autotune.cache_clear()
    dev = torch.device("cpu")
    cfg = [0] * autotune.N_BUCKETS
    assert autotune.lookup(dev, 64, 512, 3) is None
    autotune.cache_put(autotune.workload_key(dev, 64, 512, 3), cfg, 1.5)
    assert autotune.lookup(dev, 64, 512, 3) == cfg
    assert autotune.lookup(dev, 64, 300, 3) == cfg  # same power-of-two point bucket
    assert autotune.lookup(dev, 64, 513, 3) is None
    assert autotune.lookup(dev, 32, 512, 3) is None
    path = tmp_path / "tune.json"
    autotune.cache_save(str(path))
    autotune.cache_clear()
    assert autotune.lookup(dev, 64, 512, 3) is None
    assert autotune.cache_load(str(path)) == 1
    assert autotune.lookup(dev, 64, 512, 3) == cfg
    autotune.cache_clear()
    with pytest.raises(ValueError):
        autotune.cache_put(autotune.workload_key(dev, 64, 512, 3), [99] * autotune.N_BUCKETS)
    with pytest.raises(ValueError):
        autotune.make_options([0, 0])
