"""The C-ABI boundary: every function declared in include/*.h is exported by
the built libraries, the headers compile as C and as C++, and the Python
loader declared every symbol it binds (no compute calls: runs without a GPU)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = {"sr_eval.h": "libsr_hip.so", "sr_mlp.h": "libsr_hip.so", "sr_pcfg.h": "libsr_host.so"}


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
    cfg = autotune.get_config()
    assert len(v) == len(cfg) >= 1
    for b, row in enumerate(v):
        assert 0 <= cfg[b] < len(row)
        for d in row:
            assert d["waves"] >= 1 and d["ppt"] >= 1
