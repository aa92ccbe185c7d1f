"""The Philox-4x32-10 generator of the device sampler (csrc/sr_philox.h),
through its host export sr_philox4x32_10 and the host sampler
sr_pcfg_sample_philox (CPU).

* The three philox4x32 R=10 known-answer vectors of the Random123 suite
  (Salmon et al., SC'11): counter/key all zeros, all ones, and the pi
  digits.
* The stream layout (key = seed, counter = (block, row), words in order)
  against PyTorch's own Philox engine, at::philox_engine(seed, row, 0),
  compiled from its header into a small helper (tests/philox_torch.cpp).
* Uniform floats are (w >> 8) * 2^-24: 24 random bits, in [0, 1).
"""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _lib():
    from cudasymbolicregression_amd import _lib

    return _lib.host()


def _block(ctr, key):
    c = np.array(ctr, np.uint32)
    k = np.array(key, np.uint32)
    out = np.zeros(4, np.uint32)
    _lib().sr_philox4x32_10(c.ctypes.data, k.ctypes.data, out.ctypes.data)
    return [int(v) for v in out]


KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,expected", KAT)
def test_philox_known_answers(ctr, key, expected):
    assert _block(ctr, key) == expected


def _stream(seed, row, nwords):
    words = []
    n = 0
    while len(words) < nwords:
        words += _block([n & 0xFFFFFFFF, n >> 32, row & 0xFFFFFFFF, row >> 32], [seed & 0xFFFFFFFF, seed >> 32])
        n += 1
    return words[:nwords]


def test_philox_streams_match_torch_engine(tmp_path):
    import torch.utils.cpp_extension as ce

    exe = str(tmp_path / "philox_torch")
    inc = [f"-I{p}" for p in ce.include_paths()]
    subprocess.run(["g++", "-O1", "-std=c++17", *inc, os.path.join(HERE, "philox_torch.cpp"), "-o", exe],
                   check=True)
    cases = [(0, 0, 12), (1, 0, 9), (123456789, 7, 16), (2**63 - 25, 2**40 + 3, 8), (2**64 - 1, 2**32 - 1, 12)]
    args = [str(v) for c in cases for v in c]
    lines = subprocess.run([exe, *args], check=True, capture_output=True, text=True).stdout.strip().split("\n")
    assert len(lines) == len(cases)
    for (seed, row, n), line in zip(cases, lines):
        assert [int(w, 16) for w in line.split()] == _stream(seed, row, n), (seed, row)
