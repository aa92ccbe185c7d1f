"""In-tree build of the native libraries (gfx950 HIP kernels + host engine).

    python -m cudasymbolicregression_amd.build

libsr_hip.so  <- csrc/*.hip   (hipcc --offload-arch=gfx950, one object per
                              source, compiled in parallel, linked together)
libsr_host.so <- csrc/pcfg_host.cpp (g++ -fopenmp)
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-I" + INCLUDE,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # a wave-uniform atomic (the work-queue fetch) reduces its lanes with DPP
    # row operations instead of a 64-iteration scalar loop over the lanes
    "-mllvm", "-amdgpu-atomic-optimizer-strategy=DPP",
    "-Wno-unused-command-line-argument",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build_hip(jobs: int = 8) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    objs = [os.path.join(OBJ, os.path.basename(s) + ".o") for s in srcs]
    with ThreadPoolExecutor(max_workers=max(1, min(jobs, len(srcs)))) as ex:
        list(ex.map(lambda so: _run([HIPCC, *HIP_FLAGS, "-c", so[0], "-o", so[1]]), zip(srcs, objs)))
    out = os.path.join(HERE, "libsr_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, out)
    return out


def build_host() -> str:
    out = os.path.join(HERE, "libsr_host.so")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-o", tmp, os.path.join(CSRC, "pcfg_host.cpp")])
    os.replace(tmp, out)
    return out


def build_all() -> None:
    with ThreadPoolExecutor(max_workers=2) as ex:
        f1 = ex.submit(build_hip)
        f2 = ex.submit(build_host)
        f1.result()
        f2.result()


if __name__ == "__main__":
    build_all()
    print("built", file=sys.stderr)
