"""Builds librayna_amd.so (the gfx950 renderer behind include/rayna_gpu.h) in-tree.

    python -m rayna_amd.build            # product library
    python -m rayna_amd.build --oracle   # also the oracle (make -C oracle)
    python -m rayna_amd.build --cpp-tests  # also the C++ API tests (tests/cpp)

Device code is compiled by hipcc for gfx950 only, without floating-point contraction, with correctly rounded
single-precision division and square root, and with packed-FP32 instructions disabled.  Host code is compiled
with the same floating-point rules.
"""
from __future__ import annotations

import argparse
import os
import shlex
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librayna_amd.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

FP_FLAGS = ["-ffp-contract=off", "-fno-fast-math"]
DEVICE_FLAGS = ["--offload-arch=gfx950", "-fhip-fp32-correctly-rounded-divide-sqrt", "-Xclang", "-target-feature",
                "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function"] + FP_FLAGS


def _run(cmd):
    print("+", " ".join(shlex.quote(c) for c in cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_product():
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    k = os.path.join(BUILD, "kernels.o")
    _run([HIPCC] + COMMON + DEVICE_FLAGS + ["-c", os.path.join(CSRC, "kernels.hip"), "-o", k])
    objs.append(k)
    for src in ("renderer.cpp", "batch.cpp", "scene_build.cpp"):
        o = os.path.join(BUILD, src.replace(".cpp", ".o"))
        _run(["g++"] + COMMON + ["-pthread", "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include", "-c",
                                 os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-pthread"] + objs + ["-o", tmp])
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    _run(["make", "-C", os.path.join(ROOT, "oracle")])


CPP_TESTS = os.path.join(ROOT, "tests", "cpp")


def build_cpp_tests():
    """The C++ API's tests (tests/cpp/*.cpp, against include/rayna/engine.hpp), linked to the in-tree library."""
    out_dir = os.path.join(CPP_TESTS, "_build")
    os.makedirs(out_dir, exist_ok=True)
    for name in sorted(os.listdir(CPP_TESTS)):
        if not name.endswith(".cpp"):
            continue
        exe = os.path.join(out_dir, name[:-4])
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", f"-I{os.path.join(ROOT, 'include')}",
              os.path.join(CPP_TESTS, name), f"-L{HERE}", "-lrayna_amd", f"-Wl,-rpath,{HERE}",
              "-Wl,-rpath,$ORIGIN/../../../rayna_amd", "-o", exe])


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--oracle", action="store_true", help="also build the oracle (CPU restatement)")
    ap.add_argument("--cpp-tests", action="store_true", help="also build the C++ API tests (tests/cpp)")
    a = ap.parse_args(argv)
    build_product()
    if a.oracle:
        build_oracle()
    if a.cpp_tests:
        build_cpp_tests()


if __name__ == "__main__":
    sys.exit(main())
