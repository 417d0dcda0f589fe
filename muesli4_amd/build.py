"""Build recipes for libmsl_hip.so (gfx950) and the oracle library.

Compiles in-tree with hipcc directly (no JIT cache), so the built files
travel with the source tree.  Run as ``python -m muesli4_amd.build``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "muesli4_amd")
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
OBJDIR = os.path.join(ROOT, "_build", "obj")
ORACLE = os.path.join(ROOT, "oracle")

INCLUDE = os.path.join(ROOT, "include")
API_INCLUDE = os.path.join(PKG, "include")
CPP_TESTS = os.path.join(ROOT, "tests", "cpp")

LIB_NAME = "libmsl_hip.so"
SOURCES = [
    "runtime.hip",
    "elementwise.hip",
    "fold.hip",
    "stencil2d.hip",
    "stencil3d.hip",
    "batch.hip",
    "serve.hip",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# -ffp-contract=off: no FMA contraction (bit-exact with the CPU loops).
# packed FP32 instructions are disabled for all device code.
CODEGEN_FLAGS = [
    "-O3",
    "-std=c++17",
    f"--offload-arch={ARCH}",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
# The strip marches fully unroll one row-sum ring (up to 33 rows at r = 16)
# so that every ring slot is a register; the default size limit of
# `#pragma unroll` gives up on the widest ones and falls back to scratch.
HIP_FLAGS = [*CODEGEN_FLAGS, "-mllvm", "-pragma-unroll-threshold=200000", "-Wall", "-Wno-unused-function",
             "-Wno-unused-variable"]


def _run(cmd, cwd=None):
    res = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + res.stdout)
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(cmd[:3])} ...")
    return res.stdout


def _compile_one(src: str) -> str:
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-c",
           os.path.join(CSRC, src), "-o", obj]
    _run(cmd)
    return obj


def build_hip(jobs: int = 8) -> str:
    """Compile every HIP translation unit for gfx950 and link libmsl_hip.so."""
    os.makedirs(OBJDIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=max(1, min(jobs, len(srcs)))) as ex:
        objs = list(ex.map(_compile_one, srcs))
    out = os.path.join(LIBDIR, LIB_NAME)
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, out)
    return out


def build_oracle() -> str:
    """Compile the oracle (test infrastructure) with make."""
    make = shutil.which("make") or "make"
    _run([make, "-C", ORACLE, "-s"])
    return os.path.join(ORACLE, "liboracle.so")


def program_flags() -> list:
    """hipcc flags for a program written against the msl:: class templates
    (muesli4_amd/include): the same code-generation flags as the library,
    the two include directories and the link line for libmsl_hip.so."""
    return [*CODEGEN_FLAGS, "-x", "hip", "-I", API_INCLUDE, "-I", INCLUDE,
            "-L", LIBDIR, "-lmsl_hip", f"-Wl,-rpath,{LIBDIR}"]


def build_program(src: str, out: str, rpath_origin: str | None = None) -> str:
    """Compile one C++ skeleton program (e.g. an example written against the
    msl:: API) into an executable linked to libmsl_hip.so.  rpath_origin, if
    given, is the library directory relative to the executable ($ORIGIN), so
    the executable keeps working when the tree is moved."""
    os.makedirs(os.path.dirname(out), exist_ok=True)
    flags = program_flags()
    if rpath_origin is not None:
        flags = [f for f in flags if not f.startswith("-Wl,-rpath")] + [f"-Wl,-rpath,$ORIGIN/{rpath_origin}"]
    tmp = out + ".tmp"
    _run([HIPCC, *flags, src, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_cpp_tests() -> str:
    """The C++ API test driver run by tests/test_gpu_cpp_api.py."""
    out = os.path.join(CPP_TESTS, "build", "skeletons_test")
    rel = os.path.relpath(LIBDIR, os.path.dirname(out))
    return build_program(os.path.join(CPP_TESTS, "skeletons_test.cpp"), out, rpath_origin=rel)


def build_cpp_bench() -> str:
    """The C++ skeleton-API benchmark run by bench.py (user functors through
    msl::DA/DM/DC beside the library kernels)."""
    out = os.path.join(PKG, "bench_cpp", "build", "skeleton_bench")
    rel = os.path.relpath(LIBDIR, os.path.dirname(out))
    return build_program(os.path.join(PKG, "bench_cpp", "skeleton_bench.cpp"), out, rpath_origin=rel)


def build_all() -> None:
    build_oracle()
    build_hip()
    build_cpp_tests()
    build_cpp_bench()


if __name__ == "__main__":
    build_all()
    print("built", os.path.join(LIBDIR, LIB_NAME))
