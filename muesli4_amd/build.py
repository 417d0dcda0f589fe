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
HIP_FLAGS = [
    "-O3",
    "-std=c++17",
    f"--offload-arch={ARCH}",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
]


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


def build_all() -> None:
    build_oracle()
    build_hip()


if __name__ == "__main__":
    build_all()
    print("built", os.path.join(LIBDIR, LIB_NAME))
