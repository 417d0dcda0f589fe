"""CPU checks of the C++ skeleton API headers (no GPU needed).

- The test driver tests/cpp/skeletons_test.cpp type-checks for gfx950.
- When the reference checkout is present, its hot-path example programs
  compile unchanged against muesli4_amd/include (type-checked; the
  stencil-heavy ones also through gfx950 code generation).  Its sources are
  read in place, never copied.
"""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

from muesli4_amd import build as B

ROOT = B.ROOT
REF_EXAMPLES = "/root/reference/examples"
HOT_PATH_EXAMPLES = [
    "da_test", "ds_test", "dc_test", "dc_test_comp", "dm_test", "dc_map_stencil_test",
    "heatdiffusion", "gameoflife", "jacobisolver", "gaussianblur", "piDA",
    "meanblur2", "meanblur4", "meanblur6", "meanblur8", "meanblur10", "meanblur12",
    "meanblur14", "meanblur16", "meanblur18", "meanblur20",
    # the other examples the reference's own build compiles (CMakeLists.txt:301-317)
    "aco_tsp", "aco_bpp", "gaussianblur-color", "minnested",
]

needs_hipcc = pytest.mark.skipif(not os.path.exists(B.HIPCC) and shutil.which("hipcc") is None,
                                 reason="hipcc not installed")


# stencil-heavy examples get a full device compile (code generation can
# fail where type-checking passes); the rest are type-checked
DEVICE_COMPILED = ["gameoflife", "heatdiffusion", "jacobisolver", "gaussianblur", "meanblur2", "dc_map_stencil_test",
                   "aco_tsp", "aco_bpp"]


def _syntax(src, codegen=False):
    cmd = [B.HIPCC, "-x", "hip", "-std=c++17", f"--offload-arch={B.ARCH}", "-I", B.API_INCLUDE, "-I", B.INCLUDE, src]
    if codegen:
        cmd += ["-O3", "-ffp-contract=off", "--offload-device-only", "-c", "-o", os.devnull]
    else:
        cmd += ["-fsyntax-only"]
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)
    return res.returncode, res.stdout


@needs_hipcc
def test_cpp_driver_typechecks():
    rc, out = _syntax(os.path.join(ROOT, "tests", "cpp", "skeletons_test.cpp"))
    assert rc == 0, out


@needs_hipcc
@pytest.mark.skipif(not os.path.isdir(REF_EXAMPLES), reason="reference checkout not present")
def test_reference_examples_compile_unchanged():
    srcs = [os.path.join(REF_EXAMPLES, e + ".cpp") for e in HOT_PATH_EXAMPLES]
    with ThreadPoolExecutor(max_workers=6) as ex:
        results = list(ex.map(lambda s: _syntax(s, os.path.basename(s)[:-4] in DEVICE_COMPILED), srcs))
    failed = [(s, out[-2000:]) for s, (rc, out) in zip(srcs, results) if rc != 0]
    assert not failed, failed
