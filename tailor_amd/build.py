"""Build libtailor.so (HIP for gfx950 + C++20 host code) in-tree with hipcc.

Every translation unit is compiled for gfx950 only, with packed FP32
instructions disabled.  assembly.hip is compiled without floating-point
contraction so its sums round exactly like a sequential CPU evaluation.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libtailor.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-std=c++20", "-fPIC", "--offload-arch=gfx950",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function", "-Werror=uninitialized",
          f"-I{PKG.parent / 'include'}"]
PER_FILE = {
    "assembly.hip": ["-ffp-contract=off"],
    "lu.hip": ["-munsafe-fp-atomics"],
    # keep MFMA accumulators in VGPRs: with the default AGPR form the
    # calibration loop copies every accumulator to and from AGPRs around
    # each round of MFMAs and measures those copies, not the MFMA rate
    "calibrate.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
}
SOURCES = ["assembly.hip", "lu.hip", "vec.hip", "capi.cpp", "eigen.cpp", "symbolic.cpp",
           "lu_plan.cpp", "dense.cpp", "mlp.hip", "noalloc.hip", "batch.cpp", "calibrate.hip"]


def _compile(src: str) -> Path:
    obj = OBJ / (src + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") or src.endswith(".cpp") else []
    cmd = [HIPCC, *lang, *COMMON, *PER_FILE.get(src, []), "-c", str(CSRC / src), "-o", str(obj)]
    subprocess.run(cmd, check=True)
    return obj


def build(verbose: bool = False) -> Path:
    OBJ.mkdir(exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp),
                    *map(str, objs)], check=True)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
