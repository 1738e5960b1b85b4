"""Build libtomhip.so for gfx950 with hipcc (in-tree, no JIT cache).

    python -m tom_amd.build
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
OUT = os.path.join(HERE, "libtomhip.so")
OBJ = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TOMHIP_ARCH", "gfx950")

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
    "-Wno-unused-variable", "-Wno-unused-result",
    # packed FP32 instructions are disabled for every kernel of this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]
# per-file extra flags: the rotation must round exactly like the reference C
# No automatic multiply-add contraction where results must not depend on how
# work is grouped (FFT butterflies; the float-exact rotation); the hot loops
# spell their fused multiply-adds out.
_NOCONTRACT = ["-ffp-contract=off"]
EXTRA = {"rotate.hip": _NOCONTRACT, "search.hip": _NOCONTRACT, "subtomo.hip": _NOCONTRACT,
         "util.hip": _NOCONTRACT, "picks.hip": _NOCONTRACT, "analysis.hip": _NOCONTRACT}
SOURCES = ["rotate.hip", "search.hip", "subtomo.hip", "picks.hip", "util.hip", "analysis.hip", "unet.hip", "api.hip",
           "em_io.cpp"]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, src + ".o")
    if src.endswith(".cpp"):  # host-only code
        cmd = ["g++", "-O2", "-fPIC", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), "-c",
               os.path.join(CSRC, src), "-o", obj]
    else:
        cmd = [HIPCC, *COMMON, *EXTRA.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs: int = 4) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    shutil.rmtree(OBJ, ignore_errors=True)
    return OUT


if __name__ == "__main__":
    print(build())
    sys.exit(0)
