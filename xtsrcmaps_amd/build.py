"""Build libxtsrcmaps_amd.so (HIP kernels + C ABI + the C++ drop-in layer)
for gfx950 with hipcc, in-tree.

    python -m xtsrcmaps_amd.build
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libxtsrcmaps_amd.so")

ARCH = os.environ.get("XTS_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
CXX = os.environ.get("XTS_CXX", "/opt/rocm/llvm/bin/clang++")

# Packed FP32 ops are disabled for every device pass: the hardware forms of
# v_pk_{add,mul,fma}_f32 are kept out of generated code.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
# Per-file flags.  xts_modelmap.hip: the surrogate network's matrix-core
# results stay in VGPRs (the AGPR form added a v_accvgpr_read per
# accumulator value: 48 per 16 pixels).
FILE_FLAGS = {"xts_modelmap.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}
COMMON = ["-O3", "-fPIC", "-std=c++20", f"--offload-arch={ARCH}",
          "-I", os.path.join(ROOT, "include"), "-I", CSRC,
          "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument"]


def sources() -> list[str]:
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith(".hip") or f.endswith(".cpp"))


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC, *COMMON, *DEVICE_FLAGS, *FILE_FLAGS.get(os.path.basename(src), []), "-c", src,
           "-o", obj]
    if src.endswith(".cpp"):
        # host-only C++23 (the Tensor API uses multi-index operator[])
        cmd = [CXX, "-x", "c++", "-O3", "-fPIC", "-std=c++23", "-I", os.path.join(ROOT, "include"),
               "-I", CSRC, "-I", os.path.join(HERE, "include"), "-Wall", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-soname,libxtsrcmaps_amd.so",
           "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    # C++ drivers of the Fermi:: API (tests/cpp), linked against the library
    for src in sorted(glob.glob(os.path.join(ROOT, "tests", "cpp", "*.cpp"))):
        exe = src[:-4]
        cmd = [CXX, "-O2", "-std=c++23", "-I", os.path.join(ROOT, "include"),
               "-I", os.path.join(HERE, "include"), "-Wall", "-x", "c++", src, "-x", "none",
               LIB, "-Wl,-rpath,$ORIGIN/" + os.path.relpath(HERE, os.path.dirname(exe)),
               "-o", exe + ".tmp"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"driver build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        os.replace(exe + ".tmp", exe)
    if verbose:
        print(f"built {LIB} ({len(objs)} objects, {ARCH})")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
