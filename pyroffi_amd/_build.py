"""In-tree build of the HIP library ``pyroffi_amd/libpyroffi_amd.so`` for gfx950.

One ``hipcc`` invocation compiles every kernel source and the C-ABI glue into a
single shared object next to this file, so it travels with the source tree.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG_DIR, "csrc")
LIB_PATH = os.path.join(PKG_DIR, "libpyroffi_amd.so")
SOURCES = ("pk_common.cpp", "pk_fk.hip", "pk_collision.hip", "pk_ik.hip", "pk_grad.hip", "pk_grid.hip")
HEADERS = ("pk_common.h", "pk_se3.h", "pk_model.h")

# Packed FP32 instructions are disabled for all device code; hipcc's host pass
# warns that it does not know the feature, which is harmless.
HIPCC_FLAGS = (
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
)


def hipcc_path() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (set HIPCC or install ROCm)")


def build_library(verbose: bool = False) -> str:
    """Compile the library (always rebuilds) and return its path."""
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc_path(), *HIPCC_FLAGS, *[os.path.join(CSRC, s) for s in SOURCES], "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"hipcc failed with exit code {proc.returncode}")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build_library(verbose=True))
