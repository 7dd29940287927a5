"""In-tree build of the HIP library ``pyroffi_amd/libpyroffi_amd.so`` for gfx950.

Every kernel source and the C-ABI glue is compiled by its own ``hipcc`` process
(in parallel, objects under ``_build/``) and linked into a single shared object
next to this file, so it travels with the source tree.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG_DIR, "csrc")
LIB_PATH = os.path.join(PKG_DIR, "libpyroffi_amd.so")
SOURCES = ("pk_common.cpp", "pk_fk.hip", "pk_collision.hip", "pk_distances.hip", "pk_ik.hip", "pk_ik_jax.hip",
           "pk_grad.hip", "pk_grid.hip", "pk_traj.hip", "pk_traj_noalloc.hip")
# pk_ik_inst.hip is compiled once per (padded DOF, padded EE count) pair.
IK_INSTANCES = tuple((m, e) for m in (8, 16) for e in (1, 2, 4))


def _units():
    """(source, extra flags, object name) of every translation unit."""
    units = [(s, (), os.path.splitext(s)[0] + ".o") for s in SOURCES]
    units += [(src, (f"-DPK_IK_MAXA={m}", f"-DPK_IK_NEE={e}"), f"{stem}_{m}_{e}.o")
              for src, stem in (("pk_ik_inst.hip", "pk_ik"), ("pk_ik_jax_inst.hip", "pk_ik_jax"))
              for m, e in IK_INSTANCES]
    return units
HEADERS = ("pk_common.h", "pk_se3.h", "pk_model.h", "pk_geom.h", "pk_grid.h", "pk_ik.h", "pk_ik_jax.h")

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


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"hipcc failed with exit code {proc.returncode}: {' '.join(cmd[-3:])}")


def build_library(verbose: bool = False, jobs: int | None = None) -> str:
    """Compile the library (always rebuilds every source, in parallel) and return its path."""
    from concurrent.futures import ThreadPoolExecutor

    hipcc = hipcc_path()
    obj_dir = os.path.join(PKG_DIR, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    compile_flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    units = _units()
    objs = [os.path.join(obj_dir, o) for _, _, o in units]
    cmds = [[hipcc, *compile_flags, *defs, "-c", os.path.join(CSRC, s), "-o", o]
            for (s, defs, _), o in zip(units, objs)]
    workers = jobs or min(len(cmds), max(1, (os.cpu_count() or 4)))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        for f in [ex.submit(_run, c, verbose) for c in cmds]:
            f.result()
    tmp = LIB_PATH + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp], verbose)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build_library(verbose=True))
