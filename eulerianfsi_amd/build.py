"""Build libfsi_hip.so (gfx950) in-tree with hipcc.

Usage: python -m eulerianfsi_amd.build   (or build() from __graft_entry__)
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
SOURCES = ["fsi_api.hip", "fsi_physics.hip", "fsi_solvers.hip", "fsi_mg.hip", "fsi_spectral.hip", "fsi_adjoint.hip",
           "fsi_adjoint_api.hip", "fsi_growth.hip"]
LIB = PKG / "libfsi_hip.so"
ARCH = os.environ.get("FSI_OFFLOAD_ARCH", "gfx950")

# -packed-fp32-ops: no v_pk_*_f32 generation (some of their forms have given
# wrong results on gfx950); harmless for this FP64 code, kept for the FP32 paths.
HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (Path(cand).exists() or cand == "hipcc"):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = True) -> Path:
    """Compile every HIP source into one shared library next to this file."""
    objs = []
    for src in SOURCES:
        obj = CSRC / (Path(src).stem + ".o")
        cmd = [hipcc(), *[f for f in HIPCC_FLAGS if f != "-shared"], "-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(str(obj))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.unlink(o)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
