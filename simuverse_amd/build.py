"""Builds the gfx950 shared library of the D2Q9 hot path with hipcc, and the
C++ host layer over its C ABI (D2Q9Node / FluidSimulator) with g++.

Both are written in-tree to ``simuverse_amd/_lib/`` (libsimuverse_lbm.so,
libsimuverse_host.so) so that they travel with the source tree to the
machine that runs them.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "_lib"
LIB = LIBDIR / "libsimuverse_lbm.so"
HOST_LIB = LIBDIR / "libsimuverse_host.so"
HOST_SOURCES = ["host/d2q9_node.cpp"]

SOURCES = ["lbm_kernels.hip", "lbm_step2.hip", "lbm_step3.hip", "lbm_learned.hip", "lbm_api.hip", "lbm_batch_api.hip",
           "lbm_mlp.hip", "lbm3d.hip"]

# -ffp-contract=off: the parity kernel must round exactly where the
# reference's WGSL rounds, and the physics kernels stay reproducible
# across tile shapes.  The packed-fp32 feature is switched off for all
# device code.
HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]
# Per-file additions.  The learned step is bound by vector instructions
# around its matrix instructions; with the accumulators in ordinary VGPRs
# (rather than AGPRs) it needs no copies between the two register files.
EXTRA_FLAGS = {
    "lbm_learned.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"],
}


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the gfx950 library cannot be built")


def build(verbose: bool = False) -> Path:
    """Compile every HIP source and link the shared library; returns its path."""
    LIBDIR.mkdir(exist_ok=True)
    objdir = LIBDIR / "obj"
    objdir.mkdir(exist_ok=True)
    cc = hipcc()
    objs, procs = [], []
    # the translation units compile in parallel (one hipcc each)
    for src in SOURCES:
        obj = objdir / (Path(src).stem + ".o")
        cmd = [cc, *HIPCC_FLAGS, *EXTRA_FLAGS.get(src, []), "-I", str(ROOT / "include"), "-I", str(CSRC), "-c",
               str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd)))
        objs.append(str(obj))
    failed = [src for src, p in procs if p.wait() != 0]
    if failed:
        raise RuntimeError(f"hipcc failed on {', '.join(failed)}")
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_host(verbose)
    return LIB


def build_host(verbose: bool = False) -> Path:
    """The C++ host layer, linked against the HIP library (rpath $ORIGIN)."""
    tmp = HOST_LIB.with_suffix(".so.tmp")
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", "-Wextra", "-I", str(ROOT / "include"),
           *[str(PKG / s) for s in HOST_SOURCES], "-L", str(LIBDIR), "-lsimuverse_lbm", "-Wl,-rpath,$ORIGIN",
           "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, HOST_LIB)
    return HOST_LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
