"""Build recipe of libtcloth_hip.so (hipcc for gfx950 + g++ for the host mesh build).

Run as ``python -m taremincloth_amd.build`` or through ``__graft_entry__.build()``.
The library is written in-tree to ``taremincloth_amd/lib/``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIBNAME = "libtcloth_hip.so"
ARCH = os.environ.get("TC_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["tc_sim.hip", "tc_xpbd.hip", "tc_selfcol.hip", "tc_fused.hip", "tc_warmstart.hip", "tc_ws_api.hip", "tc_sdf_bake.hip", "tc_frames.hip", "tc_pins.hip", "tc_jacobi.hip", "tc_edge_collision.hip", "tc_wrinkle.hip", "tc_bone_sdf.hip", "tc_sewing.hip", "tc_dyn_sdf.hip"]
CXX_SOURCES = ["tc_mesh.cpp", "tc_tiles.cpp", "tc_batch.cpp"]

# Strict IEEE fp32 (no FMA contraction) so results follow the reference's
# operation order; packed-fp32 instructions are disabled.
HIP_FLAGS = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-Wall"]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm toolchain required to build libtcloth_hip.so)")


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")


def build(verbose: bool = False) -> str:
    """Compile every source and link the shared library; returns its path."""
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    hipcc = _hipcc()
    objs = []
    for src in CXX_SOURCES:
        obj = os.path.join(objdir, src + ".o")
        _run(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, src + ".o")
        _run([hipcc, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    out = os.path.join(LIBDIR, LIBNAME)
    tmp = out + ".tmp"
    _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp])
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    build(verbose=True)
