"""Build recipe for the gfx950 library (``libsynchrad_amd.so``).

One ``hipcc`` invocation, in-tree, so the shared object travels with the
source tree.  ``-ffp-contract=off`` keeps the FP64 expression order of the
kernels exact; packed FP32 instructions are disabled for the whole build.
"""
from __future__ import annotations

import os
import shutil
import subprocess

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
SRCS = [os.path.join(PKG_DIR, "csrc", f) for f in ("spectrum.hip", "adjoint.hip")]
INCLUDE = os.path.join(REPO_DIR, "include")
LIB = os.path.join(PKG_DIR, "libsynchrad_amd.so")
ARCH = os.environ.get("SYNCHRAD_AMD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build synchrad_amd)")


def command(out: str = LIB) -> list:
    return [
        hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", INCLUDE, *SRCS, "-o", out,
    ]


def build(verbose: bool = False) -> str:
    """Compile the library; returns its path.  Writes to a temporary name and
    renames, so a process that already loaded the old file keeps working."""
    tmp = LIB + ".tmp"
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
