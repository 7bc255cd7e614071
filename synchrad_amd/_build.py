"""Build recipe for the gfx950 library (``libsynchrad_amd.so``).

One ``hipcc`` invocation, in-tree, so the shared object travels with the
source tree.  ``-ffp-contract=off`` keeps the FP64 expression order of the
kernels exact.  ``-Xclang -target-feature -Xclang -packed-fp32-ops`` turns off
the packed FP32 instructions (v_pk_fma_f32, v_pk_mul_f32, v_pk_add_f32) in
the device compilation: without it hipcc -O3 forms them from ordinary float
code on gfx950; with it the device code has none (``device_asm`` below lets
the tests check the built source).  The host pass of the same command does
not know the feature and prints "'-packed-fp32-ops' is not a recognized
feature for this target (ignoring feature)"; that warning concerns the x86
host code only.

``-Rpass-analysis=kernel-resource-usage`` makes the device compiler report
each kernel's registers, scratch (spill) bytes per lane and occupancy; the
report of the last build is kept next to the library (``RESOURCE_LOG``).
"""
from __future__ import annotations

import os
import re
import shutil
import subprocess
from typing import Dict, List

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
SRCS = [os.path.join(PKG_DIR, "csrc", f) for f in ("spectrum.hip", "adjoint.hip")]
INCLUDE = os.path.join(REPO_DIR, "include")
LIB = os.path.join(PKG_DIR, "libsynchrad_amd.so")
RESOURCE_LOG = os.path.join(PKG_DIR, "libsynchrad_amd.resources.txt")
ARCH = os.environ.get("SYNCHRAD_AMD_ARCH", "gfx950")

DEVICE_FLAGS = ["-O3", "-std=c++17", "-ffp-contract=off",
                "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build synchrad_amd)")


def command(out: str = LIB) -> list:
    return [hipcc(), f"--offload-arch={ARCH}", *DEVICE_FLAGS, "-fPIC", "-shared",
            "-Rpass-analysis=kernel-resource-usage", "-I", INCLUDE, *SRCS, "-o", out]


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
    with open(RESOURCE_LOG, "w") as f:
        f.write(res.stderr)
    return LIB


def device_asm(src: str) -> str:
    """gfx950 assembly of one source file, compiled with the library's flags."""
    res = subprocess.run([hipcc(), f"--offload-arch={ARCH}", *DEVICE_FLAGS, "--cuda-device-only", "-S",
                          "-I", INCLUDE, src, "-o", "-"], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc -S failed ({res.returncode}):\n{res.stderr}")
    return res.stdout


def kernel_resources(log: str) -> Dict[str, Dict[str, int]]:
    """{kernel: {"vgpr": .., "scratch": .., "occupancy": ..}} from a
    kernel-resource-usage report."""
    out: Dict[str, Dict[str, int]] = {}
    name = None
    for line in log.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
            out[name] = {}
            continue
        for key, pat in (("vgpr", r" VGPRs: (\d+)"), ("scratch", r"ScratchSize \[bytes/lane\]: (\d+)"),
                         ("occupancy", r"Occupancy \[waves/SIMD\]: (\d+)")):
            m = re.search(pat, line)
            if m and name:
                out[name][key] = int(m.group(1))
    return out


if __name__ == "__main__":
    print(build(verbose=True))
