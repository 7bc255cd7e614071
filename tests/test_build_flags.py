"""The device code is built without packed FP32 instructions, and the flag that
removes them is effective (hipcc -O3 otherwise forms v_pk_*_f32 from ordinary
float code on gfx950).  Compiles to assembly only; no GPU needed."""
import re
import shutil

import pytest

from synchrad_amd import _build

pytestmark = pytest.mark.skipif(shutil.which("hipcc") is None and not __import__("os").path.exists(
    "/opt/rocm/bin/hipcc"), reason="needs hipcc")

PACKED = re.compile(r"\bv_pk_(fma|mul|add)_f32\b")

PROBE = """
#include <hip/hip_runtime.h>
__global__ void probe(float2 *a, const float2 *b, float s) {
  const int i = threadIdx.x + blockIdx.x * blockDim.x;
  float2 x = a[i], y = b[i];
  x.x = x.x * s + y.x;
  x.y = x.y * s + y.y;
  a[i] = x;
}
"""


@pytest.mark.parametrize("src", _build.SRCS, ids=lambda s: s.rsplit("/", 1)[-1])
def test_library_device_code_has_no_packed_fp32(src):
    asm = _build.device_asm(src)
    assert "v_fma_f32" in asm or "v_fmac_f32" in asm
    assert not PACKED.search(asm)


def test_flag_removes_packed_fp32(tmp_path):
    import subprocess
    probe = tmp_path / "probe.hip"
    probe.write_text(PROBE)
    base = [_build.hipcc(), f"--offload-arch={_build.ARCH}", "-O3", "--cuda-device-only", "-S", str(probe), "-o", "-"]
    plain = subprocess.run(base, capture_output=True, text=True, check=True).stdout
    flagged = subprocess.run(base + ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"],
                             capture_output=True, text=True, check=True).stdout
    assert PACKED.search(plain), "without the flag the compiler forms packed FP32 here"
    assert not PACKED.search(flagged)
