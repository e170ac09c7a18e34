"""The library's Halton StartPixel forms the per-pixel offset in 32-bit
arithmetic; pine sums two 64-bit products and reduces them modulo the stride
(sampler.cpp:57-70).  tests/halton_offset_check.cpp compiles the library's
header (HIP source: hipcc, host code only) and compares the two for every pixel of 324 film sizes.
(The device build, whose base-2 digits take a bit reverse instead of the
loop, is checked bit for bit against pine's sampler values on the GPU.)"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "pine_amd", "csrc")


HIPCC = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not found")
def test_start_pixel_offset_matches_the_64_bit_sum(tmp_path):
    exe = str(tmp_path / "halton_offset_check")
    subprocess.run([HIPCC, "-x", "hip", "--cuda-host-only", "-std=c++17", "-O2", "-I", CSRC, "-o", exe,
                    os.path.join(HERE, "halton_offset_check.cpp")], check=True)
    p = subprocess.run([exe], stdout=subprocess.PIPE, text=True, timeout=300)
    assert p.returncode == 0, p.stdout
    assert "0 mismatches" in p.stdout, p.stdout
