"""Host compile-and-run check of the register FFTs used by the plane-wave
kernels (sternheimergw_amd/csrc/regfft.h), FP64 and FP32, against a
long-double DFT (tests/native/regfft_check.cpp)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_register_fft_against_long_double_dft(tmp_path):
    exe = tmp_path / "regfft_check"
    src = os.path.join(ROOT, "tests", "native", "regfft_check.cpp")
    inc = os.path.join(ROOT, "sternheimergw_amd", "csrc")
    subprocess.run([HIPCC, "-std=c++20", "-O2", "-I", inc, src, "-o", str(exe)], check=True,
                   capture_output=True, timeout=300)
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    print(res.stdout)
    assert res.returncode == 0, res.stdout
