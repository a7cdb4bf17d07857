"""The reference's solver unit test, run as Fortran through the drop-in
modules (sternheimergw_amd/fortran/*.f90 -> ISO_C_BINDING -> libsgw_hip.so)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "sternheimergw_amd", "fortran", "sgw_fixture_test")

pytestmark = pytest.mark.gpu


def test_fortran_fixture_driver(lin_prob, tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("Fortran driver not built (amdflang missing at build time)")
    path = tmp_path / "lin_prob.bin"
    with open(path, "wb") as f:
        np.array([lin_prob["DIMENSION"], lin_prob["NUMBER_SHIFT"]], dtype="<i4").tofile(f)
        lin_prob["LIST_SHIFT"].astype("<c16").tofile(f)
        np.asfortranarray(lin_prob["LINEAR_OPERATOR"]).astype("<c16").T.tofile(f)
        lin_prob["RIGHT_HAND_SIDE"].astype("<c16").tofile(f)
    res = subprocess.run([EXE, str(path)], capture_output=True, text=True, timeout=170)
    print(res.stdout[-2000:])
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-2000:]
    assert "PASS" in res.stdout
