"""The C-ABI library loads and exports every entry point include/sgw.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sgw.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(sgw_[a-z0-9_]+)\s*\(", text)
    return sorted(set(n for n in names if not n.endswith("_fn")))


@pytest.fixture(scope="module")
def libpath():
    from sternheimergw_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from sternheimergw_amd import build
        build.build_hip()
    return _lib.LIB_PATH


def test_header_declares_entry_points():
    names = _declared()
    for n in ("sgw_select_solver", "sgw_bicgstab", "sgw_linear_solver", "sgw_operator_pw",
              "sgw_batch_solver_run", "sgw_residual_norms_device"):
        assert n in names


def test_library_exports_every_declared_symbol(libpath):
    handle = ctypes.CDLL(libpath)
    missing = [n for n in _declared() if not hasattr(handle, n)]
    assert not missing, missing


def test_binding_covers_every_symbol(libpath):
    from sternheimergw_amd import _lib
    bound = {n for n, _, _ in _lib.SIGNATURES}
    assert set(_declared()) == bound


def test_library_is_gfx950_code_object(libpath):
    out = subprocess.run(["/opt/rocm/llvm/bin/llvm-readelf", "-S", libpath], capture_output=True,
                         text=True)
    if out.returncode != 0:
        pytest.skip("llvm-readelf unavailable")
    assert ".hip_fatbin" in out.stdout
    blob = open(libpath, "rb").read()
    assert b"gfx950" in blob


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from sternheimergw_amd import _lib
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _lib.lib()


def test_fortran_shim_declares_reference_modules():
    src = open(os.path.join(ROOT, "sternheimergw_amd", "fortran", "select_solver_module.f90")).read()
    for mod in ("MODULE select_solver_module", "MODULE bicgstab_module",
                "MODULE linear_solver_module"):
        assert mod in src
