"""Build recipes for the in-tree native libraries.

* libsgw_hip.so   — the product: HIP kernels for gfx950 and the C ABI of
                    include/sgw.h (built with hipcc).
* Fortran shims   — ISO_C_BINDING modules mirroring the reference's Fortran
                    interfaces plus a test driver (built with amdflang when the
                    compiler is present).

Run ``python -m sternheimergw_amd.build`` or call :func:`build_all`.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsgw_hip.so")
FORTRAN = os.path.join(HERE, "fortran")
OBJDIR = os.path.join(HERE, "_build")

HIP_SOURCES = ["operators.hip", "bicgstab.hip", "subspace.hip", "sigma.hip", "adjoint.hip",
               "density.hip", "boxfft.hip", "scf.hip", "fft6.hip", "analytic.hip", "aaa.hip", "ktimer.hip", "tune.hip", "mixed.hip", "capi.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++20",
    "-Wall",
    "-Wno-unused-result",
    # never emit packed FP32 (v_pk_*_f32) instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
    "-I" + CSRC,
]


def _hipcc() -> str:
    for cand in ("/opt/rocm/bin/hipcc", shutil.which("hipcc") or ""):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def _run(cmd, **kw):
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, **kw)
    if res.returncode != 0:
        sys.stderr.write(res.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd[:3]) + " ...")
    return res.stdout


# per-file additions: the plane-wave operator's FFT kernels hold no signed
# zeros, NaNs or infinities that matter, so x + 0 and x * 0 may fold (this is
# what lets the band-specialised transforms drop their zero butterflies)
FILE_FLAGS = {"operators.hip": ["-fno-signed-zeros", "-ffinite-math-only"]}


def build_hip(verbose: bool = False) -> str:
    hipcc = _hipcc()
    os.makedirs(OBJDIR, exist_ok=True)
    objs = []
    procs = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJDIR, src.replace(".hip", ".o"))
        cmd = [hipcc, *HIPCC_FLAGS, *FILE_FLAGS.get(src, []), "-c", os.path.join(CSRC, src),
               "-o", obj]
        if verbose:
            print(" ".join(cmd))
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    failed = False
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            failed = True
        elif verbose and out.strip():
            print(out)
    if failed:
        raise RuntimeError("hipcc compilation failed")
    tmp = LIB + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_fortran(verbose: bool = False):
    """ISO_C_BINDING shims and their test driver (needs amdflang)."""
    fc = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang"
                                      if os.path.exists("/opt/rocm/bin/amdflang") else None)
    if fc is None:
        return None
    mod = os.path.join(OBJDIR, "fmod")
    os.makedirs(mod, exist_ok=True)
    modules = ["sgw_kinds.f90", "select_solver_module.f90", "sgw_callers.f90"]
    exes = []
    for driver in ("sgw_fixture_test", "sgw_callers_test"):
        exe = os.path.join(FORTRAN, driver)
        cmd = [fc, "-O2", "-J", mod, *[os.path.join(FORTRAN, s) for s in modules],
               os.path.join(FORTRAN, driver + ".f90"),
               "-L" + HERE, "-lsgw_hip", "-Wl,-rpath,$ORIGIN/..", "-o", exe]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        exes.append(exe)
    return exes


def build_oracle(verbose: bool = False) -> str:
    """The CPU oracle (test infrastructure, not part of the product)."""
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "liboracle.so")


def build_all(verbose: bool = False):
    build_hip(verbose)
    if os.path.isdir(FORTRAN) and os.path.exists(os.path.join(FORTRAN, "sgw_kinds.f90")):
        build_fortran(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print("built", LIB)
