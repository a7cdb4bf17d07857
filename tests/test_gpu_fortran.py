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


CALLERS = os.path.join(ROOT, "sternheimergw_amd", "fortran", "sgw_callers_test")


def test_fortran_callers_batched_loops(c1, tmp_path):
    """sgw_callers_module (compiled Fortran on the C ABI): green_columns (the
    G' loop of green.f90:196-218) on the Green operator and linter_bands (the
    band loop of solve_linter.f90:325-374 with the +/- omega average) on the
    Coulomb operator of the 24^3 cell, each one batched device solve,
    against dense direct solves at 1e-9."""
    if not os.path.exists(CALLERS):
        pytest.skip("Fortran driver not built (amdflang missing at build time)")
    w = c1
    n1, n2, n3 = w.box
    ld, ng, nproj, nb = w.ld, len(w.spheres), w.nproj, w.nproj_beta
    npw = w.npw.astype("<i4")
    cols = np.array([1, 4, 18, int(npw[0])], dtype="<i4")      # 1-based, the last plane wave
    omega = np.array([0.1 + 0.3j, -0.5 + 0.2j, 0.4j])
    freq = 1j * np.array([0.0, 0.3, 1.1])
    nsys = w.nbnd
    et = w.eps_v[:nsys].astype("<f8")
    wg = np.full(nsys, 2.0)
    wk = np.ones(nsys)
    pin, pout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(pin, "wb") as f:
        np.array([n1, n2, n3, ng, nproj, nb], dtype="<i4").tofile(f)
        np.array([ld], dtype="<i8").tofile(f)
        npw.tofile(f)
        np.ascontiguousarray(w.vloc, dtype="<f8").tofile(f)
        np.ascontiguousarray(w.mill_padded(), dtype="<i4").tofile(f)   # (g, ld, 3) = mill(3, ld, g)
        np.ascontiguousarray(w.kin_padded(), dtype="<f8").tofile(f)
        np.ascontiguousarray(w.proj, dtype="<c16").tofile(f)           # (g, p, ld) = proj(ld, p, g)
        np.ascontiguousarray(w.dcoef, dtype="<f8").tofile(f)
        np.array([0, cols.size, omega.size], dtype="<i4").tofile(f)
        cols.tofile(f)
        omega.astype("<c16").tofile(f)
        np.array([nsys, freq.size], dtype="<i4").tofile(f)
        w.group[:nsys].astype("<i4").tofile(f)
        et.tofile(f)
        wg.astype("<f8").tofile(f)
        wk.astype("<f8").tofile(f)
        freq.astype("<c16").tofile(f)
        np.ascontiguousarray(w.rhs[:nsys], dtype="<c16").tofile(f)
    print("Fortran callers driver", flush=True)
    res = subprocess.run([CALLERS, str(pin), str(pout)], capture_output=True, text=True,
                         timeout=170)
    print(res.stdout[-2000:])
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-2000:]
    n = int(npw[0])
    raw = np.fromfile(pout, dtype=np.uint8)
    off = 0
    green = np.frombuffer(raw[off:off + 16 * n * omega.size * cols.size], dtype="<c16")
    green = green.reshape(cols.size, omega.size, n)
    off += green.nbytes
    ierr_g = np.frombuffer(raw[off:off + 4 * cols.size], dtype="<i4")
    off += ierr_g.nbytes
    dpsi = np.frombuffer(raw[off:off + 16 * ld * freq.size * nsys], dtype="<c16")
    dpsi = dpsi.reshape(nsys, freq.size, ld)
    off += dpsi.nbytes
    ierr_l = np.frombuffer(raw[off:off + 4 * nsys], dtype="<i4")
    assert np.all(ierr_g == 0) and np.all(ierr_l == 0)
    h = w.dense_h[0]
    for i, c in enumerate(cols):
        e = np.zeros(n, dtype=np.complex128)
        e[c - 1] = -1.0
        for k, om in enumerate(omega):
            ref = np.linalg.solve(h - om * np.eye(n), e)
            assert np.linalg.norm(green[i, k] - ref) <= 1e-9 * np.linalg.norm(ref), (c, om)
    ev = w.evq[0]
    hc = h + w.alpha_pv * (ev.T @ ev.conj())
    for v in range(nsys):
        b = w.rhs[v, :n]

        def solve(om):
            return np.linalg.solve(hc - (et[v] + om) * np.eye(n), b) * wg[v] / wk[v]

        for k, om in enumerate(freq):
            ref = solve(om) if k == 0 else 0.5 * (solve(om) + solve(-om))
            assert np.linalg.norm(dpsi[v, k, :n] - ref) <= 1e-9 * np.linalg.norm(ref), (v, k)
