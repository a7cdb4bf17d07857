"""GPU parity of the 48^3 kernels that the batched 48^3 workload runs.

The 48^3 H-apply picks transforms specialised to the sphere's half-width
M = max |Miller index| (8, 9, 10 or 11 on a 48 box).  The 25 Ha batched
workload has M = 8; higher cutoffs reach the other three.  Every test here
compares a device result on such a cell with the oracle (the C restatement,
oracle/sgw_oracle.c), not with another device result:

* H-apply of ragged vector sets (several 16-vector projector tiles, an odd
  count, two k+q groups) at 1e-12 relative, for each M and for the general
  kernels (SGW_PW_BAND=0) as well;
* whole multishift solves of the 25 Ha cell (32 shifts) at threshold 1e-12:
  every shift within 1e-10 relative of the oracle's bicgstab
  (algo/linear_solver/src/bicgstab.f90:154-275);
* the streamed solve that the benchmark times (more systems than slots) and
  the device residual check that the benchmark uses to count solves, against
  residuals the oracle's operator recomputes.
"""
import functools

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from sternheimergw_amd import _lib, cell, operators
from sternheimergw_amd.batch import BatchSolver

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")
# cutoff (Ha) giving sphere half-width M on the 48^3 box (all three axes)
CUTOFF = {8: 25.0, 9: 33.0, 10: 40.0, 11: 47.0}
# exit-iteration spread allowed near the attainable accuracy (see test_gpu_pw)
ITER_JITTER = 2


@functools.lru_cache(maxsize=None)
def _cell(m: int):
    """Two k+q groups (ragged N_pw), 4 bands, 32 projectors, 32 GL shifts."""
    q, scale = cell.c2_qpoints(2)
    w, _ = cell.gauss_legendre(0.0, 4.0, 32)
    wl = cell.build_workload(f"c2-m{m}", box=(48, 48, 48), ecut_ha=CUTOFF[m], nbnd=4,
                             lchan=[0, 1, 2, 3], qpoints=q, ecut_scale=scale,
                             omega_ha=1j * w, seed=cell.SEED_BASE + 2)
    mmax = max(int(np.abs(s.mill).max()) for s in wl.spheres)
    assert mmax == m, (m, mmax)
    return wl


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


@pytest.mark.parametrize("band", ["1", "0"])
@pytest.mark.parametrize("m", [8, 9, 10, 11])
def test_happly_48cubed_matches_oracle(m, band, monkeypatch):
    """37 vectors over two groups: three projector tiles (the last ragged),
    an odd vector count and mixed shifts.  band=0 runs the general kernels."""
    w = _cell(m)
    monkeypatch.setenv("SGW_PW_BAND", band)
    print(f"H-apply 48^3 M={m} band={band}", flush=True)
    op = operators.PwOperator.from_workload(w)
    rng = np.random.Generator(np.random.PCG64(100 + m))
    nv = 37
    grp = (np.arange(nv) * 7 % 3 == 0).astype(np.int32)
    x = np.zeros((nv, w.ld), dtype=np.complex128)
    for v in range(nv):
        n = w.npw[grp[v]]
        x[v, :n] = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    sig = rng.standard_normal(nv) + 1j * rng.standard_normal(nv)
    ax = op.apply(x, sig, group=grp)
    ops = [orc.PwOp(w, g) for g in range(2)]
    worst = 0.0
    for v in range(nv):
        n = w.npw[grp[v]]
        ref = ops[grp[v]].apply(x[v, :n], sig[v])
        worst = max(worst, np.linalg.norm(ax[v, :n] - ref) / np.linalg.norm(ref))
        assert np.all(ax[v, n:] == 0), v
    assert worst <= 1e-12, worst


@pytest.fixture(scope="module")
def solved_c2():
    """All 8 systems of the 25 Ha cell solved on the device at 1e-12 and at
    1e-10."""
    w = _cell(8)
    op = operators.PwOperator.from_workload(w)
    print("48^3 solves at 1e-12 and 1e-10: 8 systems x 32 shifts", flush=True)
    x, ierr, iters = op.solve_batched(w.rhs, w.sigma, group=w.group, threshold=1e-12)
    _, ierr10, iters10 = op.solve_batched(w.rhs, w.sigma, group=w.group, threshold=1e-10)
    return w, x, ierr, iters, ierr10, iters10


@pytest.mark.parametrize("s", [0, 3, 5])
def test_solve_48cubed_matches_oracle(solved_c2, s):
    """Bands 0 and 3 of q 0 and band 1 of q 1: every one of the 32 shifted
    solutions of the 1e-12 solve within 1e-10 relative of the oracle's FP64
    multishift solve, and the exit iteration of the 1e-10 solve within
    ITER_JITTER of the oracle's.  (Near 1e-12 the last iterations stagnate at
    the attainable accuracy and the exit iteration follows the rounding: two
    FP64 builds of the H-apply 1 ulp apart exit at 26 and 28 on system 5,
    the oracle at 25; at 1e-10 the recurrences still decide it.)"""
    w, x, ierr, iters, ierr10, iters10 = solved_c2
    assert ierr[s] == 0 and ierr10[s] == 0
    g = int(w.group[s])
    n = int(w.npw[g])
    print(f"oracle bicgstab 48^3 system {s}", flush=True)
    xo, eo, io = orc.bicgstab(orc.PwOp(w, g), w.rhs[s, :n], w.sigma[s], lmax=4, threshold=1e-12)
    assert eo == 0
    rel = np.linalg.norm(x[s, :, :n] - xo, axis=1) / np.linalg.norm(xo, axis=1)
    assert rel.max() <= 1e-10, rel.max()
    _, eo10, io10 = orc.bicgstab(orc.PwOp(w, g), w.rhs[s, :n], w.sigma[s], lmax=4,
                                 threshold=1e-10)
    assert eo10 == 0 and abs(io10 - iters10[s]) <= ITER_JITTER, (io10, iters10[s])


def test_streamed_solve_and_residual_check_against_oracle():
    """The benchmark's path: a queue of 12 systems through 5 slots with the
    shift test at 1e-8, then sgw_residual_norms_device.  Each device
    residual equals the one the oracle's operator gives for the same
    solution, and every solution meets 1e-8 by the oracle's residual."""
    w = _cell(8)
    op = operators.PwOperator.from_workload(w)
    order = np.array([0, 4, 1, 5, 2, 6, 3, 7, 0, 7, 4, 3])
    rhs, sig, grp = w.rhs[order], np.ascontiguousarray(w.sigma[order]), w.group[order]
    nq, ns, ld = len(order), w.nshift, w.ld
    sol = BatchSolver(op, 5, ns, ld, lmax=4, threshold=1e-8, shift_test=True)
    d_x = torch.zeros((nq, ns, ld), dtype=torch.complex128, device=DEV)
    d_rhs, d_sig, d_g = _t(rhs), _t(sig), _t(grp)
    print("streamed 48^3: 12 systems through 5 slots", flush=True)
    ierr, _ = sol.solve_stream(grp, d_rhs, d_sig, sig, d_x)
    assert np.all(ierr == 0)
    d_res = torch.zeros((nq, ns), dtype=torch.float64, device=DEV)
    _lib.check(_lib.lib().sgw_residual_norms_device(
        op.handle, nq, ns, ld, _lib.c_void_p(d_g.data_ptr()), _lib.c_void_p(d_rhs.data_ptr()),
        _lib.c_void_p(d_sig.data_ptr()), _lib.c_void_p(d_x.data_ptr()),
        _lib.c_void_p(d_res.data_ptr()), None))
    torch.cuda.synchronize()
    res = d_res.cpu().numpy()
    x = d_x.cpu().numpy()
    ops = [orc.PwOp(w, g) for g in range(2)]
    for i in range(nq):
        g = int(grp[i])
        n = int(w.npw[g])
        bn = np.linalg.norm(rhs[i, :n])
        for k in range(ns):
            r = np.linalg.norm(rhs[i, :n] - ops[g].apply(x[i, k, :n], sig[i, k])) / bn
            assert r <= 1e-8, (i, k, r)
            assert abs(res[i, k] - r) <= 1e-3 * r + 1e-13, (i, k, res[i, k], r)


def test_autotuned_shapes_agree_bitwise(monkeypatch):
    """Every launch shape the autotuner may pick for the 48^3 plane
    transforms and the assembly (SGW_AUTOTUNE_FORCE=i) gives the default
    shape's H-apply bit for bit, and the tuned run records a choice."""
    w = _cell(8)
    op = operators.PwOperator.from_workload(w)
    rng = np.random.Generator(np.random.PCG64(5))
    nv = 21
    grp = (np.arange(nv) % 2).astype(np.int32)
    x = np.zeros((nv, w.ld), dtype=np.complex128)
    for v in range(nv):
        n = w.npw[grp[v]]
        x[v, :n] = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    sig = rng.standard_normal(nv) + 1j * rng.standard_normal(nv)
    out = []
    for c in ("0", "1", "2", "3"):
        monkeypatch.setenv("SGW_AUTOTUNE_FORCE", c)
        out.append(op.apply(x, sig, group=grp))
    monkeypatch.delenv("SGW_AUTOTUNE_FORCE")
    tuned = op.apply(x, sig, group=grp)
    for o in out[1:] + [tuned]:
        assert np.array_equal(o, out[0])
    L = _lib.lib()
    assert L.sgw_autotune_choice(b"plane48/M8/b8/n32") in (0, 1, 2, 3)
    assert L.sgw_autotune_choice(b"assemble/n32/g2/p36") in (0, 1, 2)
    ref = orc.PwOp(w, 0).apply(x[0, :w.npw[0]], sig[0])
    assert np.linalg.norm(tuned[0, :w.npw[0]] - ref) <= 1e-12 * np.linalg.norm(ref)
