"""GPU density response (sgw_drho_accumulate) against the CPU oracle on the
24^3 cell and, at the 48^3 batch size, against Parseval's box sum."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import density_oracle as do
from sternheimergw_amd import _lib, cell, operators

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _gpu_drho(op, w, group, dpsi_dev, nbnd, nshift, weight):
    psi_r = np.stack([do.sphere_to_real(w.box, w.sphere_k, w.psi_k[v]) for v in range(nbnd)])
    d_psi = torch.from_numpy(np.ascontiguousarray(psi_r)).to(DEV)
    d_out = torch.zeros((nshift,) + tuple(reversed(w.box)), dtype=torch.complex128, device=DEV)
    _lib.check(_lib.lib().sgw_drho_accumulate(
        op.handle, group, nbnd, nshift, ctypes.c_void_p(dpsi_dev.data_ptr()), w.ld,
        ctypes.c_void_p(d_psi.data_ptr()), weight, ctypes.c_void_p(d_out.data_ptr()), 0, None))
    torch.cuda.synchronize()
    return d_out.cpu().numpy()


def test_drho_matches_oracle(c1):
    w = c1
    op = operators.PwOperator.from_workload(w)
    x, ierr, _ = op.solve_batched(w.rhs, w.sigma, group=w.group, lmax=4, threshold=1e-10)
    assert np.all(ierr == 0)
    n = int(w.npw[0])
    weight = 2.0 / 270.011  # 2 wk / Omega with wk = 1
    got = _gpu_drho(op, w, 0, torch.from_numpy(np.ascontiguousarray(x)).to(DEV), w.nbnd,
                    w.nshift, weight)
    ref = do.drho(w, 0, x[:, :, :n], weight)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_drho_parseval_48cubed():
    w = cell.workload_c2(nq=1, nbnd=16)
    print("drho 48^3: 16 bands x 32 shifts", flush=True)
    op = operators.PwOperator.from_workload(w)
    x, ierr, _ = op.solve_batched(w.rhs, w.sigma, group=w.group, lmax=4, threshold=1e-8,
                                  shift_test=True)
    assert np.all(ierr == 0)
    n = int(w.npw[0])
    got = _gpu_drho(op, w, 0, torch.from_numpy(np.ascontiguousarray(x)).to(DEV), w.nbnd,
                    w.nshift, 1.0)
    s = got.reshape(w.nshift, -1).sum(1)
    ref = do.box_sum(w, 0, x[:, :, :n], 1.0)
    assert np.allclose(s, ref, rtol=1e-11, atol=1e-11 * np.abs(ref).max())


def test_drho_matches_oracle_48cubed():
    """Element by element on the 48^3 box (the half-width-8 split kernel):
    2 bands x 4 shifts of random coefficients on the k+q sphere, at 1e-12
    of the largest element.  Catches misplaced values (an x/y transpose)
    that the box sum above cannot see."""
    w = cell.workload_c2(nq=1, nbnd=2)
    print("drho 48^3 element-wise: 2 bands x 4 shifts", flush=True)
    op = operators.PwOperator.from_workload(w)
    n = int(w.npw[0])
    rng = np.random.Generator(np.random.PCG64(17))
    x = np.zeros((2, 4, w.ld), dtype=np.complex128)
    x[:, :, :n] = rng.standard_normal((2, 4, n)) + 1j * rng.standard_normal((2, 4, n))
    got = _gpu_drho(op, w, 0, torch.from_numpy(x).to(DEV), 2, 4, 0.5)
    ref = do.drho(w, 0, x[:, :, :n], 0.5)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
