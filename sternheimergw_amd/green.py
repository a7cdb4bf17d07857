"""The Green's function of the self-energy on the device: green_function
(phys/green/src/green.f90:105-226) as one batched multishift solve over all
G' columns, its assembly into G(G, G'; omega) on the correlation grid, and
the 6-D transforms of fft6.f90 to G(r, r'; omega) (sigma.f90:688-693).

Arrays are torch tensors in the device layout green[f][c][r] = Fortran
green(r, c, f); the arithmetic is the library's (sgw_green_assemble,
sgw_invfft6, sgw_fwfft6, the batched solver)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .batch import BatchSolver


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class SolverError(RuntimeError):
    pass


def box_index(mill, n):
    m = np.asarray(mill, dtype=np.int64).reshape(-1, 3)
    return ((m[:, 0] % n) + n * ((m[:, 1] % n) + n * (m[:, 2] % n))).astype(np.int32)


class CorrGrid:
    """A correlation FFT box (QE's corr_fft / corr_par_fft): cubic edge n and
    its G list (Miller indices, in the order of the reference's gvect)."""

    def __init__(self, n: int, mill):
        import torch
        self.n = int(n)
        self.mill = np.asarray(mill, dtype=np.int64).reshape(-1, 3)
        self.nr = self.n ** 3
        self.num_g = self.mill.shape[0]
        self.nl = box_index(self.mill, self.n)
        self.d_nl = torch.from_numpy(self.nl).to(torch.device("cuda", torch.cuda.current_device()))


def green_maps(sphere_mill, corr_mill, corr_par_mill):
    """Row and column maps of green.f90: rowmap[r] = 0-based position of
    correlation G r in the sphere, -1 where `0 < map < num_g` (:212) fails;
    colidx[c] = position of G'_c in the sphere, -1 where the column is
    skipped (:199-200)."""
    key = {tuple(m): i + 1 for i, m in enumerate(np.asarray(sphere_mill))}
    num_g = len(sphere_mill)
    mp = np.array([key.get(tuple(m), 0) for m in np.asarray(corr_mill)], dtype=np.int64)
    rowmap = np.where((mp > 0) & (mp < num_g), mp - 1, -1).astype(np.int32)
    col = np.array([key.get(tuple(m), 0) for m in np.asarray(corr_par_mill)], dtype=np.int64)
    return rowmap, (col - 1).astype(np.int32)


def green_function(op, group: int, sphere_mill, omega, grid: CorrGrid, grid_p: CorrGrid = None,
                   *, threshold: float = 1e-10, lmax: int = 4, max_iter: int = 10000,
                   shift_test: bool = True):
    """G(G, G'; omega) for every G' of grid_p (default: grid) as a device
    tensor green[f][c][r] (nfreq, n_p^3, n^3).  `op` is the Green operator
    (alpha_pv = 0), `group` the k-point's group, `sphere_mill` its plane
    waves.  Raises when a column does not converge (green.f90:208)."""
    import torch
    grid_p = grid_p or grid
    dev = grid.d_nl.device
    omega = np.atleast_1d(np.asarray(omega, dtype=np.complex128))
    nfreq = omega.size
    rowmap, colidx = green_maps(sphere_mill, grid.mill, grid_p.mill)
    cols = np.nonzero(colidx >= 0)[0]
    colsys = np.full(grid_p.num_g, -1, dtype=np.int32)
    colsys[cols] = np.arange(cols.size, dtype=np.int32)
    nsys = int(cols.size)
    ld = op.ld
    green = torch.empty((nfreq, grid_p.nr, grid.nr), dtype=torch.complex128, device=dev)
    if nsys == 0:
        x = torch.zeros((1, nfreq, ld), dtype=torch.complex128, device=dev)
    else:
        rhs = np.zeros((nsys, ld), dtype=np.complex128)
        rhs[np.arange(nsys), colidx[cols]] = -1.0                 # green.f90:203-204
        sigma = np.tile(-omega, (nsys, 1))                        # :207
        bs = BatchSolver(op, nsys, nfreq, ld, lmax=lmax, max_iter=max_iter, threshold=threshold,
                         shift_test=shift_test)
        d_group = torch.full((nsys,), group, dtype=torch.int32, device=dev)
        x = torch.zeros((nsys, nfreq, ld), dtype=torch.complex128, device=dev)
        ierr, _it = bs.solve(d_group, torch.from_numpy(rhs).to(dev),
                             torch.from_numpy(sigma).to(dev), sigma, x)
        if np.any(ierr != 0):
            raise SolverError("the linear solver for G did not converge")
    d_row = torch.from_numpy(rowmap).to(dev)
    d_col = torch.from_numpy(colsys).to(dev)
    _lib.check(_lib.lib().sgw_green_assemble(nfreq, grid.n, grid_p.n, grid.num_g, _p(d_row),
                                             grid_p.num_g, _p(d_col), _p(x), nfreq, ld,
                                             _p(green), _stream()))
    return green


def invfft6(f, grid: CorrGrid, omega_cell: float, grid_p: CorrGrid = None):
    """In place on f (nmat, nr_p, nr): f(r, r') of fft6.f90's invfft6."""
    grid_p = grid_p or grid
    nmat = f.numel() // (grid.nr * grid_p.nr)
    _lib.check(_lib.lib().sgw_invfft6(grid.n, grid_p.n, grid.num_g, _p(grid.d_nl), grid_p.num_g,
                                      _p(grid_p.d_nl), omega_cell, nmat, _p(f), _stream()))
    return f


def fwfft6(f, grid: CorrGrid, omega_cell: float, grid_p: CorrGrid = None):
    """In place on f (nmat, nr_p, nr): f(G, G') of fft6.f90's fwfft6 in the
    (num_g, num_g_p) block."""
    grid_p = grid_p or grid
    nmat = f.numel() // (grid.nr * grid_p.nr)
    _lib.check(_lib.lib().sgw_fwfft6(grid.n, grid_p.n, grid.num_g, _p(grid.d_nl), grid_p.num_g,
                                     _p(grid_p.d_nl), omega_cell, nmat, _p(f), _stream()))
    return f
