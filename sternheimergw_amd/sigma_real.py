"""The reference form of the correlation self-energy on the device
(phys/corr/src/sigma.f90:528-745): W continued analytically from its solver
frequencies (analytic.f90: Godby-Needs, Pade, robust Pade or AAA), then for every Sigma
frequency and every Green's-function frequency the real-space product of
sigma_prod, accumulated into Sigma(G, G').  The GEMM restatement of the
same sum over frequencies is sigma.py.

Device tensors use the layout [f][c][r] of Fortran arrays (r, c, f); the
arithmetic is the library's (sgw_analytic_coeff, sgw_analytic_eval,
sgw_sigma_prod)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .green import CorrGrid

NO_SYMMETRY, EVEN_SYMMETRY, SQUARE_SYMMETRY = 0, 1, 2


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _cplx(z):
    return _lib.sgw_complex(float(np.real(z)), float(np.imag(z)))


def freqbins_symm(solver, symm):
    """The symmetrised mesh of freqbins.f90:333-425 and, for every point,
    the solver frequency whose W it carries."""
    solver = np.asarray(solver, dtype=np.complex128)
    n = solver.size
    if symm == NO_SYMMETRY:
        return solver.copy(), np.arange(n, dtype=np.int32)
    if symm == SQUARE_SYMMETRY:
        return solver ** 2, np.arange(n, dtype=np.int32)
    zero = np.abs(solver) < 1e-14
    if zero.sum() > 1:
        raise ValueError("only a single frequency may be smaller than 1e-14")
    keep = np.nonzero(~zero)[0]
    return (np.concatenate([solver, -solver[keep]]),
            np.concatenate([np.arange(n), keep]).astype(np.int32))


class ScreenedCoulomb:
    """Coefficients of W's analytic continuation on the device
    (analytic_coeff, analytic.f90:50-208)."""

    def __init__(self, d_w, solver, model: int = _lib.PADE_APPROX, symm: int = NO_SYMMETRY,
                 thres: float = 0.0):
        """d_w: (nfreq_sym, num_g, num_g) complex128 device tensor with W at
        the solver frequencies in its first len(solver) slices; overwritten
        by the coefficients.  thres: analytic_coeff's threshold (tr2_gw),
        used by the AAA models."""
        solver = np.asarray(solver, dtype=np.complex128)
        self.model, self.symm = model, symm
        self.num_g = d_w.shape[-1]
        if model == _lib.GODBY_NEEDS:
            self.z = solver[:2].copy()
            src = None
        elif model == _lib.PADE_ROBUST:           # the circle itself (analytic.f90:158-160)
            self.z = solver.copy()
            src = None
        else:
            self.z, src = freqbins_symm(solver, symm)
        nz = self.z.size
        if d_w.shape[0] < nz:
            raise ValueError("W needs one slice per point of the symmetrised mesh")
        self._hz_buf = np.ascontiguousarray(self.z, dtype=np.complex128)
        hz = ctypes.c_void_p(self._hz_buf.ctypes.data)
        self._src_buf = None if src is None else np.ascontiguousarray(src, dtype=np.int32)
        hs = None if src is None else ctypes.c_void_p(self._src_buf.ctypes.data)
        _lib.check(_lib.lib().sgw_analytic_coeff(model, float(thres), self.num_g, nz, hz, hs,
                                                 _p(d_w), _stream()))
        self._hz = hz
        self.coeff = d_w

    def symmetrize(self, freq):
        return freq ** 2 if self.symm == SQUARE_SYMMETRY else freq

    def eval_into(self, freq, grid: CorrGrid, grid_p: CorrGrid, d_gmap, d_gmap_p, out):
        """analytic_eval (analytic.f90:211-348) into out (nr_p, nr)."""
        _lib.check(_lib.lib().sgw_analytic_eval(
            self.model, self.num_g, self.z.size, self._hz, _p(self.coeff),
            _cplx(self.symmetrize(freq)), grid.n, grid_p.n, grid.num_g, grid_p.num_g,
            _p(d_gmap), _p(d_gmap_p), _p(out), _stream()))
        return out


def sigma_prod(grid: CorrGrid, grid_p: CorrGrid, omega_cell: float, alpha, d_green_r, d_array,
               d_sigma=None):
    """sigma.f90:417-523 on one W (d_array, W(G, G') block in, alpha Sigma
    block out); adds into d_sigma (num_g_p, num_g) when given."""
    _lib.check(_lib.lib().sgw_sigma_prod(
        grid.n, grid_p.n, grid.num_g, _p(grid.d_nl), grid_p.num_g, _p(grid_p.d_nl), omega_cell,
        _cplx(alpha), _p(d_green_r), _p(d_array), _p(d_sigma),
        d_sigma.shape[-1] if d_sigma is not None else 0, _stream()))
    return d_array


def sigma_correlation(grid: CorrGrid, grid_p: CorrGrid, omega_cell: float, alpha, mu, coul,
                      weight, sigma_freq, d_green_r, w: ScreenedCoulomb, gmap=None, gmap_p=None):
    """Sigma_c(G, G'; omega_sigma) of sigma.f90:653-745.  d_green_r:
    G(r, r'; omega) (2 num_coul, nr_p, nr) at freq_green = [mu + coul,
    mu - coul] (freqbins.f90:302-322).  Returns (nsigma, num_g_p, num_g)."""
    import torch
    dev = d_green_r.device
    coul = np.asarray(coul, dtype=np.complex128)
    num_coul = coul.size
    freq_green = np.concatenate([mu + coul, mu - coul])
    freq_sigma = mu + np.asarray(sigma_freq, dtype=np.complex128)
    gm = np.arange(grid.num_g, dtype=np.int32) if gmap is None else np.asarray(gmap, np.int32)
    gmp = np.arange(grid_p.num_g, dtype=np.int32) if gmap_p is None else np.asarray(gmap_p,
                                                                                      np.int32)
    d_gm, d_gmp = torch.from_numpy(gm).to(dev), torch.from_numpy(gmp).to(dev)
    sigma = torch.zeros((freq_sigma.size, grid_p.num_g, grid.num_g), dtype=torch.complex128,
                        device=dev)
    work = torch.empty((grid_p.nr, grid.nr), dtype=torch.complex128, device=dev)
    for isig, fs in enumerate(freq_sigma):
        for ig, fg in enumerate(freq_green):
            fc = fs - fg                                             # sigma.f90:709
            if fc.real * fc.imag < 0.0:                              # :711
                fc = np.conj(fc)
            w.eval_into(fc, grid, grid_p, d_gm, d_gmp, work)         # :714
            icoul = ig % num_coul                                    # :722
            sigma_prod(grid, grid_p, omega_cell, alpha * weight[icoul], d_green_r[ig], work,
                       sigma[isig])                                  # :725, :741
    return sigma
