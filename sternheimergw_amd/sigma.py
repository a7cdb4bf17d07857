"""Sigma_c frequency convolution on the device (host side of sgw_sigma_conv_device).

    Sigma(i omega_i) = alpha * sum_j w_j G(i omega_i + i omega'_j) W(i omega'_j)

is the frequency sum of sigma_correlation (phys/corr/src/sigma.f90:705-745)
written as one complex GEMM per output frequency over the G'' index; the
library runs it on FP64 MFMA.  `SigmaWorkload` builds the seeded synthetic
inputs: N_G reciprocal-lattice vectors of the fcc cell, Gauss-Legendre
frequencies, a Hermitian positive W decaying as 1/(1 + |G|^2) and a Green's
function stack generated on the device from a counter hash.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

from . import _lib, cell

SIGMA_FP64 = _lib.SIGMA_FP64
SIGMA_FP32_SPLIT = _lib.SIGMA_FP32_SPLIT


def g_vectors(n: int, alat: float = 10.26) -> np.ndarray:
    """The n shortest reciprocal-lattice vectors of the fcc cell (bohr^-1),
    sorted by length, then by Miller index."""
    _, b, _ = cell.fcc_lattice(alat)
    return np.ascontiguousarray(cell.shortest_millers(n, alat) @ b)


@dataclass
class SigmaWorkload:
    n: int
    nout: int
    nin: int
    omega: np.ndarray      # output frequencies (Ha), imaginary axis
    omega_p: np.ndarray    # convolution frequencies (Ha)
    weight: np.ndarray     # Gauss-Legendre weights of omega_p
    nu: np.ndarray         # (nout, nin) omega_i + omega'_j
    gvec: np.ndarray       # (n, 3)
    theta: np.ndarray      # (n,) seeded phases of W
    cw: np.ndarray         # (nin,) 1 / (1 + omega'^2)
    length: float
    alpha: complex
    seed: int


def sigma_workload(n: int = 1500, nfreq: int = 64, wmax_ha: float = 4.0, length: float = 1.5,
                   seed: int = cell.SEED_BASE + 4) -> SigmaWorkload:
    """N_G = n, nfreq Gauss-Legendre imaginary frequencies on [0, wmax] for
    both the output and the convolution grid; alpha = -1/(2 pi)."""
    om, wt = cell.gauss_legendre(0.0, wmax_ha, nfreq)
    rng = np.random.Generator(np.random.PCG64(seed))
    theta = rng.uniform(0.0, 2.0 * math.pi, n)
    return SigmaWorkload(n=n, nout=nfreq, nin=nfreq, omega=om, omega_p=om.copy(), weight=wt,
                         nu=np.add.outer(om, om), gvec=g_vectors(n), theta=theta,
                         cw=1.0 / (1.0 + om ** 2), length=length, alpha=-1.0 / (2.0 * math.pi),
                         seed=seed)


def synthetic_device(w: SigmaWorkload, i0: int, ni: int, d_g, d_w, d_nu, d_gvec, d_theta, d_cw,
                     ld: int, stream=None):
    """Fill G[i0:i0+ni] (d_g: (ni, nin, n, ld) complex128, column-major
    matrices) and/or W (d_w: (nin, n, ld)) on the device.  Device arrays are
    torch tensors or None."""
    n, mat = w.n, ld * w.n
    ptr = (lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None)
    _lib.check(_lib.lib().sgw_sigma_synthetic_device(
        n, ld, i0, ni, w.nout, w.nin, w.seed, ptr(d_nu), ptr(d_g), w.nin * mat, mat,
        ptr(d_gvec), ptr(d_theta), ptr(d_cw), w.length, ptr(d_w), mat,
        ctypes.c_void_p(stream) if stream else None))


def conv_device(d_g, d_w, d_weight, d_sigma, n: int, ld: int, nout: int, nin: int,
                alpha: complex, accumulate: bool = False, precision: int = SIGMA_FP64,
                g_strides=None, stream=None):
    """Sigma[i] (+)= alpha sum_j w_j G[i][j] W[j] on device tensors (complex128,
    each matrix n x n column-major with leading dimension ld)."""
    mat = ld * n
    gsi, gsj = g_strides if g_strides is not None else (nin * mat, mat)
    desc = _lib.sgw_sigma_conv_desc(n, ld, nout, nin, gsi, gsj, mat, mat, precision,
                                    int(accumulate))
    _lib.check(_lib.lib().sgw_sigma_conv_device(
        ctypes.byref(desc), ctypes.c_void_p(d_g.data_ptr()), ctypes.c_void_p(d_w.data_ptr()),
        ctypes.c_void_p(d_weight.data_ptr()), _lib.sgw_complex(alpha.real, alpha.imag),
        ctypes.c_void_p(d_sigma.data_ptr()), ctypes.c_void_p(stream) if stream else None))
