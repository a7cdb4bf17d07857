"""CPU oracle of the Green's function of the self-energy (test infrastructure
only):

* green_function — phys/green/src/green.f90:105-226 statement by statement:
  the row map over the correlation G list into the k-point's sphere, the
  columns (skipped when the G' vector is not in the sphere, :199-200), the
  right-hand side -e_{G'} (:203-204), the shifts -omega (:207) and the masked
  copy `WHERE (map > 0 .AND. map < num_g)` (:211-213), which also leaves out
  the last plane wave of the sphere.  The linear systems are solved densely
  (numpy), as the iterative solver is pinned on its own.
* invfft6 / fwfft6 — data/fft/src/fft6.f90:84-181 and 215-345, steps 1-6
  with the conjugations around QE's fwfft/invfft, on the oracle's FFT.

Arrays use the device layout F[c][r] = Fortran f(r, c) (first index
fastest).  Pins (tests/test_green_oracle.py): the plane-wave closed form of
invfft6, the fwfft6(invfft6) round trip, and the spectral sum
-sum_n psi_n(G) psi_n(G')^* / (e_n - omega) looked up through explicit
Miller-index searches.
"""
from __future__ import annotations

import numpy as np

from . import oracle as orc


def _fft(box_flat, n, sign):
    return orc.fft3d(np.array(box_flat, dtype=np.complex128).reshape(n, n, n), sign).reshape(-1)


def box_index(mill, n):
    m = np.asarray(mill, dtype=np.int64)
    return ((m[:, 0] % n) + n * ((m[:, 1] % n) + n * (m[:, 2] % n))).astype(np.int32)


def invfft6(F, n, n_p, nl, nl_p, omega):
    """F (nr_p, nr) complex, in place and returned."""
    num_g, num_g_p = len(nl), len(nl_p)
    nr_p = n_p ** 3
    for ig in range(num_g):                                   # fft6.f90:297-311
        work = np.zeros(nr_p, dtype=np.complex128)
        work[nl_p] = F[:num_g_p, ig] / omega
        F[:, ig] = _fft(work, n_p, +1)
    for ir in range(nr_p):                                    # :320-338
        work = np.zeros(n ** 3, dtype=np.complex128)
        work[nl] = np.conj(F[ir, :num_g])
        F[ir, :] = np.conj(_fft(work, n, +1))
    return F


def fwfft6(F, n, n_p, nl, nl_p, omega):
    """F (nr_p, nr) complex, in place and returned; only F[:num_g_p, :num_g]
    is defined on return."""
    num_g, num_g_p = len(nl), len(nl_p)
    nr_p = n_p ** 3
    for ir in range(nr_p):                                    # fft6.f90:146-163
        work = np.conj(F[ir, :])
        work = _fft(work, n, -1)
        F[ir, :num_g] = np.conj(work[nl])
    for ig in range(num_g):                                   # :172-186
        work = F[:, ig].copy()
        work = _fft(work, n_p, -1)
        F[:num_g_p, ig] = work[nl_p] * omega
    return F


def green_maps(sphere_mill, corr_mill, corr_par_mill):
    """(rowmap, colidx): 0-based sphere positions; rowmap is -1 where the
    reference's mask 0 < map < num_g fails, colidx is -1 where the column is
    skipped (map == 0)."""
    key = {tuple(m): i + 1 for i, m in enumerate(np.asarray(sphere_mill))}
    num_g = len(sphere_mill)
    mp = np.array([key.get(tuple(m), 0) for m in np.asarray(corr_mill)], dtype=np.int64)
    rowmap = np.where((mp > 0) & (mp < num_g), mp - 1, -1).astype(np.int32)   # :212
    col = np.array([key.get(tuple(m), 0) for m in np.asarray(corr_par_mill)], dtype=np.int64)
    colidx = (col - 1).astype(np.int32)                                      # :199-200
    return rowmap, colidx


def green_function(hmat, sphere_mill, corr_mill, corr_par_mill, omega, n, n_p):
    """green[f][c][r] (nfreq, n_p^3, n^3) of green.f90 with dense solves of
    (H - omega_f) x = -e_{G'}."""
    omega = np.atleast_1d(np.asarray(omega, dtype=np.complex128))
    rowmap, colidx = green_maps(sphere_mill, corr_mill, corr_par_mill)
    num_g = len(sphere_mill)
    green = np.zeros((omega.size, n_p ** 3, n ** 3), dtype=np.complex128)
    rows = np.nonzero(rowmap >= 0)[0]
    for igp, ig in enumerate(colidx):
        if ig < 0:
            continue
        bb = np.zeros(num_g, dtype=np.complex128)
        bb[ig] = -1.0
        for f, om in enumerate(omega):
            part = np.linalg.solve(hmat - om * np.eye(num_g), bb)   # sigma = -omega
            green[f, igp, rows] = part[rowmap[rows]]
    return green
