"""CPU oracle of the density response accumulated after the solve (test
infrastructure only): QE's incdrhoscf as called by solve_linter
(phys/coul/src/solve_linter.f90:486-497),

    drho_k(r) = weight * sum_v conj(psi_v(r)) dpsi_{v,k}(r),

with psi_v(r), dpsi(r) the inverse FFTs (QE sign, no factor) of the
plane-wave coefficients on their own spheres (k for psi, k+q for dpsi).
QE's incdrhoscf is external to the reference tree, so the formula is its
published one; the FFT is the oracle's (pinned by brute-force DFTs in
tests/test_oracle.py) and Parseval pins the sum over the box.
"""
from __future__ import annotations

import numpy as np

from . import oracle as orc


def sphere_to_real(box, sphere, coef):
    """invfft of coefficients on `sphere` into a (n3, n2, n1) box."""
    n1, n2, n3 = box
    b = np.zeros(n1 * n2 * n3, dtype=np.complex128)
    b[sphere.box_index(box)] = coef[: sphere.npw]
    return orc.fft3d(b.reshape(n3, n2, n1), +1)


def drho(w, group, dpsi, weight):
    """dpsi (nbnd, nshift, npw_q) of k+q group `group` -> (nshift, n3, n2, n1)."""
    nbnd, nshift = dpsi.shape[:2]
    sq = w.spheres[group]
    out = np.zeros((nshift,) + tuple(reversed(w.box)), dtype=np.complex128)
    for v in range(nbnd):
        pr = sphere_to_real(w.box, w.sphere_k, w.psi_k[v])
        for k in range(nshift):
            out[k] += weight * pr.conj() * sphere_to_real(w.box, sq, dpsi[v, k])
    return out


def box_sum(w, group, dpsi, weight):
    """Parseval: sum_r drho_k(r) = N weight sum_v sum_m conj(psi_v(m)) dpsi_{v,k}(m)
    over the Miller indices m present in both spheres."""
    n = int(np.prod(w.box))
    key = {tuple(m): i for i, m in enumerate(w.sphere_k.mill)}
    sq = w.spheres[group]
    idx = np.array([key.get(tuple(m), -1) for m in sq.mill])
    hit = idx >= 0
    nbnd, nshift = dpsi.shape[:2]
    out = np.zeros(nshift, dtype=np.complex128)
    for v in range(nbnd):
        pk = w.psi_k[v][idx[hit]]
        out += weight * n * (pk.conj()[None, :] * dpsi[v, :, : sq.npw][:, hit]).sum(1)
    return out
