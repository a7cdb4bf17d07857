"""CPU oracle of the reference form of the correlation self-energy (test
infrastructure only):

* freqbins_symm (algo/grid/src/freqbins.f90:333-425) and the symmetrised
  output frequency (:451-474);
* godby_needs_coeffs / godby_needs_model (algo/analytic/src/godby_needs.f90);
* pade_coeff / pade_eval (algo/analytic/src/pade.f90:41-147, Vidberg and
  Serene, J. Low Temp. Phys. 29, 179 (1977)), with `protect`;
* analytic_coeff / analytic_eval (algo/analytic/src/analytic.f90:50-348) for
  those two models, the robust Pade approximant (pade_robust_oracle.py) and
  AAA (aaa_oracle.py);
* sigma_prod (phys/corr/src/sigma.f90:417-523) on the oracle's fft6;
* the frequency loops of sigma_correlation (sigma.f90:653-745).

Pins (tests/test_sigma_real_oracle.py): the Pade approximant reproduces a
rational function with K <= N/2 poles everywhere and interpolates its
points; Godby-Needs recovers the pole of a single-pole W exactly; sigma_prod
of single reciprocal-space elements is the closed form alpha/Omega at
(G_a + G_b, G'_a + G'_b); the loop is linear in G for a frequency-constant W.
"""
from __future__ import annotations

import numpy as np

from . import aaa_oracle as aa
from . import green_oracle as go
from . import pade_robust_oracle as pr

GODBY_NEEDS, PADE_APPROX, PADE_ROBUST, AAA_APPROX, AAA_POLE = 1, 2, 3, 4, 5
NO_SYMMETRY, EVEN_SYMMETRY, SQUARE_SYMMETRY = 0, 1, 2
EPS8, EPS14, EPS24 = 1e-8, 1e-14, 1e-24


def freqbins_symm(solver, symm):
    """(z, src): the symmetrised mesh and, for each of its points, the
    index of the solver frequency whose W it takes (freqbins.f90:333-425)."""
    solver = np.asarray(solver, dtype=np.complex128)
    n = solver.size
    if symm == NO_SYMMETRY:
        return solver.copy(), np.arange(n, dtype=np.int32)
    if symm == SQUARE_SYMMETRY:
        return solver ** 2, np.arange(n, dtype=np.int32)
    zero = np.abs(solver) < EPS14
    if zero.sum() > 1:
        raise ValueError("only a single frequency may be smaller than 1e-14")
    z = list(solver)
    src = list(range(n))
    for i in range(n):
        if not zero[i]:
            z.append(-solver[i])
            src.append(i)
    return np.array(z), np.array(src, dtype=np.int32)


def symmetrize(freq, symm):
    return freq ** 2 if symm == SQUARE_SYMMETRY else freq


def _protect(x):
    return np.where(np.abs(x) > EPS24, x, EPS24 + 0j)


def pade_coeff(z, u):
    """u (N, ...) values at z (N,) -> a (N, ...), element-wise."""
    z = np.asarray(z)
    N = z.size
    g = np.zeros((N,) + u.shape, dtype=np.complex128)     # g[p, i, ...]
    a = np.zeros((N,) + u.shape[1:], dtype=np.complex128)
    for p in range(N):                                       # pade.f90:73-87
        if p == 0:
            g[0] = _protect(u)
        else:
            for i in range(p, N):
                tmp1 = g[p - 1, p - 1] / g[p - 1, i]
                tmp2 = g[p - 1, i] / g[p - 1, i]
                g[p, i] = _protect((tmp1 - tmp2) / (z[i] - z[p - 1]))
        a[p] = g[p, p]
    return a


def pade_eval(z, a, w):
    """Continued fraction of pade.f90:128-138 at w, element-wise over a[p]."""
    N = len(z)
    acap_m2 = np.zeros(a.shape[1:], dtype=np.complex128)
    acap_m1 = a[0].astype(np.complex128)
    bcap_m2 = np.ones(a.shape[1:], dtype=np.complex128)
    bcap_m1 = np.ones(a.shape[1:], dtype=np.complex128)
    for i in range(1, N):
        acap = acap_m1 + (w - z[i - 1]) * a[i] * acap_m2
        bcap = bcap_m1 + (w - z[i - 1]) * a[i] * bcap_m2
        acap_m2, acap_m1, bcap_m2, bcap_m1 = acap_m1, acap, bcap_m1, bcap
    return acap_m1 / bcap_m1


def godby_needs_coeffs(omega_p, w):
    """w (2, ...) -> coefficients (2, ...) (godby_needs.f90)."""
    if omega_p < 0:
        raise ValueError("plasmon frequency must be positive")
    out = np.zeros_like(w, dtype=np.complex128)
    work = w[0] - w[1]
    ok = ~(np.abs(work.real) < EPS8)
    ratio = np.where(ok, w[1] / np.where(ok, work, 1.0), 0.0)
    ok &= ~(ratio.real < EPS8)
    c2 = np.sqrt(ratio + 0j) * omega_p
    out[1] = np.where(ok, c2, 0.0)
    out[0] = np.where(ok, 0.5 * w[0] * c2, 0.0)
    return out


def godby_needs_model(freq, coeff):
    c1, c2 = coeff[0], coeff[1]
    with np.errstate(divide="ignore", invalid="ignore"):
        res = c1 * (1.0 / (c2 + freq) + 1.0 / (c2 - freq))
    return np.where(np.abs(c1) > EPS8, res, 0.0)


def analytic_coeff(model, w, solver, symm=NO_SYMMETRY, thres=0.0):
    """w (nfreq_sym, num_g, num_g) with the first len(solver) slices filled
    (device layout [f][igp][ig]); returns (coeff, z)."""
    if model == GODBY_NEEDS:
        return godby_needs_coeffs(np.asarray(solver)[1].imag, w[:2]), np.asarray(solver)[:2]
    if model == PADE_ROBUST:                      # analytic.f90:158-160, no symmetrised mesh
        solver = np.asarray(solver, dtype=np.complex128)
        n = solver.size
        coeff = pr.pade_coeff_robust(solver, np.moveaxis(w[:n], 0, -1))
        return np.ascontiguousarray(np.moveaxis(coeff, -1, 0)), solver
    z, src = freqbins_symm(solver, symm)
    if model in (AAA_APPROX, AAA_POLE):            # analytic.f90:162-199
        vals = w[src]
        out = np.zeros_like(vals)
        for c in range(vals.shape[1]):
            for r in range(vals.shape[2]):
                out[:, c, r] = aa.aaa_coeff(model, thres, z, vals[:, c, r])
        return out, z
    return pade_coeff(z, w[src]), z


def analytic_eval(model, coeff, z, freq, gmap, gmap_p, n, n_p):
    """(n_p^3, n^3) array, device layout [c][r], zero outside the block."""
    out = np.zeros((n_p ** 3, n ** 3), dtype=np.complex128)
    sub = coeff[:, np.asarray(gmap_p)[:, None], np.asarray(gmap)[None, :]]
    if model == GODBY_NEEDS:
        val = godby_needs_model(freq, sub)
    elif model in (AAA_APPROX, AAA_POLE):
        flat = np.moveaxis(sub, 0, -1)
        val = np.array([[aa.aaa_eval(model, flat[i, j], freq) for j in range(flat.shape[1])]
                        for i in range(flat.shape[0])])
    elif model == PADE_ROBUST:
        flat = np.moveaxis(sub, 0, -1)
        val = np.array([[pr.pade_eval_robust(flat[i, j], freq) for j in range(flat.shape[1])]
                        for i in range(flat.shape[0])])
    else:
        val = pade_eval(z, sub, freq)
    out[: len(gmap_p), : len(gmap)] = val
    return out


def sigma_prod(n, n_p, nl, nl_p, omega, alpha, green_r, array):
    """array (n_p^3, n^3): W(G, G') block in, alpha Sigma block out."""
    go.invfft6(array, n, n_p, nl, nl_p, omega)
    array *= green_r
    go.fwfft6(array, n, n_p, nl, nl_p, omega)
    k, kp = len(nl), len(nl_p)
    array[:kp, :k] *= alpha
    return array


def sigma_correlation(n, n_p, nl, nl_p, omega, alpha, mu, coul, weight, sigma_freq, green_r,
                      model, coeff, z, symm, gmap, gmap_p):
    """sigma[isigma] (num_g_p, num_g) of sigma.f90:653-745; green_r
    (2 num_coul, n_p^3, n^3) at freq_green = [mu + coul, mu - coul]."""
    coul = np.asarray(coul, dtype=np.complex128)
    num_coul = coul.size
    freq_green = np.concatenate([mu + coul, mu - coul])            # freqbins.f90:315-321
    freq_sigma = mu + np.asarray(sigma_freq, dtype=np.complex128)  # sigma.f90:660
    k, kp = len(nl), len(nl_p)
    out = np.zeros((freq_sigma.size, kp, k), dtype=np.complex128)
    for isig, fs in enumerate(freq_sigma):
        for ig, fg in enumerate(freq_green):
            fc = fs - fg                                             # :709
            if fc.real * fc.imag < 0.0:                              # :711
                fc = np.conj(fc)
            work = analytic_eval(model, coeff, z, symmetrize(fc, symm), gmap, gmap_p, n, n_p)
            icoul = ig % num_coul                                    # :722
            sigma_prod(n, n_p, nl, nl_p, omega, alpha * weight[icoul], green_r[ig], work)
            out[isig] += work[:kp, :k]                               # :741
    return out
