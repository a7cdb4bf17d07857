"""CPU oracle of the AAA continuation (test infrastructure only): the
vendored AAA library vendor/analytic/src/aaa.f90 (Nakatsukasa, Sete and
Trefethen, "The AAA algorithm for rational approximation", SIAM J. Sci.
Comput. 40, A1494 (2018)) and its two uses in algo/analytic/src/analytic.f90:
the barycentric coefficients (aaa_approx, :162-193, eval :351-386) and the
pole-residue form (aaa_pole, pole_correction :389-427, eval :430-459).

aaa_generate (aaa.f90:53-302), for values f on the points z:
  thres_abs = thres * max |f|; fit = mean(f) everywhere; repeat
    j = first index of max |f - fit|; z_j becomes a support point (fit_j = f_j);
    weights w = the last right singular vector of the Loewner submatrix
      L(i, j) = (f_i - f_j) / (z_i - z_j), rows the other points, columns the
      support points (ZGESDD; an empty matrix gives the identity, so w = e_m);
    fit = N / D on the other points, N, D = sum_j w_j (f_j, 1) / (z - z_j)
      (1 / eps14 where |z - z_j| <= eps14, construct_Cauchy_matrix);
  until every |fit - f| <= thres_abs or max_point support points.
aaa_evaluate (:71-90) is N / D, with f_j returned within eps14 of a support
point.  Poles (:387-455): the finite eigenvalues alpha / beta (|beta| > eps14)
of the (m+1) pencil A = [[0, 1^T], [w, diag(z_j)]], B = diag(0, 1, ..., 1)
(ZGGEV, here scipy.linalg.eig); residues (:458-500): the mean of
r(p + s) s over s = +-1e-6, +-1e-6 i.

numpy's SVD is ZGESDD and scipy's generalized eigensolver ZGGEV, the LAPACK
routines the reference's wrappers call (vendor/analytic/src/lapack.f90).

Pins (tests/test_aaa_oracle.py): the reference's own test cases
(vendor/analytic/test/testAAA.pf, values in tests/golden/aaa_reference.json):
the GW example's support points and weights, the tangent's support points,
values, weights and evaluations, the tangent's poles and residues, the
threshold-to-step table, constant and rational functions, max_point.
"""
from __future__ import annotations

import numpy as np

EPS6, EPS12, EPS14 = 1e-6, 1e-12, 1e-14
SHIFTS = np.array([EPS6, EPS6 * 1j, -EPS6, -EPS6 * 1j])


def loewner(z, f):
    """L(i, j) = (f_i - f_j) / (z_i - z_j), zero diagonal (aaa.f90:331-351)."""
    z, f = np.asarray(z, np.complex128), np.asarray(f, np.complex128)
    n = z.size
    L = np.zeros((n, n), np.complex128)
    for j in range(n):
        for i in range(j + 1, n):
            L[i, j] = (f[i] - f[j]) / (z[i] - z[j])
            L[j, i] = L[i, j]
    return L


def cauchy(x, y):
    """1 / (x_i - y_j), eps14 where |x_i - y_j| <= eps14 (aaa.f90:308-327)."""
    d = np.asarray(x, np.complex128)[:, None] - np.asarray(y, np.complex128)[None, :]
    d = np.where(np.abs(d) <= EPS14, EPS14 + 0j, d)
    return 1.0 / d


def evaluate_raw(pos, val, wgt, z):
    c = cauchy(np.atleast_1d(z), pos)
    return (c @ (wgt * val)) / (c @ wgt)


def aaa_evaluate(pos, val, wgt, z):
    """aaa_evaluate (aaa.f90:71-90)."""
    z = np.atleast_1d(np.asarray(z, np.complex128))
    out = evaluate_raw(pos, val, wgt, z)
    for i in range(z.size):
        d = np.abs(z[i] - pos)
        k = int(np.argmin(d))
        if d[k] < EPS14:
            out[i] = val[k]
    return out


def aaa_generate(thres, max_point, z, f):
    """(position, value, weight) of aaa_generate (aaa.f90:53-302); max_point
    -1 means no restriction."""
    z, f = np.asarray(z, np.complex128), np.asarray(f, np.complex128)
    n = z.size
    if f.size != n:
        raise ValueError("input_error: points and values differ in size")
    thres_abs = thres * np.max(np.abs(f))
    mp = n if max_point == -1 else max_point
    fit = np.full(n, f.sum() / n, np.complex128)
    sup = np.zeros(n, bool)
    L = loewner(z, f)
    while True:
        j = int(np.argmax(np.abs(f - fit)))
        fit[j] = f[j]
        sup[j] = True
        pos, val = z[sup], f[sup]
        sub = L[np.ix_(~sup, sup)]
        if sub.size == 0:
            vh = np.eye(int(sup.sum()), dtype=np.complex128)
        else:
            _, _, vh = np.linalg.svd(sub)          # ZGESDD, V^H (m x m)
        wgt = np.conj(vh[-1, :])
        ev = evaluate_raw(pos, val, wgt, z)
        fit[~sup] = ev[~sup]
        if np.all(np.abs(fit - f) <= thres_abs) or pos.size >= mp:
            return pos.copy(), val.copy(), wgt


def poles(pos, wgt):
    """find_pole (aaa.f90:387-455): finite generalized eigenvalues, ZGGEV order."""
    import scipy.linalg
    m = pos.size + 1
    A = np.zeros((m, m), np.complex128)
    B = np.zeros((m, m), np.complex128)
    A[0, 1:] = 1.0
    A[np.arange(1, m), np.arange(1, m)] = pos
    A[1:, 0] = wgt
    B[np.arange(1, m), np.arange(1, m)] = 1.0
    ab = scipy.linalg.eig(A, B, right=False, homogeneous_eigvals=True)
    alpha, beta = ab[0], ab[1]
    keep = np.abs(beta) > EPS14
    return alpha[keep] / beta[keep]


def residues(pos, val, wgt, pole):
    """calculate_residual (aaa.f90:458-500)."""
    pole = np.asarray(pole, np.complex128)
    zz = (pole[:, None] + SHIFTS[None, :]).ravel()
    ff = aaa_evaluate(pos, val, wgt, zz).reshape(pole.size, SHIFTS.size)
    return (ff * SHIFTS[None, :]).sum(axis=1) / SHIFTS.size


def pole_correction(thres, pos, val, wgt, nslice):
    """pole_correction (analytic.f90:389-427): [poles, residues] of the
    residues above thres, each half nslice // 2 long."""
    p = poles(pos, wgt)
    r = residues(pos, val, wgt, p)
    half = nslice // 2
    keep = np.abs(r) > thres
    if keep.sum() > half:
        raise ValueError("two many relevant poles")
    out = np.zeros(nslice, np.complex128)
    k = int(keep.sum())
    out[:k] = p[keep]
    out[half: half + k] = r[keep]
    return out


def aaa_coeff(model, thres, z, f):
    """One element's coefficient slices (analytic.f90:162-199): model 4
    [position, value, weight] in thirds of mmax = N // 3 each; model 5 the
    pole-residue halves."""
    n = z.size
    if model == 4:
        mmax = n // 3
        pos, val, wgt = aaa_generate(thres, mmax, z, f)
        out = np.zeros(n, np.complex128)
        mm = pos.size
        out[:mm] = pos
        out[mmax: mmax + mm] = val
        out[2 * mmax: 2 * mmax + mm] = wgt
        return out
    pos, val, wgt = aaa_generate(thres, n, z, f)
    return pole_correction(thres, pos, val, wgt, n)


def aaa_eval(model, coeff, freq):
    """aaa_approx_eval (analytic.f90:351-386) or aaa_pole_eval (:430-459)."""
    coeff = np.asarray(coeff, np.complex128)
    n = coeff.size
    if model == 4:
        mmax = n // 3
        mm = int(np.count_nonzero(np.abs(coeff[2 * mmax: 3 * mmax]) > EPS12))
        return aaa_evaluate(coeff[:mm], coeff[mmax: mmax + mm],
                            coeff[2 * mmax: 2 * mmax + mm], freq)[0]
    half = n // 2
    npole = int(np.count_nonzero(np.abs(coeff[half:]) > 0.0))
    res = 0j
    for i in range(npole):
        res += coeff[half + i] / (freq - coeff[i])
    return res
