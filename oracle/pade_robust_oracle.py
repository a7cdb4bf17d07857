"""CPU oracle of the robust Pade continuation (test infrastructure only):
pade_coeff_robust / pade_robust / pade_eval_robust of
algo/analytic/src/pade_robust.f90, the algorithm of P. Gonnet, S. Guettel
and L. N. Trefethen, "Robust Pade approximation via SVD", SIAM Rev. 55,
101-117 (2013).  numpy's SVD and QR call the LAPACK routines the reference's
wrappers call (ZGESDD; ZGEQRF + ZUNGQR, data/algebra/src/lapack.f90:34-225).

Steps for the values f_j on the circle z_j = R e^{2 pi i j / N}:

1. Taylor coefficients c_k = (1/N) sum_j f_j e^{-2 pi i j k / N} / R^k,
   k < m + n + 1 (pade_derivative, :440-500); entries below 1e-14 ||c|| are
   dropped, and c is made real when its imaginary part is below that;
2. the (m, n) = (N/2 - 2, N/2 - 2) approximant; zero function if
   ||c(0:m)||_inf <= 1e-14 ||c||_inf (:287-306);
3. diagonal hopping (:318-346): the rows m+1..m+n of the Toeplitz matrix
   of c (first row c_0, 0, ...) have numerical rank rho (singular values
   above 1e-14 ||c||_2); while rho < n, m -= n - rho, n = rho;
4. (:351-423) b from the reweighted QR of (C D)^T with D = |v| + sqrt(eps),
   v the last right singular vector of C, b = D q_{n+1}, normalised;
   a = Z(0:m, 0:n) b; leading zeros of b (and a) and trailing zeros of b
   dropped (1e-14);
5. trailing zeros of a dropped (1e-14 ||c||); a, b divided by b_0.

q_{n+1} of the QR of the plain transpose satisfies (C D) conj(q) = 0, so for
complex c the b of step 4 is the conjugate of C's null vector; for real c
(real functions on the circle) it is the null vector.  The restatement keeps
that.  The reference runs step 4 only for deg_num > 1 and leaves the
coefficients unset when the hopping ends with deg_num = deg_den = 1 (a
degree-(1,1) function); here, as in the published algorithm, step 4 runs for
every deg_den > 0.  Parity for that case is unpinned (the reference has no
defined result).

Coefficient layout (pade_coeff_robust, :143-160): [m, n, a_0..a_m, b_0..b_n]
in the frequency slices; pade_eval_robust (:37-86) evaluates
(sum a_i z^i) / (sum b_i z^i) by Horner.

Pins (tests/test_pade_robust_oracle.py): real rational functions of low
degree are recovered (degrees and values); the hopping reduces the degree of
a polynomial and of a lower-degree rational function to the true one; the
zero function gives (0, 0); values on and inside the circle match.
"""
from __future__ import annotations

import numpy as np

EPS14 = 1e-14
EPS12 = 1e-12
EPS = np.finfo(np.float64).eps


def _derivatives(radius, func, tol, num_deriv):
    n = func.size
    work = np.fft.fft(func) / n                      # QE cft_1z, isign = -1
    deriv = np.zeros(num_deriv, dtype=np.complex128)
    k = min(num_deriv, n)
    deriv[:k] = work[:k]
    if abs(radius - 1.0) > EPS14:
        deriv[:k] = deriv[:k] * radius ** (-np.arange(k, dtype=np.float64))
    abs_tol = tol * np.linalg.norm(deriv)
    deriv[np.abs(deriv) < abs_tol] = 0.0
    if np.max(np.abs(deriv.imag), initial=0.0) < abs_tol:
        deriv = deriv.real.astype(np.complex128)
    return deriv


def _toeplitz(col, ncol):
    nrow = col.size
    z = np.zeros((nrow, ncol), dtype=np.complex128)
    for j in range(ncol):
        z[j:, j] = col[: nrow - j]
    return z


def pade_robust(radius, func, deg_num, deg_den, rel_tol=EPS14):
    """(a, b) with deg a = m, deg b = n (pade_robust.f90:164-442)."""
    func = np.asarray(func, dtype=np.complex128)
    m, n = int(deg_num), int(deg_den)
    coeff = _derivatives(radius, func, rel_tol, m + n + 1)
    abs_tol = rel_tol * np.linalg.norm(coeff)
    if np.max(np.abs(coeff[: m + 1])) <= rel_tol * np.max(np.abs(coeff)):
        return np.zeros(1, dtype=np.complex128), np.ones(1, dtype=np.complex128)
    while True:
        if n == 0:
            return coeff[: m + 1].copy(), np.ones(1, dtype=np.complex128)
        z = _toeplitz(coeff[: m + n + 1], n + 1)
        c = z[m + 1: m + n + 1, :]
        sigma = np.linalg.svd(c, compute_uv=False)
        rho = int(np.sum(sigma > abs_tol))
        if rho == n:
            break
        m -= n - rho
        n = rho
    # (the reference runs this for deg_num > 1 only and leaves the
    # coefficients unset otherwise; the published algorithm runs it for every
    # n > 0, which is what happens here)
    _, _, vh = np.linalg.svd(c)
    b = vh[n, :]                                     # row n+1 of V^H (ZGESDD)
    d = np.abs(b) + np.sqrt(EPS)
    q, _ = np.linalg.qr((c * d[None, :]).T, mode="complete")
    b = d * q[:, n]
    b = b / np.linalg.norm(b)
    a = z[: m + 1, : n + 1] @ b
    nz = np.nonzero(np.abs(b) > rel_tol)[0]
    lam = int(nz[0]) if nz.size else 0
    if lam > 0:
        m -= lam
        n -= lam
        a = a[lam:]
        b = b[lam:]
    nz = np.nonzero(np.abs(b) > rel_tol)[0]
    last = int(nz[-1]) if nz.size else -1
    if last != n:
        n = last
        b = b[: n + 1]
    nz = np.nonzero(np.abs(a) > abs_tol)[0]
    last = int(nz[-1]) if nz.size else -1
    if last != m:
        m = last
        a = a[: m + 1]
    return a / b[0], b / b[0]


def pade_coeff_robust(freq, func):
    """Coefficient slices of every element (pade_coeff_robust, :89-160):
    func (..., N) on the circle freq -> (..., N) = [m, n, a, b, 0...]."""
    freq = np.asarray(freq, dtype=np.complex128)
    n_f = freq.size
    if n_f < 10:
        raise ValueError("use at least 10 frequencies to form the circle")
    radius = freq[0].real
    if np.any(np.abs(np.abs(freq) - radius) > EPS12):
        raise ValueError("frequencies must span circle in the complex plane")
    func = np.asarray(func, dtype=np.complex128)
    out = np.zeros_like(func)
    flat_in = func.reshape(-1, n_f)
    flat_out = out.reshape(-1, n_f)
    for e in range(flat_in.shape[0]):
        deg = n_f // 2 - 2
        a, b = pade_robust(radius, flat_in[e], deg, deg)
        flat_out[e, 0] = a.size - 1
        flat_out[e, 1] = b.size - 1
        flat_out[e, 2: 2 + a.size] = a
        flat_out[e, 2 + a.size: 2 + a.size + b.size] = b
    return out


def pade_eval_robust(coeff, freq):
    """pade_eval_robust (:37-86) for one coefficient slice vector."""
    m = int(round(abs(coeff[0])))
    n = int(round(abs(coeff[1])))
    a = coeff[2: 3 + m]
    b = coeff[3 + m: 4 + m + n]
    num = 0j
    for c in a[::-1]:
        num = c + num * freq
    den = 0j
    for c in b[::-1]:
        den = c + den * freq
    return num / den


def circle(radius, n):
    """n equally spaced points on the circle of radius R, the first (R, 0)."""
    return radius * np.exp(2j * np.pi * np.arange(n) / n)
