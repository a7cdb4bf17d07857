"""CPU oracle of the Sigma_c frequency convolution (test infrastructure only).

    Sigma[i] = alpha * sum_j w_j G[i][j] @ W[j]

This is the restatement of the frequency sum of sigma_correlation
(phys/corr/src/sigma.f90:705-745: sigma(:,:,isigma) += alpha * w * [G (x) W])
as a matrix product over the G'' index.  The reference forms the product in
real space (sigma_prod, sigma.f90:417-523), so parity with the reference's own
numbers is unpinned; the oracle is pinned by explicit loops on small sizes
and by the closed forms in tests/test_sigma_oracle.py.

It also restates the seeded synthetic inputs that the library generates on
the device (sgw_sigma_synthetic_device): a splitmix64 counter hash for G and a
Gaussian-kernel W.  Matrices here are numpy arrays M[a, b] (row a, column b);
the device stores element (a, b) at a + b * ld.
"""
from __future__ import annotations

import numpy as np

_MASK = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(z: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser (Steele, Lea, Flood 2014) on uint64 arrays."""
    with np.errstate(over="ignore"):
        z = (z + np.uint64(0x9E3779B97F4A7C15)) & _MASK
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _MASK
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _MASK
        return z ^ (z >> np.uint64(31))


def unit_sym(z: np.ndarray) -> np.ndarray:
    """Top 53 bits of a uint64 as a double in [-1, 1)."""
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 / 9007199254740992.0) - 1.0


def gen_g(n: int, i: int, j: int, nout: int, nin: int, seed: int, nu_ij: float) -> np.ndarray:
    """G[i][j](a, b) = (u(2c) + i u(2c + 1)) / ((1 + nu_ij) sqrt(n)),
    c = (((seed * nout + i) * nin + j) * n + b) * n + a."""
    a = np.arange(n, dtype=np.uint64)[:, None]
    b = np.arange(n, dtype=np.uint64)[None, :]
    with np.errstate(over="ignore"):
        base = ((np.uint64(seed) * np.uint64(nout) + np.uint64(i)) * np.uint64(nin)
                + np.uint64(j)) & _MASK
        ctr = ((base * np.uint64(n) + b) * np.uint64(n) + a) & _MASK
        re = unit_sym(splitmix64((np.uint64(2) * ctr) & _MASK))
        im = unit_sym(splitmix64((np.uint64(2) * ctr + np.uint64(1)) & _MASK))
    sc = 1.0 / ((1.0 + nu_ij) * np.sqrt(float(n)))
    return (re * sc) + 1j * (im * sc)


def gen_w(gvec: np.ndarray, theta: np.ndarray, cw_j: float, length: float) -> np.ndarray:
    """W(a, b) = cw_j exp(-|g_a - g_b|^2 / (2 L^2)) e^{i(th_a - th_b)} / ((1 + |g_a|^2)(1 + |g_b|^2))."""
    d = gvec[:, None, :] - gvec[None, :, :]
    g2 = (gvec ** 2).sum(1)
    mag = cw_j / ((1.0 + g2)[:, None] * (1.0 + g2)[None, :]) * np.exp(
        -(d ** 2).sum(2) * (0.5 / (length * length)))
    ph = theta[:, None] - theta[None, :]
    return mag * np.cos(ph) + 1j * mag * np.sin(ph)


def sigma_conv(G: np.ndarray, W: np.ndarray, weight: np.ndarray, alpha: complex) -> np.ndarray:
    """G (nout, nin, n, n), W (nin, n, n), weight (nin,) -> Sigma (nout, n, n), FP64."""
    nout, nin = G.shape[:2]
    out = np.zeros((nout,) + G.shape[2:], dtype=np.complex128)
    for i in range(nout):
        for j in range(nin):
            out[i] += weight[j] * (G[i, j] @ W[j])
    return alpha * out


def sigma_conv_loops(G, W, weight, alpha):
    """Explicit loops of the same sum (pins sigma_conv on tiny sizes)."""
    nout, nin, n, _ = G.shape
    out = np.zeros((nout, n, n), dtype=np.complex128)
    for i in range(nout):
        for a in range(n):
            for b in range(n):
                s = 0j
                for j in range(nin):
                    for c in range(n):
                        s += weight[j] * G[i, j, a, c] * W[j, c, b]
                out[i, a, b] = alpha * s
    return out
