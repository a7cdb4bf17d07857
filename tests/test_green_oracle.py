"""Pins of the Green's-function oracle (oracle/green_oracle.py):

* invfft6 of a single reciprocal-space element f(G_a, G'_b) = 1 is the
  closed form e^{-iG_a r} e^{+iG_b r'} / Omega on the two boxes;
* fwfft6(invfft6(F)) returns the (num_g, num_g_p) block of F;
* green_function equals the spectral sum -sum_n psi_n(G) psi_n(G')^* /
  (e_n - omega), looked up by explicit Miller searches, with the reference's
  mask (0 < map < num_g) leaving out the sphere's last plane wave and
  columns outside the sphere left zero.
"""
import numpy as np
import pytest

from oracle import green_oracle as go


def sphere(n_max_sq, rng_shuffle=None):
    r = np.arange(-3, 4)
    m = np.array(np.meshgrid(r, r, r, indexing="ij")).reshape(3, -1).T
    m = m[(m ** 2).sum(1) <= n_max_sq]
    order = np.lexsort((m[:, 2], m[:, 1], m[:, 0], (m ** 2).sum(1)))
    return m[order]


OMEGA = 270.011


def test_invfft6_plane_wave_closed_form():
    n, n_p = 16, 18
    g = sphere(3)
    gp = sphere(2)
    nl, nl_p = go.box_index(g, n), go.box_index(gp, n_p)
    a, b = 5, 3
    F = np.zeros((n_p ** 3, n ** 3), dtype=np.complex128)
    F[b, a] = 1.0                      # f(G_a, G'_b)
    go.invfft6(F, n, n_p, nl, nl_p, OMEGA)
    i = np.arange(n)
    i3, i2, i1 = np.meshgrid(i, i, i, indexing="ij")
    ph = np.exp(-2j * np.pi * (g[a, 0] * i1 + g[a, 1] * i2 + g[a, 2] * i3) / n).reshape(-1)
    j = np.arange(n_p)
    j3, j2, j1 = np.meshgrid(j, j, j, indexing="ij")
    php = np.exp(2j * np.pi * (gp[b, 0] * j1 + gp[b, 1] * j2 + gp[b, 2] * j3) / n_p).reshape(-1)
    want = php[:, None] * ph[None, :] / OMEGA
    assert np.abs(F - want).max() < 1e-14


def test_fft6_round_trip():
    n = n_p = 16
    g = sphere(5)
    nl = go.box_index(g, n)
    rng = np.random.default_rng(0)
    k = len(g)
    F = np.zeros((n_p ** 3, n ** 3), dtype=np.complex128)
    block = rng.standard_normal((k, k)) + 1j * rng.standard_normal((k, k))
    F[:k, :k] = block
    go.invfft6(F, n, n_p, nl, nl, OMEGA)
    go.fwfft6(F, n, n_p, nl, nl, OMEGA)
    assert np.abs(F[:k, :k] - block).max() < 1e-12 * np.abs(block).max()


def test_green_function_spectral_sum():
    rng = np.random.default_rng(1)
    sm = sphere(4)                      # the k-point's sphere
    ns = len(sm)
    h = rng.standard_normal((ns, ns)) + 1j * rng.standard_normal((ns, ns))
    h = h + h.conj().T
    corr = sphere(3)                    # correlation G list (inside the sphere)
    corr_par = np.vstack([sphere(2), [[3, 3, 3]]])  # one column outside the sphere
    corr = np.vstack([corr, sm[-1:]])   # include the sphere's last plane wave as a row
    omega = np.array([0.3 + 0.2j, -1.0 + 0.5j])
    n = n_p = 16
    green = go.green_function(h, sm, corr, corr_par, omega, n, n_p)
    e, u = np.linalg.eigh(h)
    pos = {tuple(m): i for i, m in enumerate(sm)}
    for f, om in enumerate(omega):
        gs = -(u / (e - om)) @ u.conj().T          # G(G, G') = -(H - omega)^-1
        for c, mc in enumerate(corr_par):
            for r, mr in enumerate(corr):
                i, j = pos.get(tuple(mr)), pos.get(tuple(mc))
                want = 0.0
                if i is not None and j is not None and i < ns - 1:
                    want = gs[i, j]
                assert abs(green[f, c, r] - want) < 1e-12, (f, c, r)
    # rows beyond the correlation list and columns beyond corr_par are zero
    assert np.all(green[:, len(corr_par):, :] == 0)
    assert np.all(green[:, :, len(corr):] == 0)
