"""Pins of the reference-form Sigma_c oracle (oracle/sigma_real_oracle.py)."""
import numpy as np

from oracle import green_oracle as go
from oracle import sigma_real_oracle as so

OMEGA = 270.011


def test_pade_reproduces_rational_function():
    rng = np.random.default_rng(0)
    K = 3
    poles = rng.standard_normal(K) + 1j * rng.standard_normal(K)
    res = rng.standard_normal(K) + 1j * rng.standard_normal(K)
    f = lambda x: (res[None, :] / (np.asarray(x)[:, None] - poles[None, :])).sum(1)  # noqa: E731
    z = 1j * np.linspace(0.1, 3.0, 8)
    a = so.pade_coeff(z, f(z))
    for w in (0.3 + 0.2j, -1.5 + 0.1j, 2.0 - 0.7j, 5j):
        assert abs(so.pade_eval(z, a, w) - f([w])[0]) < 1e-10 * abs(f([w])[0])
    # interpolation at the mesh
    for zi, ui in zip(z, f(z)):
        assert abs(so.pade_eval(z, a, zi) - ui) < 1e-12 * abs(ui)


def test_pade_even_symmetry_mesh():
    solver = np.array([0.0, 0.5j, 1.0j, 2.0j])
    z, src = so.freqbins_symm(solver, so.EVEN_SYMMETRY)
    assert np.allclose(z, [0, 0.5j, 1j, 2j, -0.5j, -1j, -2j])
    assert list(src) == [0, 1, 2, 3, 1, 2, 3]
    zs, _ = so.freqbins_symm(solver, so.SQUARE_SYMMETRY)
    assert np.allclose(zs, solver ** 2)
    # an even function is reproduced on the symmetrised mesh
    f = lambda x: 2.0 / (1.3 ** 2 - np.asarray(x) ** 2)          # noqa: E731
    w = np.zeros((z.size, 1, 1), dtype=np.complex128)
    w[: solver.size, 0, 0] = f(solver)
    coeff, zz = so.analytic_coeff(so.PADE_APPROX, w, solver, so.EVEN_SYMMETRY)
    out = so.analytic_eval(so.PADE_APPROX, coeff, zz, 0.7 + 0.3j, [0], [0], 16, 16)
    assert abs(out[0, 0] - f(0.7 + 0.3j)) < 1e-10


def test_godby_needs_recovers_single_pole():
    c1 = np.array([[0.8 - 0.2j, -0.3 + 0.1j]])
    c2 = np.array([[1.7 + 0.4j, 0.9 + 0.05j]])
    wp = 1.2
    W = lambda x: c1 * (1 / (c2 + x) + 1 / (c2 - x))             # noqa: E731
    w = np.stack([W(0.0), W(1j * wp)])
    coeff = so.godby_needs_coeffs(wp, w)
    assert np.abs(coeff[0] - c1).max() < 1e-13
    assert np.abs(coeff[1] - c2).max() < 1e-13
    x = 0.4 + 2.0j
    assert np.abs(so.godby_needs_model(x, coeff) - W(x)).max() < 1e-13
    # the reference's zero branch: W(0) = W(i wp)
    z = so.godby_needs_coeffs(wp, np.stack([W(0.0), W(0.0)]))
    assert np.all(z == 0)


def _sphere(nmax2):
    r = np.arange(-3, 4)
    m = np.array(np.meshgrid(r, r, r, indexing="ij")).reshape(3, -1).T
    m = m[(m ** 2).sum(1) <= nmax2]
    return m[np.lexsort((m[:, 2], m[:, 1], m[:, 0], (m ** 2).sum(1)))]


def test_sigma_prod_single_elements():
    n = 16
    g = _sphere(5)
    nl = go.box_index(g, n)
    pos = {tuple(m): i for i, m in enumerate(g)}
    a, ap, b, bp = 1, 2, 4, 7          # G_a = g[1], G'_a = g[2], G_b = g[4], G'_b = g[7]
    G = np.zeros((n ** 3, n ** 3), dtype=np.complex128)
    G[ap, a] = 1.0
    go.invfft6(G, n, n, nl, nl, OMEGA)
    W = np.zeros_like(G)
    W[bp, b] = 1.0
    alpha = 0.3 - 0.1j
    so.sigma_prod(n, n, nl, nl, OMEGA, alpha, G, W)
    r = pos[tuple(g[a] + g[b])]
    c = pos[tuple(g[ap] + g[bp])]
    want = np.zeros((len(g), len(g)), dtype=np.complex128)
    want[c, r] = alpha / OMEGA
    assert np.abs(W[: len(g), : len(g)] - want).max() < 1e-15


def test_sigma_correlation_linear_in_green_for_constant_w():
    n = 16
    g = _sphere(2)
    k = len(g)
    nl = go.box_index(g, n)
    rng = np.random.default_rng(2)
    coul = 1j * np.array([0.2, 1.5])
    weight = np.array([0.3, 0.7])
    green_r = (rng.standard_normal((4, n ** 3, n ** 3))
               + 1j * rng.standard_normal((4, n ** 3, n ** 3))) * 1e-2
    # a frequency-independent W: Pade of equal values
    solver = 1j * np.array([0.0, 0.5, 1.0, 2.0])
    wc = rng.standard_normal((k, k)) + 1j * rng.standard_normal((k, k))
    w = np.repeat(wc[None], solver.size, axis=0)
    coeff, z = so.analytic_coeff(so.PADE_APPROX, w, solver, so.NO_SYMMETRY)
    gm = np.arange(k)
    sig = so.sigma_correlation(n, n, nl, nl, OMEGA, 0.5, 0.1, coul, weight, [0.4j], green_r,
                               so.PADE_APPROX, coeff, z, so.NO_SYMMETRY, gm, gm)
    gsum = sum(weight[i % 2] * green_r[i] for i in range(4))
    arr = np.zeros((n ** 3, n ** 3), dtype=np.complex128)
    arr[:k, :k] = wc
    ref = so.sigma_prod(n, n, nl, nl, OMEGA, 0.5, gsum, arr)[:k, :k]
    assert np.abs(sig[0] - ref).max() < 1e-10 * np.abs(ref).max()
