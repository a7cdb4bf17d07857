"""GPU reference-form Sigma_c (analytic continuation, sigma_prod and the
frequency loops of sigma_correlation) against the CPU oracle."""
import numpy as np
import pytest
import torch

from oracle import green_oracle as go
from oracle import sigma_real_oracle as so
from sternheimergw_amd import _lib, green, sigma_real

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")
OMEGA = 270.011


def _sphere(nmax2):
    r = np.arange(-3, 4)
    m = np.array(np.meshgrid(r, r, r, indexing="ij")).reshape(3, -1).T
    m = m[(m ** 2).sum(1) <= nmax2]
    return m[np.lexsort((m[:, 2], m[:, 1], m[:, 0], (m ** 2).sum(1)))]


def _rand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("model,symm", [(_lib.GODBY_NEEDS, 0), (_lib.PADE_APPROX, 0),
                                        (_lib.PADE_APPROX, 1), (_lib.PADE_APPROX, 2)])
def test_analytic_matches_oracle(model, symm):
    print(f"analytic model={model} symm={symm}", flush=True)
    rng = np.random.default_rng(10 * model + symm)
    k = 27
    if model == _lib.GODBY_NEEDS:
        solver = np.array([0.0, 1.1j])
        # single-pole data with Re(c2) > 0 plus a few elements on the zero branch
        c1, c2 = _rand(rng, k, k), 1.0 + 0.5 * rng.random((k, k)) + 0.1j * rng.random((k, k))
        W = np.stack([c1 * 2 / c2, 2 * c1 * c2 / (c2 ** 2 + 1.1 ** 2)])
        W[1, :3, :3] = W[0, :3, :3]
        nz = 2
    else:
        solver = 1j * np.array([0.0, 0.3, 0.8, 1.5, 3.0]) if symm == 1 else \
            1j * np.array([0.1, 0.3, 0.8, 1.5, 3.0])
        z, _src = so.freqbins_symm(solver, symm)
        nz = z.size
        W = np.zeros((nz, k, k), dtype=np.complex128)
        W[: solver.size] = _rand(rng, solver.size, k, k)
    d_w = torch.from_numpy(np.ascontiguousarray(W)).to(DEV)
    scr = sigma_real.ScreenedCoulomb(d_w, solver, model, symm)
    coeff_ref, zref = so.analytic_coeff(model, W.copy(), solver, symm)
    assert np.abs(d_w.cpu().numpy()[: coeff_ref.shape[0]] - coeff_ref).max() <= \
        1e-12 * np.abs(coeff_ref).max()
    n = 16
    g = _sphere(2)
    grid = green.CorrGrid(n, g)
    gm = rng.permutation(k)[: grid.num_g].astype(np.int32)   # a symmetry-like remapping
    d_gm = torch.from_numpy(gm).to(DEV)
    out = torch.empty((n ** 3, n ** 3), dtype=torch.complex128, device=DEV)
    for freq in (0.2 + 0.4j, 1.3j, 0.7 + 0.05j):
        scr.eval_into(freq, grid, grid, d_gm, d_gm, out)
        ref = so.analytic_eval(model, coeff_ref, zref, so.symmetrize(freq, symm), gm, gm, n, n)
        got = out.cpu().numpy()
        assert np.abs(got - ref).max() <= 1e-11 * max(1.0, np.abs(ref).max())
    _ = nz


def test_sigma_prod_matches_oracle():
    n = 16
    g = _sphere(5)
    k = len(g)
    grid = green.CorrGrid(n, g)
    rng = np.random.default_rng(3)
    G = _rand(rng, n ** 3, n ** 3)
    W = np.zeros((n ** 3, n ** 3), dtype=np.complex128)
    W[:k, :k] = _rand(rng, k, k)
    d_g = torch.from_numpy(G).to(DEV)
    d_w = torch.from_numpy(W.copy()).to(DEV)
    sig = torch.zeros((k, k), dtype=torch.complex128, device=DEV)
    sigma_real.sigma_prod(grid, grid, OMEGA, 0.3 - 0.2j, d_g, d_w, sig)
    ref = so.sigma_prod(n, n, grid.nl, grid.nl, OMEGA, 0.3 - 0.2j, G, W.copy())[:k, :k]
    assert np.abs(d_w.cpu().numpy()[:k, :k] - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.abs(sig.cpu().numpy() - ref).max() <= 1e-12 * np.abs(ref).max()


def test_sigma_correlation_matches_oracle():
    n = 16
    g = _sphere(2)
    k = len(g)
    grid = green.CorrGrid(n, g)
    rng = np.random.default_rng(4)
    coul = 1j * np.array([0.2, 1.5])
    weight = np.array([0.3, 0.7])
    sigma_freq = np.array([-0.3, 0.4j])
    G = _rand(rng, 4, n ** 3, n ** 3) * 1e-2
    solver = 1j * np.array([0.0, 0.5, 1.0, 2.0])
    z, _src = so.freqbins_symm(solver, 1)
    W = np.zeros((z.size, k, k), dtype=np.complex128)
    W[: solver.size] = _rand(rng, solver.size, k, k)
    coeff_ref, zref = so.analytic_coeff(_lib.PADE_APPROX, W.copy(), solver, 1)
    d_w = torch.from_numpy(W).to(DEV)
    scr = sigma_real.ScreenedCoulomb(d_w, solver, _lib.PADE_APPROX, 1)
    d_g = torch.from_numpy(G).to(DEV)
    print("sigma_correlation 16^3, 2 x 4 products", flush=True)
    got = sigma_real.sigma_correlation(grid, grid, OMEGA, 0.5 + 0.1j, 0.05, coul, weight,
                                       sigma_freq, d_g, scr).cpu().numpy()
    gm = np.arange(k)
    ref = so.sigma_correlation(n, n, grid.nl, grid.nl, OMEGA, 0.5 + 0.1j, 0.05, coul, weight,
                               sigma_freq, G, _lib.PADE_APPROX, coeff_ref, zref, 1, gm, gm)
    assert np.abs(got - ref).max() <= 1e-10 * np.abs(ref).max()
