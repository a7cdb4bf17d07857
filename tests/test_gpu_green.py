"""GPU Green's function of the self-energy: sgw_invfft6 / sgw_fwfft6 against
the oracle's fft6, and green_function (batched solve + sgw_green_assemble)
against the oracle's green_function on the 24^3 cell, then through invfft6
to G(r, r'; omega)."""
import numpy as np
import pytest
import torch

from oracle import green_oracle as go
from sternheimergw_amd import cell, green, operators

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")
OMEGA = 270.011


def sphere(n_max_sq):
    r = np.arange(-4, 5)
    m = np.array(np.meshgrid(r, r, r, indexing="ij")).reshape(3, -1).T
    m = m[(m ** 2).sum(1) <= n_max_sq]
    order = np.lexsort((m[:, 2], m[:, 1], m[:, 0], (m ** 2).sum(1)))
    return m[order]


@pytest.mark.parametrize("n,n_p", [(16, 16), (18, 16), (24, 18), (16, 27)])
def test_fft6_matches_oracle(n, n_p):
    print(f"fft6 n={n} n_p={n_p}", flush=True)
    g, gp = sphere(6), sphere(5)
    grid, grid_p = green.CorrGrid(n, g), green.CorrGrid(n_p, gp)
    rng = np.random.default_rng(n + n_p)
    F = np.zeros((2, n_p ** 3, n ** 3), dtype=np.complex128)
    F[:, : len(gp), : len(g)] = (rng.standard_normal((2, len(gp), len(g)))
                                 + 1j * rng.standard_normal((2, len(gp), len(g))))
    d = torch.from_numpy(F.copy()).to(DEV)
    green.invfft6(d, grid, OMEGA, grid_p)
    got = d.cpu().numpy()
    ref = np.stack([go.invfft6(F[m].copy(), n, n_p, grid.nl, grid_p.nl, OMEGA) for m in range(2)])
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    # and back: the reciprocal block is recovered
    green.fwfft6(d, grid, OMEGA, grid_p)
    back = d.cpu().numpy()[:, : len(gp), : len(g)]
    assert np.abs(back - F[:, : len(gp), : len(g)]).max() <= 1e-12 * np.abs(F).max()


def test_green_function_matches_oracle(c1):
    w = c1
    op = operators.PwOperator.from_workload(w, with_valence=False)
    sq = w.spheres[0]
    corr = sphere(3)
    corr_par = np.vstack([sphere(2), [[9, 9, 9]]])        # last column outside the sphere
    corr = np.vstack([corr, sq.mill[-1:]])                 # the sphere's last plane wave as a row
    omega = np.array([0.1 + 0.3j, -0.5 + 0.2j, 0.4j])
    n = 16
    grid, grid_p = green.CorrGrid(n, corr), green.CorrGrid(n, corr_par)
    print("green_function on the 24^3 cell", flush=True)
    d_green = green.green_function(op, 0, sq.mill, omega, grid, grid_p, threshold=1e-12)
    got = d_green.cpu().numpy()
    ref = go.green_function(w.dense_h[0], sq.mill, corr, corr_par, omega, n, n)
    assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max()
    # to real space (sigma.f90:688-693)
    green.invfft6(d_green, grid, OMEGA, grid_p)
    ref_r = np.stack([go.invfft6(ref[f].copy(), n, n, grid.nl, grid_p.nl, OMEGA)
                      for f in range(omega.size)])
    assert np.abs(d_green.cpu().numpy() - ref_r).max() <= 1e-9 * np.abs(ref_r).max()
