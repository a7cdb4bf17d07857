"""GPU robust Pade continuation (pade_coeff_robust / pade_eval_robust of
algo/analytic/src/pade_robust.f90 through sgw_analytic_coeff/eval with
SGW_PADE_ROBUST) against the CPU oracle, on elements of every kind the
algorithm separates: generic complex values (full rank), real and complex
rational functions of low degree (diagonal hopping), polynomials, constants
and zeros."""
import numpy as np
import pytest
import torch

from oracle import sigma_real_oracle as so
from sternheimergw_amd import _lib, green, sigma_real

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _sphere(nmax2):
    r = np.arange(-3, 4)
    m = np.array(np.meshgrid(r, r, r, indexing="ij")).reshape(3, -1).T
    m = m[(m ** 2).sum(1) <= nmax2]
    return m[np.lexsort((m[:, 2], m[:, 1], m[:, 0], (m ** 2).sum(1)))]


def _element(rng, kind, z, radius):
    """(values on the circle z, the function) of one element of the kind"""
    npts = z.size
    if kind == 0:                                   # generic complex values
        v = rng.standard_normal(npts) + 1j * rng.standard_normal(npts)
        return v, None
    if kind in (1, 2):                              # rational, poles >= 3R away
        dn, dd = rng.integers(0, 4, size=2)
        cplx = kind == 2
        poles = 3 * radius * (1 + rng.random(dd)) * np.exp(
            1j * np.pi * (rng.random(dd) if cplx else rng.integers(0, 2, dd)))
        b = np.polynomial.polynomial.polyfromroots(poles) if dd else np.ones(1)
        b = b / b[0]
        a = rng.standard_normal(dn + 1) + (1j * rng.standard_normal(dn + 1) if cplx else 0)
        if not cplx:
            b = b.real
        fn = lambda x, a=a, b=b: np.polyval(a[::-1], x) / np.polyval(b[::-1], x)
        return fn(z), fn
    if kind == 3:                                   # real polynomial
        a = rng.standard_normal(rng.integers(1, 6))
        fn = lambda x, a=a: np.polyval(a[::-1], x) + 0j
        return fn(z), fn
    if kind == 4:                                   # constant
        c = rng.standard_normal() + 1j * rng.standard_normal()
        return np.full(npts, c), lambda x, c=c: c
    if kind == 5:                                   # real Taylor coefficients, full rank
        c = rng.standard_normal(npts) * 0.8 ** np.arange(npts)
        return np.polyval(c[::-1], z / radius), None
    return np.zeros(npts, np.complex128), lambda x: 0j   # zero function


@pytest.mark.parametrize("npts,radius", [(16, 1.0), (32, 2.0), (64, 1.0), (11, 0.5)])
def test_robust_coeff_and_eval_match_oracle(npts, radius):
    print(f"pade_robust npts={npts} radius={radius}", flush=True)
    rng = np.random.default_rng(npts)
    k = 21
    z = radius * np.exp(2j * np.pi * np.arange(npts) / npts)
    kinds = rng.integers(0, 7, size=(k, k))
    W = np.zeros((npts, k, k), dtype=np.complex128)
    fns = {}
    for i in range(k):
        for j in range(k):
            W[:, i, j], fns[i, j] = _element(rng, kinds[i, j], z, radius)
    d_w = torch.from_numpy(np.ascontiguousarray(W)).to(DEV)
    scr = sigma_real.ScreenedCoulomb(d_w, z, _lib.PADE_ROBUST)
    torch.cuda.synchronize()
    got = d_w.cpu().numpy()
    ref, zref = so.analytic_coeff(so.PADE_ROBUST, W.copy(), z)
    # Degrees: equal, except where the numerator degree of a rational
    # function is below its denominator's: the (n, n) block then carries
    # trailing numerator coefficients that are zero up to rounding, ~1e-14 of
    # the Taylor coefficients, and the drop test against 1e-14 ||c|| is
    # decided by rounding (in the reference as well).  At most 2% of the
    # elements, only of those kinds; the approximants agree below.
    assert np.all(got[:2].imag == 0)
    same = (got[0].real == ref[0].real) & (got[1].real == ref[1].real)
    assert np.all(np.isin(kinds[~same], (1, 2))) and (~same).mean() <= 0.02, kinds[~same]
    # coefficients to 1e-9 of the element's largest one where degrees agree
    scale = np.abs(ref).max(axis=0)
    err = np.abs(got - ref).max(axis=0)
    bad = same & (err > 1e-9 * scale)
    assert not bad.any(), f"kinds {kinds[bad]}, rel err {(err / scale)[bad]}"
    # hopping found the true degrees of the low-degree real elements
    deg = ref[0].real.astype(int)
    assert np.all(deg[kinds == 6] == 0) and np.all(deg[kinds == 4] == 0)
    assert np.all(deg[kinds == 0] == npts // 2 - 2)
    n = 16
    grid = green.CorrGrid(n, _sphere(2))
    gm = rng.permutation(k)[: grid.num_g].astype(np.int32)
    d_gm = torch.from_numpy(gm).to(DEV)
    out = torch.empty((n ** 3, n ** 3), dtype=torch.complex128, device=DEV)
    for freq in (0.2 * radius + 0.4j * radius, 0.7j * radius, -0.3 * radius):
        scr.eval_into(freq, grid, grid, d_gm, d_gm, out)
        want = so.analytic_eval(so.PADE_ROBUST, ref, zref, freq, gm, gm, n, n)
        have = out.cpu().numpy()
        assert np.abs(have - want).max() <= 1e-9 * max(1.0, np.abs(want).max())
        # polynomials (degree <= 4 <= N/2 - 2), constants and zeros, and real
        # rational functions where the circle's aliasing 3^-N is negligible,
        # continue to the function itself: out[c, r] is element (gm[c], gm[r])
        for ci, c in enumerate(gm):
            for ri, r in enumerate(gm):
                if (kinds[c, r] in (3, 4, 6) and npts >= 16) or (kinds[c, r] == 1 and npts >= 32):
                    assert abs(have[ci, ri] - fns[c, r](freq)) <= 1e-10 * max(
                        1.0, abs(fns[c, r](freq))), (kinds[c, r], freq)


def test_robust_coeff_rejects_bad_meshes():
    d_w = torch.zeros((12, 2, 2), dtype=torch.complex128, device=DEV)
    with pytest.raises(RuntimeError, match="at least 10"):
        sigma_real.ScreenedCoulomb(d_w[:8], np.exp(2j * np.pi * np.arange(8) / 8),
                                   _lib.PADE_ROBUST)
    z = np.exp(2j * np.pi * np.arange(12) / 12)
    z[4] *= 1.001
    with pytest.raises(RuntimeError, match="circle"):
        sigma_real.ScreenedCoulomb(d_w, z, _lib.PADE_ROBUST)
