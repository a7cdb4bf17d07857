"""CPU pin of the density-response oracle: the box sum of
weight * conj(psi(r)) dpsi(r) equals N * weight * the coefficient overlap
over the Miller indices both spheres hold (Parseval)."""
import numpy as np

from oracle import density_oracle as do


def test_parseval(c1):
    w = c1
    rng = np.random.Generator(np.random.PCG64(4))
    n = int(w.npw[0])
    dpsi = rng.standard_normal((w.nbnd, 3, n)) + 1j * rng.standard_normal((w.nbnd, 3, n))
    d = do.drho(w, 0, dpsi, 0.37)
    s = d.reshape(3, -1).sum(1)
    ref = do.box_sum(w, 0, dpsi, 0.37)
    assert np.allclose(s, ref, rtol=1e-12, atol=0)
    # the k and k+q spheres differ, so the overlap is a proper subset
    assert w.sphere_k.npw != n or not np.array_equal(w.sphere_k.mill, w.spheres[0].mill)
