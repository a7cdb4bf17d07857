"""CPU pin of the adjoint oracle: its V_loc gradient equals central finite
differences of the objective, each objective a plain forward solve."""
import numpy as np
import pytest

from oracle import adjoint_oracle as ao
from sternheimergw_amd import cell


@pytest.fixture(scope="module")
def small_c5():
    """The adjoint workload's construction on the 24^3 box: 2 bands, 1 q,
    2 plane-wave perturbations, 3 shifts."""
    return cell.workload_c5(nq=1, nbnd=2, nshift=3, ng=2, box=(24, 24, 24))


def _weights(w):
    _, wt = cell.gauss_legendre(0.0, 4.0, w.nshift)
    p = 1.0 / (1.0 + w.pert_q2)
    return p[:, None] * wt[None, :]


def test_gradient_matches_central_differences(small_c5):
    w = small_c5
    systems = list(range(w.nsys))
    weight = _weights(w)
    c = -w.rhs[systems]
    L, grad = ao.l_and_grad(w, systems, c, weight)
    assert grad.shape == tuple(reversed(w.box))
    rng = np.random.Generator(np.random.PCG64(9))
    h = 1e-5
    for _ in range(2):
        d = rng.standard_normal(w.vloc.shape)
        lp = ao.objective_at(w, w.vloc + h * d, systems, c, weight)
        lm = ao.objective_at(w, w.vloc - h * d, systems, c, weight)
        fd = (lp - lm) / (2 * h)
        an = float(np.sum(grad * d))
        assert abs(fd - an) <= 1e-6 * max(abs(fd), 1e-12), (fd, an)
    # and the objective is the plain forward value
    assert abs(L - ao.objective_at(w, w.vloc, systems, c, weight)) <= 1e-12 * abs(L)


def test_workload_perturbations(small_c5):
    """System layout s = (iq * nbnd + v) * npert + ip; G = 0 reproduces the
    single-perturbation right-hand sides; b is orthogonal to the valence."""
    w = small_c5
    assert w.npert == 2 and w.nsys == 4
    w0 = cell.workload_c5(nq=1, nbnd=2, nshift=3, ng=1, box=(24, 24, 24))
    assert np.array_equal(w.rhs[0::2], w0.rhs)
    for s in range(w.nsys):
        g = int(w.group[s])
        n = int(w.npw[g])
        assert np.abs(w.evq[g].conj() @ w.rhs[s, :n]).max() < 1e-12
