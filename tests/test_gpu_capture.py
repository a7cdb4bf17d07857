"""Failure capture and replay (green.f90:415-446): the dense operator is
assembled on the device and written as an iotk LINEAR_PROBLEM file; reading
it back and solving with the dense operator reproduces the plane-wave solve."""
import numpy as np
import pytest

from sternheimergw_amd import callers, iotk, operators

pytestmark = pytest.mark.gpu


def test_capture_and_replay(c1, tmp_path):
    w = c1
    g, s = 0, 1
    n = int(w.npw[g])
    op = operators.PwOperator.from_workload(w)
    x, ierr, _ = op.solve_batched(w.rhs[s:s + 1], w.sigma[s:s + 1], group=w.group[s:s + 1],
                                  lmax=4, threshold=1e-12)
    assert ierr[0] == 0
    path = str(tmp_path / "green_solver0.xml")
    hamil = callers.capture_linear_problem(op, g, w.rhs[s], w.sigma[s], x[0, :, :n].T, path)
    # the assembled matrix is the Coulomb operator of the cell: the dense
    # Hamiltonian plus alpha_pv P_v (linear_op.f90:131-138)
    ev = w.evq[g]
    ref = w.dense_h[g] + w.alpha_pv * (ev.T @ ev.conj())
    assert np.abs(hamil - ref).max() <= 1e-12 * np.abs(ref).max()
    lp = iotk.read_linear_problem(path)
    assert lp["DIMENSION"] == n and lp["NUMBER_SHIFT"] == w.nshift
    assert np.array_equal(lp["LINEAR_OPERATOR"], hamil)
    dense = operators.DenseOperator(lp["LINEAR_OPERATOR"])
    xr, er, _ = dense.solve_batched(lp["RIGHT_HAND_SIDE"][None, :], lp["LIST_SHIFT"][None, :],
                                    lmax=4, threshold=1e-12)
    assert er[0] == 0
    rel = np.linalg.norm(xr[0, :, :n] - x[0, :, :n], axis=1) / np.linalg.norm(x[0, :, :n], axis=1)
    assert rel.max() <= 1e-10
