"""GPU parity on the reference's own solver fixture.

Restates algo/linear_solver/test/linear_solver.pf (test_bicgstab :125-259,
test_linear_solver :264-345) against the HIP library, and compares every
result with the CPU oracle (oracle/sgw_oracle.c) on the same inputs.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from sternheimergw_amd import operators, solvers

pytestmark = pytest.mark.gpu

THRESHOLD = 1e-6  # linear_solver.pf:169
# At the reference's loose threshold the two implementations take the same
# iterations but their rounding differs (parallel vs sequential sums); the
# Krylov polynomials of degree up to 2l amplify that difference, so iterates
# agree to 1e-10..2e-8 relative (measured; largest at l >= 11) while each
# solution is only accurate to ~cond * 1e-6.  The 1e-10 bound of the plain
# FP64 form is asserted on converged solves (test_bicgstab_converged_parity).
LOOSE_PARITY = 1e-7
TIGHT_PARITY = 1e-10


def _residual(a, sigma, x, b):
    return np.linalg.norm(a @ x + sigma * x - b)


@pytest.fixture(scope="module")
def dense(lin_prob):
    return operators.DenseOperator(lin_prob["LINEAR_OPERATOR"])


@pytest.mark.parametrize("ll", list(range(1, 16)))
def test_bicgstab_single_shift(lin_prob, dense, ll):
    """linear_solver.pf:199-220: every shift alone, |(A + sigma) x - b| < 1e-6.
    All 70 single-shift systems run as one batch (independent systems)."""
    a, b, sig = lin_prob["LINEAR_OPERATOR"], lin_prob["RIGHT_HAND_SIDE"], lin_prob["LIST_SHIFT"]
    ns = sig.size
    x, ierr, iters = dense.solve_batched(np.tile(b, (ns, 1)), sig[:, None], lmax=ll,
                                         threshold=THRESHOLD)
    assert np.all(ierr == 0)
    assert not np.isnan(x).any()
    op = orc.DenseOp(a)
    for k in range(ns):
        assert _residual(a, sig[k], x[k, 0], b) < THRESHOLD
    # parity with the oracle on a spread of shifts: same iteration count, same x
    for k in (0, ns // 3, ns - 1):
        xo, eo, io = orc.bicgstab(op, b, sig[k:k + 1], lmax=ll, threshold=THRESHOLD)
        assert eo == 0 and io == iters[k]
        assert np.linalg.norm(x[k, 0] - xo[0]) <= LOOSE_PARITY * np.linalg.norm(xo[0])


@pytest.mark.parametrize("ll", list(range(1, 16)))
def test_bicgstab_multishift(lin_prob, dense, ll):
    """linear_solver.pf:229-252: all 70 shifts at once, residual < 10 * 1e-6;
    parity with the oracle (iterations and every shifted solution)."""
    a, b, sig = lin_prob["LINEAR_OPERATOR"], lin_prob["RIGHT_HAND_SIDE"], lin_prob["LIST_SHIFT"]
    xx, ierr, iters = solvers.bicgstab(solvers.BicgstabType(lmax=ll, threshold=THRESHOLD), dense,
                                       b, sig, return_iters=True)
    assert ierr == 0
    assert not np.isnan(xx).any()
    for k in range(sig.size):
        assert _residual(a, sig[k], xx[:, k], b) < 10.0 * THRESHOLD
    xo, eo, io = orc.bicgstab(orc.DenseOp(a), b, sig, lmax=ll, threshold=THRESHOLD)
    assert eo == 0 and io == iters
    rel = np.linalg.norm(xx.T - xo, axis=1) / np.linalg.norm(xo, axis=1)
    assert rel.max() <= LOOSE_PARITY


@pytest.mark.parametrize("ll", [1, 2, 4, 8, 15])
def test_bicgstab_converged_parity(lin_prob, dense, ll):
    """Plain FP64 form: solved to 1e-12 the GPU solution is within 1e-10
    relative of the oracle's FP64 CPU solve for every shift."""
    a, b, sig = lin_prob["LINEAR_OPERATOR"], lin_prob["RIGHT_HAND_SIDE"], lin_prob["LIST_SHIFT"]
    xx, ierr = solvers.bicgstab(solvers.BicgstabType(lmax=ll, threshold=1e-12), dense, b, sig)
    xo, eo, _ = orc.bicgstab(orc.DenseOp(a), b, sig, lmax=ll, threshold=1e-12)
    assert ierr == 0 and eo == 0
    rel = np.linalg.norm(xx.T - xo, axis=1) / np.linalg.norm(xo, axis=1)
    assert rel.max() <= TIGHT_PARITY


@pytest.mark.parametrize("ll", [1, 2, 4])
def test_batched_shift_test(lin_prob, dense, ll):
    """The batched solver's shift test (no counterpart in the reference): with
    the fixture's shifts reversed the seed is the easiest system, and the
    solve must go on until every shifted residual passes.  Same iterations
    and iterates as the oracle with the same test; two systems (b, i b).
    The test reads the residuals off Frommer's collinearity factors; for
    l >= 9 rounding in the recurrences breaks that relation (the oracle
    itself ends at a true shifted residual of 1.6e-7 for l = 9 on this
    problem), so it is exercised at the degrees where it holds."""
    a, b = lin_prob["LINEAR_OPERATOR"], lin_prob["RIGHT_HAND_SIDE"]
    sig = lin_prob["LIST_SHIFT"][::-1].copy()
    thr = 1e-8
    x, ierr, iters = dense.solve_batched(np.stack([b, 1j * b]), np.stack([sig, sig]), lmax=ll,
                                         threshold=thr, shift_test=True)
    assert np.all(ierr == 0) and iters[0] == iters[1]
    for s, bs in enumerate((b, 1j * b)):
        for k in range(sig.size):
            assert _residual(a, sig[k], x[s, k], bs) < 1.01 * thr
    xo, eo, io = orc.bicgstab(orc.DenseOp(a), b, sig, lmax=ll, threshold=thr, shift_test=True)
    assert eo == 0 and abs(io - iters[0]) <= 1
    if io == iters[0]:
        rel = np.linalg.norm(x[0] - xo, axis=1) / np.linalg.norm(xo, axis=1)
        assert rel.max() <= LOOSE_PARITY


def test_bicgstab_callback_matches_device_operator(lin_prob, dense):
    """The host-callback entry (the reference's procedure argument) and the
    device operator give the same multishift solution."""
    a, b, sig = lin_prob["LINEAR_OPERATOR"], lin_prob["RIGHT_HAND_SIDE"], lin_prob["LIST_SHIFT"]
    cfg = solvers.BicgstabType(lmax=4, threshold=THRESHOLD)

    def AA(sigma, x):  # linear_problem_apply, linear_solver.pf:100-113
        return a @ x + sigma * x

    x_cb, e_cb, it_cb = solvers.bicgstab(cfg, AA, b, sig[:8], return_iters=True)
    x_dev, e_dev, it_dev = solvers.bicgstab(cfg, dense, b, sig[:8], return_iters=True)
    assert e_cb == 0 and e_dev == 0 and it_cb == it_dev
    assert np.abs(x_cb - x_dev).max() <= 1e-12 * np.abs(x_dev).max()


def test_bicgstab_tight_vs_dense_solve(lin_prob, dense):
    """Converged far below tolerance the multishift solution equals the
    direct solve of every shifted system (closed form)."""
    a, b, sig = lin_prob["LINEAR_OPERATOR"], lin_prob["RIGHT_HAND_SIDE"], lin_prob["LIST_SHIFT"]
    xx, ierr = solvers.bicgstab(solvers.BicgstabType(lmax=4, threshold=1e-13), dense, b, sig)
    assert ierr == 0
    n = b.size
    for k in range(sig.size):
        ex = np.linalg.solve(a + sig[k] * np.eye(n), b)
        assert np.linalg.norm(xx[:, k] - ex) <= 1e-10 * np.linalg.norm(ex)


def test_linear_solver_fixture(lin_prob):
    """linear_solver.pf:264-345: default config, relative residual < 1e-4 for
    every shift; parity with the oracle's subspace solver."""
    a, b, sig = lin_prob["LINEAR_OPERATOR"], lin_prob["RIGHT_HAND_SIDE"], lin_prob["LIST_SHIFT"]

    def AA(sigma, x):
        return a @ x + sigma * x

    cfg = solvers.LinearSolverConfig()
    xx, ierr = solvers.linear_solver(cfg, AA, b, sig)
    assert ierr == 0
    assert abs(cfg.abs_threshold - 1e-4 * np.linalg.norm(b)) < 1e-18
    for k in range(sig.size):
        assert _residual(a, sig[k], xx[:, k], b) / np.linalg.norm(b) < cfg.threshold
    xo, eo, nb = orc.linear_solver(orc.DenseOp(a), b, sig)
    assert eo == 0
    rel = np.linalg.norm(xx.T - xo, axis=1) / np.linalg.norm(xo, axis=1)
    assert rel.max() <= 1e-8


@pytest.mark.parametrize("priority,max_iter,thr", [([1, 3], 2, 1e-12), ([7, 1], 10000, 1e-6),
                                                   ([2], 10000, 1e-6), ([3, 1], 3, 1e-10)])
def test_select_solver_fallback(lin_prob, dense, priority, max_iter, thr):
    """select_solver.f90:121-159: solvers tried in priority order until ierr == 0;
    same ierr and solution as the oracle's dispatch."""
    a, b, sig = lin_prob["LINEAR_OPERATOR"], lin_prob["RIGHT_HAND_SIDE"], lin_prob["LIST_SHIFT"][:6]
    cfg = solvers.SelectSolverType(priority=priority, max_iter=max_iter, threshold=thr)
    xx, ierr = solvers.select_solver(cfg, dense, b, sig)
    xo, eo = orc.select_solver(orc.DenseOp(a), b, sig, priority, threshold=thr, max_iter=max_iter)
    assert ierr == eo
    if ierr == 0:
        rel = np.linalg.norm(xx.T - xo, axis=1) / np.linalg.norm(xo, axis=1)
        assert rel.max() <= 1e-8


def test_nan_operator_gives_ierr_2(lin_prob):
    """An operator producing NaN ends with ierr 2 (bicgstab.f90:264-267)."""
    b, sig = lin_prob["RIGHT_HAND_SIDE"], lin_prob["LIST_SHIFT"][:3]

    def AA(sigma, x):
        return np.full_like(x, np.nan)

    xx, ierr = solvers.bicgstab(solvers.BicgstabType(lmax=2, threshold=1e-6, max_iter=5), AA, b, sig)
    assert ierr == 2


def test_max_iter_gives_ierr_1(lin_prob, dense):
    b, sig = lin_prob["RIGHT_HAND_SIDE"], lin_prob["LIST_SHIFT"][:3]
    xx, ierr, it = solvers.bicgstab(solvers.BicgstabType(lmax=1, threshold=1e-14, max_iter=2),
                                    dense, b, sig, return_iters=True)
    assert ierr == 1 and it == 3
