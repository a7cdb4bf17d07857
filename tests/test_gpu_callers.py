"""GPU tests of the batched mirrors of the solver's two callers
(sternheimergw_amd/callers.py) against dense direct solves of the 24^3 cell:

* green_function (phys/green/src/green.f90:196-218): (H - omega) G(:, G') =
  -delta_{G'} for every listed column and frequency;
* solve_linter_direct (phys/coul/src/solve_linter.f90:237-252, 325-374,
  462-480): per band, the solves at -(eps_v + omega) over [freq, -freq(2:)],
  the wg/wk factor and the +/-omega average.
"""
import numpy as np
import pytest

from sternheimergw_amd import callers, operators, solvers

pytestmark = pytest.mark.gpu


def _config(thr):
    return solvers.SelectSolverType(priority=[1, 3], max_iter=10000, threshold=thr, bicg_lmax=4)


def test_callers_green_function_dense(c1):
    w = c1
    op = operators.PwOperator.from_workload(w, with_valence=False)
    n = int(w.npw[0])
    omega = np.array([0.1 + 0.3j, -0.5 + 0.2j, 0.4j])
    cols = np.array([0, 3, 17, n - 1])
    print("callers.green_function: 4 columns x 3 frequencies", flush=True)
    g = callers.green_function(op, 0, omega, cols, _config(1e-12))
    h = w.dense_h[0]
    for i, c in enumerate(cols):
        e = np.zeros(n, dtype=np.complex128)
        e[c] = -1.0
        for f, om in enumerate(omega):
            ref = np.linalg.solve(h - om * np.eye(n), e)
            assert np.linalg.norm(g[i, f] - ref) <= 1e-9 * np.linalg.norm(ref), (c, om)


def test_callers_solve_linter_direct_dense(c1):
    w = c1
    op = operators.PwOperator.from_workload(w)
    n = int(w.npw[0])
    freq = 1j * np.array([0.0, 0.3, 1.1])
    eps = w.eps_v[: w.nbnd]
    wg = np.full(w.nbnd, 2.0)
    wk = np.full(w.nbnd, 1.0)
    print("callers.solve_linter_direct: 4 bands x 3 frequencies (5 shifts)", flush=True)
    out = callers.solve_linter_direct(op, w.rhs, w.group, eps, freq, wg, wk, _config(1e-12))
    ev = w.evq[0]
    h = w.dense_h[0] + w.alpha_pv * (ev.T @ ev.conj())
    for v in range(w.nbnd):
        b = w.rhs[v, :n]

        def solve(om):
            return np.linalg.solve(h - (eps[v] + om) * np.eye(n), b) * wg[v] / wk[v]

        for f, om in enumerate(freq):
            ref = solve(om) if f == 0 else 0.5 * (solve(om) + solve(-om))
            assert np.linalg.norm(out[v, f, :n] - ref) <= 1e-9 * np.linalg.norm(ref), (v, f)
