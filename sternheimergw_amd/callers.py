"""The two callers of the solver, batched: Green's function and the
first (direct) Sternheimer iteration of the screened Coulomb interaction.

* :func:`green_function` — phys/green/src/green.f90:105-226.  For every G'
  column the reference solves (H - omega) G(:, G') = -delta_{G G'} with the
  multishift solver at shifts -omega (:196-218) using the Green operator
  (alpha_pv = 0, :282-283).  Here all columns form one batch.
* :func:`coulomb_frequencies` / :func:`solve_linter_direct` — the first
  iteration of phys/coul/src/solve_linter.f90 (:237-252, :325-374,
  :462-480): shift list [freq, -freq(2:)], per band sigma = -(eps_v + omega),
  rescaling by wg/wk and the +/-omega average.  All bands of all k+q points
  form one batch.

The QE bookkeeping around these loops (FFT-grid maps, buffers, MPI) is not
part of the path; inputs and outputs are plain arrays in sphere order.
Solver failures raise, as the reference's errore calls abort
(green.f90:208, solve_linter.f90:370).
"""
from __future__ import annotations

import numpy as np

from . import solvers


class SolverError(RuntimeError):
    pass


def _batched_select(op, rhs, sigma, group, config: solvers.SelectSolverType):
    """select_solver semantics over a batch: every system starts with the
    first solver of the priority list (batched multishift BiCGStab when it is
    1 or 2); systems left with ierr != 0 continue down the list one by one
    (select_solver.f90:121-159)."""
    prio = list(config.priority or [])
    if not prio:
        raise ValueError("priority of the solvers not specified")
    nsys = rhs.shape[0]
    x = np.zeros((nsys, sigma.shape[1], op.ld), dtype=np.complex128)
    ierr = np.ones(nsys, dtype=np.int32)
    start = 0
    if prio[0] in (solvers.bicgstab_multi, solvers.bicgstab_no_multi):
        x, ierr, _ = op.solve_batched(rhs, sigma, group=group, lmax=config.bicg_lmax,
                                      threshold=config.threshold, max_iter=config.max_iter)
        start = 1
    rest = solvers.SelectSolverType(priority=prio[start:], max_iter=config.max_iter,
                                    threshold=config.threshold, bicg_lmax=config.bicg_lmax)
    for s in np.nonzero(ierr != 0)[0]:
        if not rest.priority:
            break
        g = int(group[s]) if group is not None else 0
        n = int(op.npw[g]) if hasattr(op, "npw") else op.ld
        xs, e = solvers.select_solver(rest, op, rhs[s, :n], sigma[s], group=g)
        x[s, :, :n] = xs.T
        ierr[s] = e
    return x, ierr


def green_function(op, group: int, omega, columns, config: solvers.SelectSolverType):
    """G(G, G'; omega) for the G' listed in `columns` (indices into the k-q
    sphere of `group`), from (H - omega) G = -delta (green.f90:196-218).

    `op` must be the Green operator (no valence projector).  Returns
    green[ncol, nomega, npw]."""
    omega = np.asarray(omega, dtype=np.complex128)
    columns = np.asarray(columns, dtype=np.int64)
    n = int(op.npw[group])
    ncol = columns.size
    rhs = np.zeros((ncol, op.ld), dtype=np.complex128)
    rhs[np.arange(ncol), columns] = -1.0                     # :203-204
    sigma = np.tile(-omega, (ncol, 1))                       # :207
    grp = np.full(ncol, group, dtype=np.int32)
    x, ierr = _batched_select(op, rhs, sigma, grp, config)
    if np.any(ierr != 0):
        raise SolverError("the linear solver for G did not converge")
    return x[:, :, :n]


def coulomb_frequencies(freq):
    """solve_linter.f90:237-252: [freq, -freq(2:)] if freq(1) == 0 else
    [freq, -freq]; returns (omega, zero_freq)."""
    freq = np.asarray(freq, dtype=np.complex128)
    zero_freq = abs(freq[0]) < 1e-14
    neg = -freq[1:] if zero_freq else -freq
    return np.concatenate([freq, neg]), zero_freq


def solve_linter_direct(op, dvpsi, group, eps, freq, wg, wk,
                        config: solvers.SelectSolverType):
    """First iteration of solve_linter for every band of every k+q group.

    dvpsi[s]: right-hand side -P_c^+ dV psi_v of system s (band v of group
    group[s]); eps[s] = et(v, k); wg[s], wk[s]: band and k-point weights.
    Returns dpsi[s, ifreq, :] after the +/-omega average (:462-480)."""
    omega, zero_freq = coulomb_frequencies(freq)
    eps = np.asarray(eps, dtype=np.float64)
    sigma = -(eps[:, None] + omega[None, :])                 # :369
    x, ierr = _batched_select(op, np.asarray(dvpsi, dtype=np.complex128), sigma,
                              np.asarray(group, dtype=np.int32), config)
    if np.any(ierr != 0):
        raise SolverError("solver did not converge")
    x *= (np.asarray(wg) / np.asarray(wk))[:, None, None]     # :373
    nf = len(freq)
    out = x[:, :nf, :].copy()
    if zero_freq:                                            # :464-471
        out[:, 1:, :] = 0.5 * out[:, 1:, :] + 0.5 * x[:, nf:, :]
    else:                                                    # :472-479
        out[:, :, :] = 0.5 * out + 0.5 * x[:, nf:, :]
    return out


def capture_linear_problem(op, group: int, rhs, sigma, x, path: str):
    """Failure capture as in green.f90:415-446: the dense operator of k+q
    group `group` is assembled column by column by applying the device
    operator to unit vectors at sigma = 0 (all columns in one batched
    application), and written with the shifts, the right-hand side and the
    failed solution x (npw, nshift) as an iotk LINEAR_PROBLEM file that the
    reference's solver test reads.  Returns the dense matrix."""
    from . import iotk
    n = int(op.npw[group]) if hasattr(op, "npw") else op.n
    ld = op.ld
    unit = np.zeros((n, ld), dtype=np.complex128)
    unit[np.arange(n), np.arange(n)] = 1.0
    cols = op.apply(unit, 0.0, group=np.full(n, group, dtype=np.int32))
    hamil = np.ascontiguousarray(cols[:, :n].T)   # column j = H e_j
    iotk.write_linear_problem(path, hamil, np.asarray(rhs)[:n], sigma,
                              np.asarray(x).reshape(-1, np.size(sigma))[:n])
    return hamil
