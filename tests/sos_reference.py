"""Sum-over-states response of the synthetic cells (checker for the
self-consistent response tests): the exact first-iteration density response
from full diagonalisation of H_{k+q}, and the self-consistent fixed point of
(1 - v_H chi0) dV = v_H chi0 dV_bare solved by GMRES.  FFTs are numpy's,
independent of the oracle's own."""
from __future__ import annotations

import numpy as np
from scipy.sparse.linalg import LinearOperator, gmres

from oracle import scf_oracle as so


def invfft(b):
    return np.fft.ifftn(b) * b.size


def fwfft(b):
    return np.fft.fftn(b) / b.size


class Chi0:
    """chi0(omega) on box potentials of one k+q group: drho = weight sum_v
    conj(psi_v) dpsi_v with dpsi_v the +/-omega average of
    sum_c |c><c|b_v>/(e_c - e_v - omega), b_v = -P_c (dV psi_v)."""

    def __init__(self, w, group, freq, weight):
        self.w, self.n = w, w.box[0]
        n = self.n
        e, u = np.linalg.eigh(w.dense_h[group])
        nb = w.nbnd
        self.c, self.ec = u[:, nb:], e[nb:]
        self.nlk = w.sphere_k.box_index(w.box)
        self.nlq = w.spheres[group].box_index(w.box)
        self.psi_r = []
        for v in range(nb):
            f = np.zeros(n ** 3, dtype=np.complex128)
            f[self.nlk] = w.psi_k[v]
            self.psi_r.append(invfft(f.reshape(n, n, n)))
        self.omegas, self.zero_freq = so.coulomb_frequencies(freq)
        self.nf = len(freq)
        self.weight = weight

    def __call__(self, dv):
        """dv (n, n, n) -> drho (nf, n, n, n)."""
        n, nf, w = self.n, self.nf, self.w
        out = np.zeros((nf, n, n, n), dtype=np.complex128)
        for v, pr in enumerate(self.psi_r):
            b = -fwfft(pr * dv).reshape(-1)[self.nlq]
            cb = self.c.conj().T @ b
            x = np.stack([self.c @ (cb / (self.ec - w.eps_v[v] - om)) for om in self.omegas])
            dpsi = so.average_pm(x[None], nf, self.zero_freq)[0]
            for f in range(nf):
                g = np.zeros(n ** 3, dtype=np.complex128)
                g[self.nlq] = dpsi[f]
                out[f] += self.weight * pr.conj() * invfft(g.reshape(n, n, n))
        return out


def fixed_point(chi0, dvbare, bg, xq, gcut2, tol=1e-13):
    """dvscf (nf, n, n, n) with dvscf = v_H chi0 (dvbare + dvscf), per frequency."""
    n, nf = chi0.n, chi0.nf
    vh = lambda rho: so.dv_of_drho(rho, bg, xq, gcut2)[0]        # noqa: E731
    rhs_all = chi0(dvbare)
    out = np.zeros((nf, n, n, n), dtype=np.complex128)
    for f in range(nf):
        def mv(x, f=f):
            x = x.reshape(n, n, n)
            return (x - vh(chi0(x)[f])).reshape(-1)
        A = LinearOperator((n ** 3, n ** 3), matvec=mv, dtype=np.complex128)
        rhs = vh(rhs_all[f]).reshape(-1)
        sol, info = gmres(A, rhs, rtol=tol, atol=0.0, restart=200, maxiter=2000)
        assert info == 0
        out[f] = sol.reshape(n, n, n)
    return out
