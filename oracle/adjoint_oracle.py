"""CPU oracle of the adjoint of the batched Sternheimer solve (test
infrastructure only).

For L(V) = sum_{s,k} w_sk Re <c_s | x_sk>, (H(V) + sigma_sk) x_sk = b_s, the
derivative with respect to the local potential V(r) at fixed b, c, sigma and
projectors is

    dL/dV(r) = -(1/N) sum_{s,k} w_sk Re[conj(y_sk(r)) x_sk(r)],
    (H + conj(sigma_sk)) y_sk = c_s,   x(r) = invfft(x(G)) (QE sign, no factor).

The reference has no derivatives; this oracle is pinned by central finite
differences of L itself (tests/test_adjoint_oracle.py), where every L is an
ordinary forward solve of the bicgstab restatement.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import oracle as orc


def objective(c, x, weight):
    """sum w_sk Re <c_s | x_sk>; c (nsys, n), x (nsys, nshift, n), weight (nsys, nshift)."""
    return float(np.sum(weight * np.real(np.einsum("sn,skn->sk", c.conj(), x))))


def with_vloc(w, vloc):
    """The workload with V_loc(r) replaced (everything else fixed)."""
    return dataclasses.replace(w, vloc=np.ascontiguousarray(vloc))


def solve(w, systems, rhs, sigma, threshold=1e-13, lmax=4):
    """Forward solves of the given systems with the oracle; returns (len, nshift, npw_max)."""
    n = max(int(w.npw[w.group[s]]) for s in systems)
    out = np.zeros((len(systems), sigma.shape[1], n), dtype=np.complex128)
    for i, s in enumerate(systems):
        g = int(w.group[s])
        op = orc.PwOp(w, g)
        m = int(w.npw[g])
        x, ierr, _ = orc.bicgstab(op, rhs[i, :m], sigma[i], lmax=lmax, threshold=threshold)
        assert ierr == 0
        out[i, :, :m] = x
    return out


def to_real(w, g, vec):
    """invfft of one plane-wave vector of group g on the box: (n3, n2, n1)."""
    n1, n2, n3 = w.box
    s = w.spheres[g]
    box = np.zeros(n1 * n2 * n3, dtype=np.complex128)
    box[s.box_index(w.box)] = vec[: s.npw]
    box = box.reshape(n3, n2, n1)
    return orc.fft3d(box, +1)


def l_and_grad(w, systems, c, weight, threshold=1e-13, lmax=4):
    """L and dL/dV (n3, n2, n1) for a subset of systems (rows of c, weight)."""
    sig = w.sigma[systems]
    x = solve(w, systems, w.rhs[systems], sig, threshold, lmax)
    y = solve(w, systems, c, sig.conj(), threshold, lmax)
    n = x.shape[2]
    L = objective(c[:, :n], x, weight)
    grad = np.zeros(tuple(reversed(w.box)))
    for i, s in enumerate(systems):
        g = int(w.group[s])
        for k in range(sig.shape[1]):
            xr = to_real(w, g, x[i, k])
            yr = to_real(w, g, y[i, k])
            grad += weight[i, k] * np.real(yr.conj() * xr)
    return L, -grad / float(np.prod(w.box))


def objective_at(w, vloc, systems, c, weight, threshold=1e-13, lmax=4):
    ww = with_vloc(w, vloc)
    x = solve(ww, systems, ww.rhs[systems], ww.sigma[systems], threshold, lmax)
    return objective(c[:, : x.shape[2]], x, weight)
