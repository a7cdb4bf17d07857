"""CPU oracle of the self-consistent response around the Sternheimer solves
(test infrastructure only): the steps of solve_linter
(phys/coul/src/solve_linter.f90:279-583) between and after the solves.

* wave_to_box / box_to_wave — QE cft_wave (+1: scatter + invfft; -1: fwfft
  + gather), as used by dvqpsi_us.f90:99-132 and solve_linter.f90:393-398.
* dvscf_psi — dV(r) psi_v(r) gathered on the k+q sphere, plus the saved
  dvbare psi term (solve_linter.f90:387-398).
* orthogonalize — QE's orthogonalize for insulators with S = 1:
  dvpsi <- -(1 - |evq><evq|) dvpsi (solve_linter.f90:337, 411).
* dv_of_drho — QE's dv_of_drho with lrpa (Hartree only; lrpa is set at
  main/src/gwq_readin.f90:223), called at solve_linter.f90:556:
  dv(G) = e2 4 pi / |q+G|^2 drho(G) on the density sphere |G|^2 <= gcut2.
* mix_potential_c — phys/coul/src/mix_pot_c.f90:25-190, statement by
  statement (modified Broyden, D. D. Johnson, PRB 38, 12807 (1988)).
* solve_linter — the whole loop (direct or iterative) for one k-point, one
  perturbation, with the oracle's BiCGStab (sgw_oracle.c).

QE's FFTXlib, cft_wave, orthogonalize and dv_of_drho are external to the
reference tree, so these follow their published formulas.  What pins each
function is in tests/test_scf_oracle.py: plane-wave closed forms (index
shifts, 8 pi/|q+G|^2), projector identities, the Anderson least-squares
equivalence of the Broyden step (Eyert, J. Comput. Phys. 124, 271 (1996)),
and the sum-over-states response from full diagonalisation for
solve_linter (direct first iteration and the self-consistent fixed point).
"""
from __future__ import annotations

import math

import numpy as np

from . import oracle as orc

E2 = 2.0  # e^2 in Rydberg units


def fold(i, n):
    """Miller index of box index i: i for i <= n/2, i - n above."""
    i = np.asarray(i)
    return np.where(i > n // 2, i - n, i)


def box_gvectors(n, bg):
    """Cartesian G (n^3, 3) of every box index in [i3][i2][i1] order and the
    Miller index triplet arrays (m1, m2, m3), each (n, n, n)."""
    m = fold(np.arange(n), n)
    m3, m2, m1 = np.meshgrid(m, m, m, indexing="ij")
    bg = np.asarray(bg, dtype=np.float64).reshape(3, 3)
    g = m1[..., None] * bg[0] + m2[..., None] * bg[1] + m3[..., None] * bg[2]
    return g.reshape(-1, 3), (m1, m2, m3)


def fwfft(box):
    """QE fwfft: sum e^{-iGr} / n^3, on a copy."""
    return orc.fft3d(np.array(box, dtype=np.complex128, copy=True), -1)


def invfft(box):
    """QE invfft: sum e^{+iGr}, no factor, on a copy."""
    return orc.fft3d(np.array(box, dtype=np.complex128, copy=True), +1)


def wave_to_box(n, nl, psi):
    """psi (nvec, >= npw) with npw = len(nl) -> boxes (nvec, n, n, n)."""
    psi = np.atleast_2d(psi)
    out = np.zeros((psi.shape[0], n, n, n), dtype=np.complex128)
    for v in range(psi.shape[0]):
        flat = np.zeros(n ** 3, dtype=np.complex128)
        flat[nl] = psi[v, : len(nl)]
        out[v] = invfft(flat.reshape(n, n, n))
    return out


def box_to_wave(n, nl, box):
    """boxes (nvec, n, n, n) -> fwfft gathered at nl, (nvec, len(nl))."""
    box = np.asarray(box).reshape(-1, n, n, n)
    return np.stack([fwfft(b).reshape(-1)[nl] for b in box])


def dvscf_psi(n, psi_r, dvscf, nlq, base=None):
    """out[f, v] = base[v] + gather_{k+q} fwfft(psi_r[v] dvscf[f])."""
    nf, nb = dvscf.shape[0], psi_r.shape[0]
    out = np.zeros((nf, nb, len(nlq)), dtype=np.complex128)
    for f in range(nf):
        out[f] = box_to_wave(n, nlq, psi_r * dvscf[f][None])
        if base is not None:
            out[f] += base[:, : len(nlq)]
    return out


def orthogonalize(dvpsi, evq):
    """dvpsi (nvec, npw), evq (nocc, npw): -(dvpsi - evq^T conj(evq) dvpsi)."""
    ps = evq.conj() @ dvpsi.T                      # <evq_j | dvpsi_v>
    return -(dvpsi - (evq.T @ ps).T)


def dv_of_drho(box, bg, xq, gcut2, zero_mean=False):
    """Hartree response (Ry) of drho boxes (nvec, n, n, n), returned new."""
    box = np.asarray(box).reshape(-1, *np.shape(box)[-3:])
    n = box.shape[-1]
    g, (m1, m2, m3) = box_gvectors(n, bg)
    g2 = np.einsum("ij,ij->i", g, g)
    qg = g + np.asarray(xq)[None, :]
    qg2 = np.einsum("ij,ij->i", qg, qg)
    keep = (g2 <= gcut2) & (qg2 > 1e-10)
    if zero_mean:
        keep &= ~((m1 == 0) & (m2 == 0) & (m3 == 0)).reshape(-1)
    fac = np.zeros(n ** 3)
    fac[keep] = E2 * 4.0 * math.pi / qg2[keep]
    fac = fac.reshape(n, n, n)
    return np.stack([invfft(fwfft(b) * fac) for b in box])


class MixState:
    """The SAVE'd history arrays df, dv of mix_potential_c (None when the
    reference has them deallocated)."""

    def __init__(self):
        self.df = None
        self.dv = None


def mix_potential_c(state, vout, vin, alphamix, tr2, it, n_iter):
    """mix_pot_c.f90:25-190 with 1-based slot arithmetic as written there.
    vout, vin: complex vectors (ndim,), modified copies are returned as
    (vout, vin, dr2, conv)."""
    maxter = 8
    vout = np.array(vout, dtype=np.complex128, copy=True)
    vin = np.array(vin, dtype=np.complex128, copy=True)
    ndim = vout.size
    w0 = 0.01
    w = np.ones(maxter)
    vout = vout - vin                                               # :93-95
    dr2 = (np.linalg.norm(vout) / ndim) ** 2                        # :96-102
    conv = dr2 < tr2
    if it == 1 and state.df is None:                                # :106-109
        state.df = np.zeros((n_iter, ndim), dtype=np.complex128)
        state.dv = np.zeros((n_iter, ndim), dtype=np.complex128)
    if conv:                                                        # :110-114
        state.df = state.dv = None
        return vout, vin, dr2, conv
    iter_used = min(it - 1, n_iter)                                 # :120
    ipos = it - 1 - int((it - 2) / n_iter) * n_iter                 # :125 (Fortran truncation)
    if it > 1:                                                      # :127-138
        state.df[ipos - 1] = vout - state.df[ipos - 1]
        state.dv[ipos - 1] = vin - state.dv[ipos - 1]
        norm = np.linalg.norm(state.df[ipos - 1])
        state.df[ipos - 1] /= norm
        state.dv[ipos - 1] /= norm
    vinsave = vin.copy()                                            # :141
    beta = np.zeros((maxter, maxter), dtype=np.complex128)
    for i in range(iter_used):                                      # :143-148
        for j in range(i + 1, iter_used):
            beta[i, j] = w[i] * w[j] * np.vdot(state.df[j], state.df[i])
        beta[i, i] = w0 ** 2 + w[i] ** 2
    if iter_used > 0:                                               # :152-163 ZHETRF/ZHETRI('U')
        k = iter_used
        a = np.triu(beta[:k, :k]) + np.triu(beta[:k, :k], 1).conj().T
        inv = np.linalg.inv(a)
        beta[:k, :k] = inv
    work = np.array([np.vdot(state.df[i], vout) for i in range(iter_used)])   # :165-167
    vin = vin + alphamix * vout                                     # :169-171
    for i in range(iter_used):                                      # :173-181
        gamma = 0.0
        for j in range(iter_used):
            gamma += beta[j, i] * w[j] * work[j]
        vin = vin - w[i] * gamma * (alphamix * state.df[i] + state.dv[i])
    inext = it - int((it - 1) / n_iter) * n_iter                    # :183
    state.df[inext - 1] = vout
    state.dv[inext - 1] = vinsave
    return vout, vin, dr2, conv


def coulomb_frequencies(freq):
    """solve_linter.f90:237-252: [freq, -freq(2:)] if freq(1) == 0 else [freq, -freq]."""
    freq = np.asarray(freq, dtype=np.complex128)
    zero_freq = abs(freq[0]) < 1e-14
    return np.concatenate([freq, -freq[1:] if zero_freq else -freq]), zero_freq


def average_pm(x, nf, zero_freq):
    """solve_linter.f90:462-480 on x[..., ishift, :] with 2nf(-1) shifts."""
    out = x[..., :nf, :].copy()
    if zero_freq:
        out[..., 1:, :] = 0.5 * out[..., 1:, :] + 0.5 * x[..., nf:, :]
    else:
        out = 0.5 * out + 0.5 * x[..., nf:, :]
    return out


def solve_linter(w, group, dvbare, freq, num_iter, *, bg, gcut2, threshold=1e-10, lmax=4,
                 alpha_mix=0.7, nmix=4, tr2=1e-14, wk=2.0, volume=None, max_iter=10000):
    """solve_linter for one k+q group and the bare potential dvbare (n, n, n)
    on the box (real space, QE's dvbarein).  Returns (drhoscf (nf, n, n, n),
    iterations).  Direct when num_iter == 1 (drhoscf = -dv_of_drho(drho));
    otherwise the self-consistent dvscf (drhoscf = dvscfin)."""
    n = w.box[0]
    sq = w.spheres[group]
    npwq = sq.npw
    nlk = w.sphere_k.box_index(w.box)
    nlq = sq.box_index(w.box)
    evq = w.evq[group]
    nbnd = w.nbnd
    op = orc.PwOp(w, group)
    omega = volume if volume is not None else abs(np.linalg.det(2 * np.pi * np.linalg.inv(
        np.asarray(bg).reshape(3, 3)).T))
    weight = 2.0 * wk / omega                                       # QE incdrhoscf wgt
    nf = len(freq)
    omegas, zero_freq = coulomb_frequencies(freq)
    direct = num_iter == 1
    psi_r = wave_to_box(n, nlk, w.psi_k)
    dvbpsi = dvscf_psi(n, psi_r, dvbare[None], nlq)[0]              # dvqpsi_us, saved
    meandvb = np.sqrt(np.sum(np.abs(dvbare) ** 2)) / n ** 3          # :532
    xq = sq.kvec - w.sphere_k.kvec
    state = MixState()
    dvscfin = np.zeros((nf, n, n, n), dtype=np.complex128)
    dr2 = 0.0
    it = 0
    for it in range(1, num_iter + 1):
        dpsi = np.zeros((nbnd, nf, npwq), dtype=np.complex128)
        if it == 1:
            rhs = orthogonalize(dvbpsi, evq)
            thresh = threshold if direct else 1e-2
            x = np.zeros((nbnd, len(omegas), npwq), dtype=np.complex128)
            for v in range(nbnd):
                xs, ierr, _it = orc.bicgstab(op, rhs[v], -(w.eps_v[v] + omegas), lmax=lmax,
                                             threshold=thresh, max_iter=max_iter)
                if ierr != 0:
                    raise RuntimeError("solver did not converge")
                x[v] = xs
            dpsi = average_pm(x, nf, zero_freq)
        else:
            thresh = min(0.1 * math.sqrt(dr2), 1e-2)
            rhs_f = dvscf_psi(n, psi_r, dvscfin, nlq, base=dvbpsi)
            x = np.zeros((nbnd, len(omegas), npwq), dtype=np.complex128)
            for f in range(nf):
                rhs = orthogonalize(rhs_f[f], evq)
                iom = f + nf - 1 if zero_freq else f + nf
                for v in range(nbnd):
                    for k in ([f] if (zero_freq and f == 0) else [f, iom]):
                        xs, ierr, _it = orc.bicgstab(op, rhs[v], -(w.eps_v[v] + omegas[k:k + 1]),
                                                     lmax=lmax, threshold=thresh,
                                                     max_iter=max_iter)
                        if ierr != 0:
                            raise RuntimeError("solver did not converge")
                        x[v, k] = xs[0]
            dpsi = average_pm(x, nf, zero_freq)
        # incdrhoscf for every frequency (:486-497)
        pr = psi_r
        drho = np.zeros((nf, n, n, n), dtype=np.complex128)
        for f in range(nf):
            dr = wave_to_box(n, nlq, dpsi[:, f, :])
            drho[f] = weight * np.sum(pr.conj() * dr, axis=0)
        dvscfout = dv_of_drho(drho, bg, xq, gcut2, zero_mean=meandvb < 1e-10)
        if direct:
            return -dvscfout, it
        _vo, vin, dr2, conv = mix_potential_c(state, dvscfout.reshape(-1), dvscfin.reshape(-1),
                                              alpha_mix, tr2 * nf, it, nmix)
        dvscfin = vin.reshape(nf, n, n, n)
        if conv:
            break
    else:
        raise RuntimeError("Iterative solver did not converge within given number of iterations")
    return dvscfin, it
