"""The self-consistent linear response around the batched solves, device
resident: solve_linter (phys/coul/src/solve_linter.f90:55-611) for one k-point
(Gamma), one k+q group and one bare perturbation, and the column of the
dielectric matrix that coulomb.f90 (phys/coul/src/coulomb.f90:115-170)
builds from it.

Every arithmetic step is an entry of include/sgw.h (box FFTs, dV psi
products, orthogonalize, the batched multishift solve, the +/-omega average,
the density response, dv_of_drho, mix_potential_c); torch supplies device
memory and the stream.  The loop follows solve_linter:

* iteration 1: dvpsi = -P_c^+ dvbare psi_v, one multishift solve per band over
  the shift list [freq, -freq(2:)] at -(eps_v + omega) (:325-374), threshold
  1e-2 for the iterative solver, the caller's for the direct one (:347-360);
* iteration > 1: dvpsi = -P_c^+ (dvbare + dvscfin(omega_f)) psi_v per
  frequency, solves at +omega_f and -omega_f with threshold
  min(0.1 sqrt(dr2), 1e-2) (:376-456).  The reference runs them as two
  single-shift solves; `pair_pm=True` (default) solves the pair as one
  two-shift system, which shares the matrix-vector products;
* the +/-omega average (:462-480), incdrhoscf (:486-497), dv_of_drho (:534-557);
* direct (num_iter == 1): drhoscf = -dvscfout (:594-598); otherwise Broyden
  mixing (:566-568) until dr2 < tr2 * nf and drhoscf = dvscfin (:600-609).

For the insulators handled here wg/wk = 1 (the :373 rescale) and QE's
incdrhoscf weight is 2 wk / Omega with wk = 2 for a Gamma-only, spin
degenerate sampling.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib
from .batch import BatchSolver


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dev():
    import torch
    return torch.device("cuda", torch.cuda.current_device())


class SolverError(RuntimeError):
    pass


class Broyden:
    """mix_potential_c with its history (phys/coul/src/mix_pot_c.f90)."""

    def __init__(self, ndim: int, n_iter: int):
        self.handle = _lib.lib().sgw_broyden_create(ndim, n_iter)
        if not self.handle:
            raise _lib.SgwError(-1, _lib.lib().sgw_last_error().decode())
        self.ndim, self.n_iter = ndim, n_iter

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            _lib.lib().sgw_broyden_destroy(h)
            self.handle = None

    def mix(self, vout, vin, alphamix: float, tr2: float, it: int):
        """In place on the device tensors; returns (dr2, conv)."""
        dr2 = ctypes.c_double(0.0)
        conv = ctypes.c_int32(0)
        _lib.check(_lib.lib().sgw_mix_potential_c(self.handle, self.ndim, _p(vout), _p(vin),
                                                  alphamix, ctypes.byref(dr2), tr2, it,
                                                  self.n_iter, ctypes.byref(conv), _stream()))
        return dr2.value, bool(conv.value)


def box_fft(box, sign: int):
    n = box.shape[-1]
    _lib.check(_lib.lib().sgw_box_fft(n, box.numel() // n ** 3, sign, _p(box), _stream()))
    return box


def wave_to_box(n: int, d_nl, d_psi, ld: int):
    import torch
    nvec = d_psi.numel() // ld
    out = torch.empty((nvec, n, n, n), dtype=torch.complex128, device=d_psi.device)
    _lib.check(_lib.lib().sgw_wave_to_box(n, nvec, d_nl.numel(), _p(d_nl), _p(d_psi), ld,
                                          _p(out), _stream()))
    return out


def box_to_wave(d_box, d_nl, d_psi, ld: int, accumulate: bool = False):
    """fwfft of d_box (overwritten) gathered at d_nl into d_psi (nvec, ld)."""
    n = d_box.shape[-1]
    nvec = d_box.numel() // n ** 3
    _lib.check(_lib.lib().sgw_box_to_wave(n, nvec, d_nl.numel(), _p(d_nl), _p(d_box), _p(d_psi),
                                          ld, int(accumulate), _stream()))
    return d_psi


def dvscf_psi(d_psi_r, d_dvscf, d_nlq, d_base, ld: int):
    """out[f, v] = base[v] + gather fwfft(psi_r[v] dvscf[f]), (nf, nbnd, ld)."""
    import torch
    n = d_psi_r.shape[-1]
    nbnd = d_psi_r.numel() // n ** 3
    nf = d_dvscf.numel() // n ** 3
    out = torch.empty((nf, nbnd, ld), dtype=torch.complex128, device=d_psi_r.device)
    _lib.check(_lib.lib().sgw_dvscf_psi(n, nbnd, nf, _p(d_psi_r), _p(d_dvscf), d_nlq.numel(),
                                        _p(d_nlq), _p(d_base), _p(out), ld, _stream()))
    return out


def orthogonalize(d_dvpsi, d_evq, npw: int, ld: int):
    """-P_c^+ in place on d_dvpsi (nvec, ld); d_evq (nocc, ld)."""
    nvec = d_dvpsi.numel() // ld
    nocc = d_evq.numel() // ld
    _lib.check(_lib.lib().sgw_orthogonalize(nvec, npw, nocc, _p(d_evq), _p(d_dvpsi), ld,
                                            _stream()))
    return d_dvpsi


def dv_of_drho(d_box, bg, xq, gcut2: float, zero_mean: bool = False):
    n = d_box.shape[-1]
    bgv = np.ascontiguousarray(np.asarray(bg, dtype=np.float64).reshape(9))
    xqv = np.ascontiguousarray(np.asarray(xq, dtype=np.float64).reshape(3))
    _lib.check(_lib.lib().sgw_dv_of_drho(n, d_box.numel() // n ** 3, _lib.ptr(bgv), _lib.ptr(xqv),
                                         gcut2, int(zero_mean), _p(d_box), _stream()))
    return d_box


def average_pm(d_x, nbnd: int, nf: int, zero_freq: bool, layout: int, ld: int,
               scale: float = 1.0):
    import torch
    out = torch.empty((nbnd, nf, ld), dtype=torch.complex128, device=d_x.device)
    _lib.check(_lib.lib().sgw_average_pm(nbnd, nf, int(zero_freq), layout, _p(d_x), ld, scale,
                                         _p(out), _stream()))
    return out


def coulomb_frequencies(freq):
    """solve_linter.f90:237-252: [freq, -freq(2:)] if freq(1) == 0 else [freq, -freq]."""
    freq = np.asarray(freq, dtype=np.complex128)
    zero_freq = abs(freq[0]) < 1e-14
    return np.concatenate([freq, -freq[1:] if zero_freq else -freq]), zero_freq


def _pad(a, ld):
    out = np.zeros((a.shape[0], ld), dtype=np.complex128)
    out[:, : a.shape[1]] = a
    return out


class LinterSetup:
    """Device copies of what solve_linter reads for one k+q group: the
    sphere maps (QE's nl(igk)), psi_v at k on the box and the occupied
    states at k+q."""

    def __init__(self, op, w, group: int):
        import torch
        dev = _dev()
        self.op, self.w, self.group = op, w, group
        self.n = w.box[0]
        if not (w.box[0] == w.box[1] == w.box[2]):
            raise ValueError("cubic FFT boxes only")
        self.ld = op.ld
        sq = w.spheres[group]
        self.npwq = sq.npw
        self.d_nlk = torch.from_numpy(w.sphere_k.box_index(w.box)).to(dev)
        self.d_nlq = torch.from_numpy(sq.box_index(w.box)).to(dev)
        ldk = ((w.sphere_k.npw + 63) // 64) * 64
        self.d_psi_r = wave_to_box(self.n, self.d_nlk,
                                   torch.from_numpy(_pad(w.psi_k, ldk)).to(dev), ldk)
        self.d_evq = torch.from_numpy(_pad(w.evq[group], self.ld)).to(dev)
        self.nbnd = w.nbnd
        self.xq = sq.kvec - w.sphere_k.kvec


SLOTS = 256  # systems solved side by side; larger batches stream through them


def _solve(setup, d_rhs, sigma, threshold, lmax, max_iter, shift_test, stats=None):
    """Batched multishift solve of rhs (nsys, ld) at sigma (nsys, nshift);
    adds the vectors the operator was applied to into stats.  More than
    SLOTS systems stream through SLOTS slots (sgw_batch_solver_run_stream)."""
    import torch
    nsys, nshift = sigma.shape
    dev = d_rhs.device
    sigma = np.ascontiguousarray(sigma)
    d_sigma = torch.from_numpy(sigma).to(dev)
    d_x = torch.zeros((nsys, nshift, setup.ld), dtype=torch.complex128, device=dev)
    if nsys > SLOTS:
        bs = BatchSolver(setup.op, SLOTS, nshift, setup.ld, lmax=lmax, max_iter=max_iter,
                         threshold=threshold, shift_test=shift_test)
        ierr, iters = bs.solve_stream(np.full(nsys, setup.group, dtype=np.int32), d_rhs, d_sigma,
                                      sigma, d_x)
    else:
        bs = BatchSolver(setup.op, nsys, nshift, setup.ld, lmax=lmax, max_iter=max_iter,
                         threshold=threshold, shift_test=shift_test)
        d_group = torch.full((nsys,), setup.group, dtype=torch.int32, device=dev)
        ierr, iters = bs.solve(d_group, d_rhs, d_sigma, sigma, d_x)
    if np.any(ierr != 0):
        raise SolverError("solver did not converge")  # errore, solve_linter.f90:370
    if stats is not None:
        stats["vector_happlies"] = stats.get("vector_happlies", 0) + int(2 * lmax * iters.sum())
    return d_x


def solve_linter(setup: LinterSetup, dvbare, freq, num_iter: int, *, bg, gcut2: float,
                 threshold: float = 1e-10, lmax: int = 4, alpha_mix=0.7, nmix: int = 4,
                 tr2: float = 1e-14, wk: float = 2.0, volume: float = None,
                 max_iter: int = 10000, pair_pm: bool = True, shift_test: bool = True,
                 stats: dict = None):
    """dvbare: (n, n, n) complex, the bare potential on the box (a host array
    or a device tensor).  Returns (drhoscf (nf, n, n, n) device tensor,
    iterations).  stats, when given, accumulates "vector_happlies"."""
    import torch
    w, n, ld, nbnd = setup.w, setup.n, setup.ld, setup.nbnd
    dev = setup.d_psi_r.device
    bgm = np.asarray(bg, dtype=np.float64).reshape(3, 3)
    omega_cell = volume if volume is not None else abs(np.linalg.det(2 * np.pi * np.linalg.inv(bgm).T))
    weight = 2.0 * wk / omega_cell
    freq = np.asarray(freq, dtype=np.complex128)
    nf = freq.size
    omegas, zero_freq = coulomb_frequencies(freq)
    direct = num_iter == 1
    alphas = np.broadcast_to(np.asarray(alpha_mix, dtype=np.float64), (max(num_iter, 1),))
    if isinstance(dvbare, torch.Tensor):
        d_dvbare = dvbare.reshape(n, n, n).contiguous()
        h_dvbare = d_dvbare.cpu().numpy()
    else:
        h_dvbare = np.ascontiguousarray(dvbare, dtype=np.complex128).reshape(n, n, n)
        d_dvbare = torch.from_numpy(h_dvbare).to(dev)
    meandvb = math.sqrt(float(np.sum(np.abs(h_dvbare) ** 2))) / n ** 3   # :532
    # dvbare psi_v on the k+q sphere, saved for every iteration (:330-332)
    d_dvbpsi = dvscf_psi(setup.d_psi_r, d_dvbare, setup.d_nlq, None, ld)[0].contiguous()
    eps = np.asarray(w.eps_v, dtype=np.float64)
    dvscfin = torch.zeros((nf, n, n, n), dtype=torch.complex128, device=dev)
    mixer = None if direct else Broyden(nf * n ** 3, nmix)
    dr2 = 0.0
    for it in range(1, num_iter + 1):
        if it == 1:
            rhs = d_dvbpsi.clone()
            orthogonalize(rhs, setup.d_evq, setup.npwq, ld)
            thresh = threshold if direct else 1e-2                      # :347-360
            sigma = -(eps[:, None] + omegas[None, :])                   # :369
            x = _solve(setup, rhs, sigma, thresh, lmax, max_iter, shift_test, stats)
            dpsi = average_pm(x, nbnd, nf, zero_freq, 0, ld)
        else:
            thresh = min(0.1 * math.sqrt(dr2), 1e-2)                     # :429-430
            rhs_f = dvscf_psi(setup.d_psi_r, dvscfin, setup.d_nlq, d_dvbpsi, ld)  # (nf, nbnd, ld)
            orthogonalize(rhs_f, setup.d_evq, setup.npwq, ld)
            if pair_pm:
                # systems (v, f), shifts (+omega_f, -omega_f); at omega = 0 both are +0
                rhs = rhs_f.permute(1, 0, 2).contiguous()               # (nbnd, nf, ld)
                om_m = np.array([omegas[0] if (zero_freq and f == 0) else
                                 omegas[f + nf - 1] if zero_freq else omegas[f + nf]
                                 for f in range(nf)])
                sig = np.empty((nbnd, nf, 2), dtype=np.complex128)
                sig[:, :, 0] = -(eps[:, None] + omegas[None, :nf])
                sig[:, :, 1] = -(eps[:, None] + om_m[None, :])
                x = _solve(setup, rhs.reshape(nbnd * nf, ld), sig.reshape(nbnd * nf, 2),
                           thresh, lmax, max_iter, shift_test, stats)
                dpsi = average_pm(x, nbnd, nf, zero_freq, 1, ld)
            else:
                # the reference's single-shift solves, systems (v, k) over the shift list
                ns = omegas.size
                kf = [k if k < nf else (k - nf + 1 if zero_freq else k - nf) for k in range(ns)]
                rhs = torch.stack([rhs_f[kf[k]] for k in range(ns)], dim=1).contiguous()  # (nbnd, ns, ld)
                sig = -(eps[:, None] + omegas[None, :])
                x = _solve(setup, rhs.reshape(nbnd * ns, ld), sig.reshape(nbnd * ns, 1),
                           thresh, lmax, max_iter, shift_test, stats)
                dpsi = average_pm(x, nbnd, nf, zero_freq, 0, ld)
        # incdrhoscf for every frequency; the direct result is -dv_of_drho(drho),
        # so the sign goes into the weight there
        drho = torch.empty((nf, n, n, n), dtype=torch.complex128, device=dev)
        _lib.check(_lib.lib().sgw_drho_accumulate(
            setup.op.handle, setup.group, nbnd, nf, _p(dpsi), ld, _p(setup.d_psi_r),
            -weight if direct else weight, _p(drho), 0, _stream()))
        dv_of_drho(drho, bg, setup.xq, gcut2, zero_mean=meandvb < 1e-10)
        if direct:
            return drho, it
        dr2, conv = mixer.mix(drho, dvscfin, float(alphas[it - 1]), tr2 * nf, it)
        if conv:
            return dvscfin, it
    raise SolverError("Iterative solver did not converge within given number of iterations")


def coulomb_column(setup: LinterSetup, pert_mill, freq, num_iter: int, g_mill, **kw):
    """One column of coulomb.f90:115-170: the response to dvbare = e^{i(q+G)r}
    for the perturbation's Miller index G, read at the Miller indices g_mill
    (num_g_corr, 3).  Returns scrcoul[iw, igp] (host): fwfft(drhoscf)(G'),
    plus dvbare(G') on the diagonal for the direct solver (eps_GG'); for the
    iterative solver the self-consistent dV (inveps - 1 in the reference's
    printout)."""
    import torch
    n = setup.n
    dev = setup.d_psi_r.device
    pm = np.asarray(pert_mill, dtype=np.int64)
    gm = np.asarray(g_mill, dtype=np.int64).reshape(-1, 3)
    # dvbare(nl(G)) = 1, invfft (coulomb.f90:130-135)
    d_dvbare = torch.zeros((n, n, n), dtype=torch.complex128, device=dev)
    d_dvbare.view(-1)[int((pm[0] % n) + n * ((pm[1] % n) + n * (pm[2] % n)))] = 1.0
    box_fft(d_dvbare, +1)
    drhoscf, it = solve_linter(setup, d_dvbare, freq, num_iter, **kw)
    nl = ((gm[:, 0] % n) + n * ((gm[:, 1] % n) + n * (gm[:, 2] % n))).astype(np.int32)
    d_nl = torch.from_numpy(nl).to(dev)
    nf = drhoscf.shape[0]
    out = torch.empty((nf, nl.size), dtype=torch.complex128, device=dev)
    box_to_wave(drhoscf, d_nl, out, nl.size)                        # :147-153
    res = out.cpu().numpy()
    if num_iter == 1:                                               # :155-159
        diag = np.all(gm == pm[None, :], axis=1)
        res[:, diag] += 1.0
    return res, it


def coulomb_columns_direct(setup: LinterSetup, pert_mills, freq, g_mill, *, bg, gcut2: float,
                           threshold: float = 1e-10, lmax: int = 4, wk: float = 2.0,
                           volume: float = None, max_iter: int = 10000, shift_test: bool = True,
                           stats: dict = None):
    """The direct solver's columns of coulomb.f90:115-170 for every G in
    pert_mills at once: the reference runs one solve_linter per G
    (coulomb.f90:118-138); with the direct solver the G's do not couple, so
    all (G, band) systems form one batched multishift solve and every other
    step runs over all G's in one call.  Returns eps[ipert, iw, igp] (host),
    the same numbers coulomb_column gives one G at a time."""
    import torch
    w, n, ld, nbnd = setup.w, setup.n, setup.ld, setup.nbnd
    dev = setup.d_psi_r.device
    pms = np.asarray(pert_mills, dtype=np.int64).reshape(-1, 3)
    npert = pms.shape[0]
    gm = np.asarray(g_mill, dtype=np.int64).reshape(-1, 3)
    bgm = np.asarray(bg, dtype=np.float64).reshape(3, 3)
    omega_cell = volume if volume is not None else abs(np.linalg.det(2 * np.pi * np.linalg.inv(bgm).T))
    weight = 2.0 * wk / omega_cell
    freq = np.asarray(freq, dtype=np.complex128)
    nf = freq.size
    omegas, zero_freq = coulomb_frequencies(freq)
    # dvbare of every G (coulomb.f90:130-135); |dvbare| = 1, so meandvb > eps10
    d_dvbare = torch.zeros((npert, n ** 3), dtype=torch.complex128, device=dev)
    idx = torch.from_numpy((pms[:, 0] % n) + n * ((pms[:, 1] % n) + n * (pms[:, 2] % n))).to(dev)
    d_dvbare[torch.arange(npert, device=dev), idx] = 1.0
    box_fft(d_dvbare.view(npert, n, n, n), +1)
    # dvbare psi_v for every (G, v), then -P_c^+ (solve_linter.f90:330-337)
    rhs = dvscf_psi(setup.d_psi_r, d_dvbare.view(npert, n, n, n), setup.d_nlq, None, ld)
    rhs = rhs.reshape(npert * nbnd, ld)
    orthogonalize(rhs, setup.d_evq, setup.npwq, ld)
    eps = np.tile(np.asarray(w.eps_v, dtype=np.float64), npert)
    sigma = -(eps[:, None] + omegas[None, :])                        # :369
    x = _solve(setup, rhs, sigma, threshold, lmax, max_iter, shift_test, stats)
    dpsi = average_pm(x, npert * nbnd, nf, zero_freq, 0, ld)         # (npert*nbnd, nf, ld)
    drho = torch.empty((npert, nf, n, n, n), dtype=torch.complex128, device=dev)
    for p in range(npert):                                           # incdrhoscf, sign of :596
        _lib.check(_lib.lib().sgw_drho_accumulate(
            setup.op.handle, setup.group, nbnd, nf, _p(dpsi[p * nbnd:(p + 1) * nbnd]), ld,
            _p(setup.d_psi_r), -weight, _p(drho[p]), 0, _stream()))
    dv_of_drho(drho.view(npert * nf, n, n, n), bg, setup.xq, gcut2, zero_mean=False)
    nl = ((gm[:, 0] % n) + n * ((gm[:, 1] % n) + n * (gm[:, 2] % n))).astype(np.int32)
    d_nl = torch.from_numpy(nl).to(dev)
    out = torch.empty((npert * nf, nl.size), dtype=torch.complex128, device=dev)
    box_to_wave(drho.view(npert * nf, n, n, n), d_nl, out, nl.size)
    res = out.cpu().numpy().reshape(npert, nf, nl.size)
    for p in range(npert):                                           # coulomb.f90:155-159
        res[p][:, np.all(gm == pms[p][None, :], axis=1)] += 1.0
    return res
