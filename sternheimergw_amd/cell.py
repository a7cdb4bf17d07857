"""Synthetic, seeded crystal cells for the Sternheimer solve (host-side setup).

This module plays the part of the ground-state calculation that precedes the
Sternheimer solve in SternheimerGW (Quantum ESPRESSO's pw.x, not in the
reference tree): it builds an fcc cell, plane-wave spheres per k+q point, a
seeded local potential V_loc(r), Gaussian beta projectors with a diagonal D,
the occupied states psi_v from an FP64 dense diagonalisation, the alpha_pv
valence shift and the right-hand sides  b = -P_c S_q psi_v.

Units: everything handed to the operator is in Rydberg atomic units, as in
SternheimerGW/QE (the input conversion is main/src/gwq_readin.f90:266-271).
The workload parameters are specified in Hartree and converted (1 Ha = 2 Ry).
Kinetic energy in Ry is |k+G|^2 (bohr^-2).

FFT conventions follow QE: invfft (G->r) f(r) = sum_G f(G) e^{+iG.r},
fwfft (r->G) divides by the number of box points.  The box index of Miller
triple (m1, m2, m3) is (m1 mod n1) + n1*((m2 mod n2) + n2*(m3 mod n3)).
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

HA2RY = 2.0
SEED_BASE = 20261202


def fcc_lattice(alat: float):
    """QE ibrav=2 fcc primitive vectors (rows), alat in bohr."""
    a = alat / 2.0 * np.array([[-1.0, 0.0, 1.0], [0.0, 1.0, 1.0], [-1.0, 1.0, 0.0]])
    b = 2.0 * math.pi * np.linalg.inv(a).T
    omega = abs(np.linalg.det(a))
    return a, b, omega


def _real_solid_harmonics(v: np.ndarray, l: int) -> np.ndarray:
    """|v|^l * Y_lm(v/|v|) for real spherical harmonics Y_lm, m = -l..l."""
    x, y, z = v[:, 0], v[:, 1], v[:, 2]
    r2 = x * x + y * y + z * z
    pi = math.pi
    if l == 0:
        return np.full((1, v.shape[0]), math.sqrt(1.0 / (4 * pi)))
    if l == 1:
        c = math.sqrt(3.0 / (4 * pi))
        return c * np.stack([y, z, x])
    if l == 2:
        return np.stack([
            math.sqrt(15.0 / (4 * pi)) * x * y,
            math.sqrt(15.0 / (4 * pi)) * y * z,
            math.sqrt(5.0 / (16 * pi)) * (3 * z * z - r2),
            math.sqrt(15.0 / (4 * pi)) * x * z,
            math.sqrt(15.0 / (16 * pi)) * (x * x - y * y),
        ])
    if l == 3:
        return np.stack([
            math.sqrt(35.0 / (32 * pi)) * (3 * x * x - y * y) * y,
            math.sqrt(105.0 / (4 * pi)) * x * y * z,
            math.sqrt(21.0 / (32 * pi)) * y * (5 * z * z - r2),
            math.sqrt(7.0 / (16 * pi)) * z * (5 * z * z - 3 * r2),
            math.sqrt(21.0 / (32 * pi)) * x * (5 * z * z - r2),
            math.sqrt(105.0 / (16 * pi)) * z * (x * x - y * y),
            math.sqrt(35.0 / (32 * pi)) * x * (x * x - 3 * y * y),
        ])
    raise ValueError("l <= 3 supported")


@dataclasses.dataclass
class Sphere:
    """Plane waves |k+G|^2/2 <= ecut (Ha) of one k-point, sorted by column.

    Sorting is by (m1, m2) column, then m3, so that every (m1, m2) column
    is a contiguous run (the layout the GPU pencil transforms use)."""
    kvec: np.ndarray          # Cartesian k (bohr^-1)
    mill: np.ndarray          # (npw, 3) int32 Miller indices
    kin: np.ndarray           # (npw,) |k+G|^2 in Ry
    kg: np.ndarray            # (npw, 3) Cartesian k+G

    @property
    def npw(self) -> int:
        return self.mill.shape[0]

    def box_index(self, box) -> np.ndarray:
        n1, n2, n3 = box
        m = self.mill
        return ((m[:, 0] % n1) + n1 * ((m[:, 1] % n2) + n2 * (m[:, 2] % n3))).astype(np.int32)


def make_sphere(b: np.ndarray, kvec: np.ndarray, ecut_ha: float) -> Sphere:
    gmax = math.sqrt(2.0 * ecut_ha) + np.linalg.norm(kvec)
    a_len = 2 * math.pi / np.linalg.norm(b, axis=1)  # not exact; bound below
    # exact bound: |m_i| <= gmax * |a_i| / (2 pi) with a = 2 pi inv(b).T
    a = 2 * math.pi * np.linalg.inv(b).T
    mmax = [int(math.floor(gmax * np.linalg.norm(a[i]) / (2 * math.pi))) + 1 for i in range(3)]
    del a_len
    r = [np.arange(-mm, mm + 1) for mm in mmax]
    m1, m2, m3 = np.meshgrid(*r, indexing="ij")
    mill = np.stack([m1.ravel(), m2.ravel(), m3.ravel()], axis=1)
    kg = kvec[None, :] + mill @ b
    e = 0.5 * np.einsum("ij,ij->i", kg, kg)
    keep = e <= ecut_ha * (1.0 + 1e-12)
    mill, kg = mill[keep], kg[keep]
    order = np.lexsort((mill[:, 2], mill[:, 1], mill[:, 0]))
    mill, kg = mill[order].astype(np.int32), kg[order]
    kin = np.einsum("ij,ij->i", kg, kg)  # Ry
    return Sphere(kvec=np.asarray(kvec, float), mill=mill, kin=kin, kg=kg)


def make_vloc(box, b: np.ndarray, rms_ha: float, rng: np.random.Generator) -> np.ndarray:
    """Seeded real Gaussian random field with spectrum ~ exp(-|G|^2/2).

    Returns V(r) in Ry on the box, array shape (n3, n2, n1) (x fastest),
    zero mean and rms = rms_ha Hartree."""
    n1, n2, n3 = box
    noise = rng.standard_normal((n3, n2, n1))
    vg = np.fft.fftn(noise)
    f = [np.fft.fftfreq(n, d=1.0 / n) for n in (n3, n2, n1)]
    k3, k2, k1 = np.meshgrid(*f, indexing="ij")
    g = (k1[..., None] * b[0] + k2[..., None] * b[1] + k3[..., None] * b[2])
    g2 = np.einsum("...i,...i->...", g, g)
    vg *= np.sqrt(np.exp(-g2 / 2.0))
    vg[0, 0, 0] = 0.0
    v = np.real(np.fft.ifftn(vg))
    v -= v.mean()
    v *= (rms_ha * HA2RY) / np.sqrt(np.mean(v * v))
    return np.ascontiguousarray(v)


def make_beta(sphere: Sphere, tau: np.ndarray, omega: float, lchan: List[int],
              rc: float = 1.0) -> np.ndarray:
    """beta_lm(k+G) = exp(-|k+G|^2 rc^2/2) |k+G|^l Y_lm e^{-i(k+G).tau} / sqrt(Omega).

    Returns (nproj, npw) complex128, ordered atom-major then l then m."""
    kg = sphere.kg
    q2 = np.einsum("ij,ij->i", kg, kg)
    radial = np.exp(-q2 * rc * rc / 2.0) / math.sqrt(omega)
    out = []
    for t in tau:
        phase = np.exp(-1j * (kg @ t))
        for l in lchan:
            ylm = _real_solid_harmonics(kg, l)
            for m in range(ylm.shape[0]):
                out.append(radial * ylm[m] * phase)
    return np.array(out, dtype=np.complex128)


def vloc_fourier(vloc: np.ndarray) -> np.ndarray:
    """fwfft of V(r): V(G) on the box, indexed [m3 % n3, m2 % n2, m1 % n1]."""
    return np.fft.fftn(vloc) / vloc.size


def dense_hamiltonian(sphere: Sphere, vg: np.ndarray, beta: np.ndarray,
                      dcoef: np.ndarray) -> np.ndarray:
    """H_GG' = |k+G|^2 d_GG' + V(G-G') + sum_p beta_p(G) D_p beta_p(G')^* (Ry).

    This is exactly the operator the FFT path applies: the box product is a
    circular convolution of the sphere coefficients with V(G)."""
    n3, n2, n1 = vg.shape
    m = sphere.mill
    d = m[:, None, :] - m[None, :, :]
    h = vg[d[..., 2] % n3, d[..., 1] % n2, d[..., 0] % n1].astype(np.complex128)
    h[np.diag_indices_from(h)] += sphere.kin
    if beta.shape[0]:
        h += (beta.T * dcoef[None, :]) @ beta.conj()
    return h


def _lowest_eigh(h: np.ndarray, nbnd: int):
    import scipy.linalg
    w, v = scipy.linalg.eigh(h, subset_by_index=[0, nbnd - 1], driver="evr")
    return w, v


@dataclasses.dataclass
class Workload:
    """Everything the batched Sternheimer solve needs, host-side (Ry)."""
    name: str
    box: tuple
    alat: float
    vloc: np.ndarray                 # (n3, n2, n1) Ry
    spheres: List[Sphere]            # one per q (k+q sphere)
    ld: int                          # padded vector length
    nproj_beta: int
    proj: np.ndarray                 # (nq, nproj, ld) complex: beta then valence states
    dcoef: np.ndarray                # (nq, nproj) Ry: D then alpha_pv
    alpha_pv: float
    eps_v: np.ndarray                # (nbnd,) occupied energies at k (Ry)
    evq: List[np.ndarray]            # per q: (nbnd, npw) occupied states at k+q
    rhs: np.ndarray                  # (nq*nbnd, ld) complex, system s = iq*nbnd + v
    group: np.ndarray                # (nsys,) int32 q index of each system
    omega: np.ndarray                # (nshift,) complex frequencies (Ry)
    sigma: np.ndarray                # (nsys, nshift) complex shifts -(eps_v + omega)
    nbnd: int
    dense_h: Optional[List[np.ndarray]] = None  # per q dense H (small cells only)
    sphere_k: Optional[Sphere] = None     # ground-state sphere at k = Gamma
    psi_k: Optional[np.ndarray] = None    # (nbnd, npw_k) occupied states at k
    npert: int = 1                   # plane-wave perturbations e^{i(q+G)r} per (q, band)
    pert_q2: Optional[np.ndarray] = None  # (nsys,) |q+G|^2 (bohr^-2) of each system's perturbation

    @property
    def nsys(self):
        return self.rhs.shape[0]

    @property
    def nshift(self):
        return self.omega.shape[0]

    @property
    def npw(self):
        return np.array([s.npw for s in self.spheres], dtype=np.int32)

    @property
    def nproj(self):
        return self.proj.shape[1]

    def mill_padded(self) -> np.ndarray:
        out = np.zeros((len(self.spheres), self.ld, 3), dtype=np.int32)
        for i, s in enumerate(self.spheres):
            out[i, : s.npw] = s.mill
        return out

    def kin_padded(self) -> np.ndarray:
        out = np.zeros((len(self.spheres), self.ld))
        for i, s in enumerate(self.spheres):
            out[i, : s.npw] = s.kin
        return out


def gauss_legendre(x1: float, x2: float, n: int):
    """Gauss-Legendre nodes/weights on [x1, x2], ascending
    (algo/grid/src/gauleg_grid.f90:23-74)."""
    z, w = np.polynomial.legendre.leggauss(n)
    xm, xl = 0.5 * (x2 + x1), 0.5 * (x2 - x1)
    return xm + xl * z, xl * w


def build_workload(name: str, *, box, ecut_ha: float, nbnd: int, lchan: List[int],
                   qpoints: np.ndarray, omega_ha: np.ndarray, ecut_scale=None,
                   alat: float = 10.26, vloc_rms_ha: float = 0.3, seed: int = SEED_BASE,
                   keep_dense: bool = False, normalize_rhs: bool = True,
                   pert_mill: Optional[np.ndarray] = None) -> Workload:
    """pert_mill (npert, 3): Miller indices G of the plane-wave perturbations
    e^{i(q+G)r}; system s = (iq * nbnd + v) * npert + ip.  Default: G = 0."""
    rng = np.random.Generator(np.random.PCG64(seed))
    a, b, omega = fcc_lattice(alat)
    tau = np.array([[0.0, 0.0, 0.0], [alat / 4, alat / 4, alat / 4]])
    vloc = make_vloc(box, b, vloc_rms_ha, rng)
    vg = vloc_fourier(vloc)
    nq = qpoints.shape[0]
    if ecut_scale is None:
        ecut_scale = np.ones(nq)
    # ground state at k = Gamma
    kvec = np.zeros(3)
    sk = make_sphere(b, kvec, ecut_ha)
    nper = sum(2 * l + 1 for l in lchan)
    dproj = np.array([(0.5 if p % 2 == 0 else -0.5) * HA2RY for p in range(nper * len(tau))])
    beta_k = make_beta(sk, tau, omega, lchan)
    hk = dense_hamiltonian(sk, vg, beta_k, dproj)
    eps_k, psi_k = _lowest_eigh(hk, nbnd)
    spheres, evqs, eps_all, dense = [], [], [eps_k], []
    betas = []
    for iq in range(nq):
        sq = make_sphere(b, kvec + qpoints[iq], ecut_ha * ecut_scale[iq])
        bq = make_beta(sq, tau, omega, lchan)
        hq = dense_hamiltonian(sq, vg, bq, dproj)
        e_q, v_q = _lowest_eigh(hq, nbnd)
        spheres.append(sq)
        betas.append(bq)
        evqs.append(np.ascontiguousarray(v_q.T))
        eps_all.append(e_q)
        if keep_dense:
            dense.append(hq)
    # QE LR_Modules setup_alpha_pv (called at algo/setup/src/gwq_setup.f90:134):
    # alpha_pv = 2 (emax_occ - emin), at least 1e-2 Ry
    emin = min(e.min() for e in eps_all)
    emax = max(e.max() for e in eps_all)
    alpha_pv = max(2.0 * (emax - emin), 1e-2)
    npw_max = max(s.npw for s in spheres)
    ld = ((npw_max + 63) // 64) * 64
    nproj = nper * len(tau) + nbnd
    proj = np.zeros((nq, nproj, ld), dtype=np.complex128)
    dcoef = np.zeros((nq, nproj))
    for iq in range(nq):
        npw = spheres[iq].npw
        proj[iq, : betas[iq].shape[0], :npw] = betas[iq]
        proj[iq, betas[iq].shape[0]:, :npw] = evqs[iq]
        dcoef[iq, : betas[iq].shape[0]] = dproj
        dcoef[iq, betas[iq].shape[0]:] = alpha_pv
    # right-hand sides b = -P_c S_q psi_v (dvqpsi_us.f90:99-132 with a G=0 plane
    # wave perturbation, then QE orthogonalize = -P_c^+; solve_linter.f90:331,337)
    omega_ry = np.asarray(omega_ha, dtype=np.complex128) * HA2RY
    pm = np.zeros((1, 3), dtype=np.int64) if pert_mill is None else np.asarray(pert_mill)
    npert = pm.shape[0]
    nsys = nq * nbnd * npert
    rhs = np.zeros((nsys, ld), dtype=np.complex128)
    group = np.zeros(nsys, dtype=np.int32)
    sigma = np.zeros((nsys, omega_ry.size), dtype=np.complex128)
    pert_q2 = np.zeros(nsys)
    key_k = {tuple(m): i for i, m in enumerate(sk.mill)}
    for iq in range(nq):
        sq = spheres[iq]
        ev = evqs[iq]
        for ip in range(npert):
            # e^{i(q+G)r} psi_v: coefficient at k+q+G2 is psi_v(G2 - G)
            idx = np.array([key_k.get(tuple(m - pm[ip]), -1) for m in sq.mill])
            hit = idx >= 0
            qg = qpoints[iq] + pm[ip] @ b
            for v in range(nbnd):
                s = (iq * nbnd + v) * npert + ip
                dv = np.zeros(sq.npw, dtype=np.complex128)
                dv[hit] = psi_k[idx[hit], v]
                dv = dv - ev.T @ (ev.conj() @ dv)
                dv = -dv
                if normalize_rhs:
                    dv /= np.linalg.norm(dv)
                rhs[s, : sq.npw] = dv
                group[s] = iq
                pert_q2[s] = float(qg @ qg)
                # solve_linter.f90:368-369: shifts -(et(v) + omega)
                sigma[s] = -(eps_k[v] + omega_ry)
    return Workload(name=name, box=tuple(box), alat=alat, vloc=vloc, spheres=spheres, ld=ld,
                    nproj_beta=nper * len(tau), proj=proj, dcoef=dcoef, alpha_pv=alpha_pv,
                    eps_v=eps_k, evq=evqs, rhs=rhs, group=group, omega=omega_ry,
                    sigma=sigma, nbnd=nbnd, dense_h=dense if keep_dense else None,
                    sphere_k=sk, psi_k=np.ascontiguousarray(psi_k.T), npert=npert,
                    pert_q2=pert_q2)


def workload_c1(keep_dense=True, box=(24, 24, 24), **kw) -> Workload:
    """Plain FP64, CPU-runnable: fcc a=10.26, 6 Ha, 24^3, s+p beta per atom
    (8 projectors), 4 bands, one q=(0.1,0,0) 2pi/a, 8 shifts i*linspace(0,2 Ha)."""
    alat = 10.26
    q = np.array([[0.1 * 2 * math.pi / alat, 0.0, 0.0]])
    return build_workload("c1", box=box, ecut_ha=6.0, nbnd=4, lchan=[0, 1], qpoints=q,
                          omega_ha=1j * np.linspace(0.0, 2.0, 8), keep_dense=keep_dense,
                          seed=SEED_BASE + 1, **kw)


def shortest_millers(n: int, alat: float = 10.26) -> np.ndarray:
    """Miller indices of the n shortest reciprocal-lattice vectors of the fcc
    cell, sorted by length, then by Miller index."""
    _, b, _ = fcc_lattice(alat)
    bmin = np.min(np.linalg.norm(b, axis=1))
    m = 1
    while True:
        r = np.arange(-m, m + 1)
        mill = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
        g2 = ((mill @ b) ** 2).sum(1)
        order = np.lexsort((mill[:, 2], mill[:, 1], mill[:, 0], np.round(g2, 10)))
        # the n-th shell must lie inside the sphere inscribed in the searched cube
        if len(order) >= n and g2[order[n - 1]] < (m * bmin / math.sqrt(3.0)) ** 2:
            return np.ascontiguousarray(mill[order[:n]])
        m += 1


def workload_c5(nq: int = 4, nbnd: int = 8, nshift: int = 16, ng: int = 500,
                box=(48, 48, 48), **kw) -> Workload:
    """Adjoint workload: the 6 Ha cell of the plain configuration on a 48^3
    box, 8 bands x 4 seeded q x ng plane-wave perturbations e^{i(q+G)r}
    (the ng shortest G) x nshift Gauss-Legendre imaginary shifts on [0, 4 Ha];
    right-hand sides -P_c e^{i(q+G)r} psi_v unnormalised (chi0 columns)."""
    q, _ = c2_qpoints(nq, spread=0.0, seed=SEED_BASE + 5)
    w, _ = gauss_legendre(0.0, 4.0, nshift)
    return build_workload("c5", box=box, ecut_ha=6.0, nbnd=nbnd, lchan=[0, 1], qpoints=q,
                          omega_ha=1j * w, seed=SEED_BASE + 5, normalize_rhs=False,
                          pert_mill=shortest_millers(ng), **kw)


def c2_qpoints(nq: int = 16, spread: float = 0.08, seed: int = SEED_BASE + 2):
    """Seeded q-points uniform in the first-BZ parallelepiped and per-q cutoff
    scales giving a +-spread plane-wave count variation (N_pw ~ Ecut^{3/2})."""
    rng = np.random.Generator(np.random.PCG64(seed))
    _, b, _ = fcc_lattice(10.26)
    frac = rng.uniform(-0.5, 0.5, size=(nq, 3))
    q = frac @ b
    u = rng.uniform(-1.0, 1.0, size=nq)
    scale = (1.0 + spread * u) ** (2.0 / 3.0)
    return q, scale


def workload_c2(nq: int = 16, nbnd: int = 16, nshift: int = 32, ecut_ha: float = 25.0,
                box=(48, 48, 48), **kw) -> Workload:
    """Batched FP64: 25 Ha, 48^3, s+p+d+f beta per atom (32 projectors),
    16 bands x 16 q (N_pw +-8%) x 32 Gauss-Legendre imaginary shifts on [0, 4 Ha]."""
    q, scale = c2_qpoints(nq)
    w, _ = gauss_legendre(0.0, 4.0, nshift)
    return build_workload("c2", box=box, ecut_ha=ecut_ha, nbnd=nbnd, lchan=[0, 1, 2, 3],
                          qpoints=q, ecut_scale=scale, omega_ha=1j * w,
                          seed=SEED_BASE + 2, **kw)
