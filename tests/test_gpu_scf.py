"""GPU self-consistent response (the scf.hip entries and the solve_linter
driver of sternheimergw_amd/scf.py) against the CPU oracle
(oracle/scf_oracle.py) and the sum-over-states response of a full
diagonalisation, on the 24^3 cell."""
import ctypes
import math
import os
import sys

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from oracle import scf_oracle as so
from sternheimergw_amd import _lib, cell, operators, scf

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import sos_reference as sr  # noqa: E402

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")
ALAT = 10.26


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def _rand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def unit_box(n, m):
    flat = np.zeros(n ** 3, dtype=np.complex128)
    flat[(m[0] % n) + n * ((m[1] % n) + n * (m[2] % n))] = 1.0
    return flat.reshape(n, n, n)


@pytest.mark.parametrize("n", [24, 32, 36, 48])
def test_box_fft_matches_oracle(n):
    print(f"box fft n={n}", flush=True)
    rng = np.random.default_rng(n)
    x = _rand(rng, 3, n, n, n)
    for sign in (1, -1):
        d = _t(x)
        scf.box_fft(d, sign)
        got = d.cpu().numpy()
        ref = np.stack([orc.fft3d(x[v].copy(), sign) for v in range(3)])
        assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max(), sign


def test_wave_box_transfers(c1):
    w = c1
    n, ld = w.box[0], w.ld
    sq = w.spheres[0]
    rng = np.random.default_rng(1)
    psi = np.zeros((5, ld), dtype=np.complex128)
    psi[:, : sq.npw] = _rand(rng, 5, sq.npw)
    nl = sq.box_index(w.box)
    d_nl = _t(nl)
    box = scf.wave_to_box(n, d_nl, _t(psi), ld)
    ref = so.wave_to_box(n, nl, psi)
    assert np.abs(box.cpu().numpy() - ref).max() <= 1e-13 * np.abs(ref).max()
    out = torch.full((5, ld), 7.0 + 0j, dtype=torch.complex128, device=DEV)
    scf.box_to_wave(box, d_nl, out, ld)
    got = out.cpu().numpy()
    assert np.abs(got[:, : sq.npw] - psi[:, : sq.npw]).max() < 1e-13
    assert np.all(got[:, sq.npw:] == 0)


@pytest.mark.parametrize("with_base", [False, True])
def test_dvscf_psi_matches_oracle(c1, with_base):
    w = c1
    n, ld = w.box[0], w.ld
    rng = np.random.default_rng(2)
    sq = w.spheres[0]
    nlk, nlq = w.sphere_k.box_index(w.box), sq.box_index(w.box)
    psi_r = so.wave_to_box(n, nlk, w.psi_k)
    dv = _rand(rng, 3, n, n, n)
    base = np.zeros((w.nbnd, ld), dtype=np.complex128)
    base[:, : sq.npw] = _rand(rng, w.nbnd, sq.npw)
    got = scf.dvscf_psi(_t(psi_r), _t(dv), _t(nlq), _t(base) if with_base else None, ld)
    ref = so.dvscf_psi(n, psi_r, dv, nlq, base if with_base else None)
    g = got.cpu().numpy()
    assert np.abs(g[:, :, : sq.npw] - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.all(g[:, :, sq.npw:] == 0)


def test_orthogonalize_matches_oracle(c1):
    w = c1
    ld = w.ld
    sq = w.spheres[0]
    rng = np.random.default_rng(3)
    x = np.zeros((6, ld), dtype=np.complex128)
    x[:, : sq.npw] = _rand(rng, 6, sq.npw)
    evq = np.zeros((w.nbnd, ld), dtype=np.complex128)
    evq[:, : sq.npw] = w.evq[0]
    d = _t(x)
    scf.orthogonalize(d, _t(evq), sq.npw, ld)
    got = d.cpu().numpy()
    ref = so.orthogonalize(x[:, : sq.npw], w.evq[0])
    assert np.abs(got[:, : sq.npw] - ref).max() <= 1e-13 * np.abs(ref).max()
    assert np.all(got[:, sq.npw:] == 0)


@pytest.mark.parametrize("zero_mean", [False, True])
def test_dv_of_drho_matches_oracle(zero_mean):
    _a, b, _o = cell.fcc_lattice(ALAT)
    n = 24
    rng = np.random.default_rng(4)
    rho = _rand(rng, 2, n, n, n)
    xq = np.array([0.1 * 2 * math.pi / ALAT, 0.02, 0.0])
    d = _t(rho)
    scf.dv_of_drho(d, b, xq, 48.0, zero_mean)
    ref = so.dv_of_drho(rho, b, xq, 48.0, zero_mean)
    assert np.abs(d.cpu().numpy() - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("zero_freq", [True, False])
def test_average_pm_layouts(zero_freq):
    rng = np.random.default_rng(5)
    nb, nf, ld = 3, 4, 64
    ns = 2 * nf - 1 if zero_freq else 2 * nf
    x0 = _rand(rng, nb, ns, ld)
    got = scf.average_pm(_t(x0), nb, nf, zero_freq, 0, ld, scale=2.0).cpu().numpy()
    ref = 2.0 * so.average_pm(x0, nf, zero_freq)
    assert np.abs(got - ref).max() < 1e-14
    x1 = np.zeros((nb, nf, 2, ld), dtype=np.complex128)
    for f in range(nf):
        x1[:, f, 0] = x0[:, f]
        if not (zero_freq and f == 0):
            x1[:, f, 1] = x0[:, f + nf - 1 if zero_freq else f + nf]
    got1 = scf.average_pm(_t(x1), nb, nf, zero_freq, 1, ld, scale=2.0).cpu().numpy()
    assert np.abs(got1 - ref).max() < 1e-14


def test_broyden_matches_oracle():
    rng = np.random.default_rng(6)
    nd = 20000
    # a contractive linear map applied through a diagonal plus low rank
    dg = 0.7 * np.exp(2j * np.pi * rng.random(nd))
    u = _rand(rng, nd, 3) / math.sqrt(nd)
    c = _rand(rng, nd)
    fmap = lambda v: dg * v + 0.5 * u @ (u.conj().T @ v) + c   # noqa: E731
    st = so.MixState()
    mixer = scf.Broyden(nd, 3)
    vin = _rand(rng, nd)
    for it in range(1, 12):
        vout = fmap(vin)
        d_out, d_in = _t(vout), _t(vin)
        dr2, conv = mixer.mix(d_out, d_in, 0.5, 1e-30, it)
        r_out, r_in, r_dr2, r_conv = so.mix_potential_c(st, vout, vin, 0.5, 1e-30, it, 3)
        assert conv == r_conv
        assert abs(dr2 - r_dr2) <= 1e-12 * r_dr2
        assert np.abs(d_out.cpu().numpy() - r_out).max() <= 1e-13 * np.abs(r_out).max()
        assert np.abs(d_in.cpu().numpy() - r_in).max() <= 1e-11 * np.abs(r_in).max(), it
        vin = r_in
    # the history is required after the first iteration
    fresh = scf.Broyden(nd, 3)
    with pytest.raises(_lib.SgwError):
        fresh.mix(_t(vout), _t(vin), 0.5, 1e-30, 2)


@pytest.fixture(scope="module")
def linter(c1):
    w = c1
    op = operators.PwOperator.from_workload(w)
    setup = scf.LinterSetup(op, w, 0)
    _a, b, omega = cell.fcc_lattice(ALAT)
    return w, op, setup, b, omega


def test_solve_linter_direct_equals_sum_over_states(linter):
    w, _op, setup, b, omega = linter
    n = w.box[0]
    freq = np.array([0.0, 0.5j, 1.5j])
    dvb = so.invfft(unit_box(n, [1, 0, 0]))
    gcut2 = 2 * 4 * 6.0
    print("solve_linter direct on the 24^3 cell", flush=True)
    d, it = scf.solve_linter(setup, dvb, freq, 1, bg=b, gcut2=gcut2, threshold=1e-11)
    chi = sr.Chi0(w, 0, freq, 2 * 2.0 / omega)
    ref = -so.dv_of_drho(chi(dvb), b, w.spheres[0].kvec, gcut2)
    got = d.cpu().numpy()
    assert it == 1
    assert np.abs(got - ref).max() <= 1e-8 * np.abs(ref).max()


@pytest.mark.parametrize("pair_pm", [True, False])
def test_solve_linter_iterative_equals_fixed_point(linter, pair_pm):
    w, _op, setup, b, omega = linter
    n = w.box[0]
    freq = np.array([0.0, 1.0j])
    dvb = so.invfft(unit_box(n, [1, 1, 0]))
    gcut2 = 2 * 4 * 6.0
    print(f"solve_linter iterative pair_pm={pair_pm}", flush=True)
    d, it = scf.solve_linter(setup, dvb, freq, 40, bg=b, gcut2=gcut2, tr2=1e-26,
                             pair_pm=pair_pm)
    chi = sr.Chi0(w, 0, freq, 2 * 2.0 / omega)
    ref = sr.fixed_point(chi, dvb, b, w.spheres[0].kvec, gcut2)
    assert 1 < it < 40
    assert np.abs(d.cpu().numpy() - ref).max() <= 1e-8 * np.abs(ref).max()


def test_coulomb_column_direct(linter):
    w, _op, setup, b, omega = linter
    n = w.box[0]
    freq = np.array([0.0, 1.0j])
    gcut2 = 2 * 4 * 6.0
    pm = np.array([1, 0, 0])
    g_mill = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [-1, 1, 1], [1, 0, 0][::-1]])
    col, it = scf.coulomb_column(setup, pm, freq, 1, g_mill, bg=b, gcut2=gcut2,
                                 threshold=1e-11)
    dvb = so.invfft(unit_box(n, pm))
    chi = sr.Chi0(w, 0, freq, 2 * 2.0 / omega)
    drhoscf = -so.dv_of_drho(chi(dvb), b, w.spheres[0].kvec, gcut2)
    nl = (g_mill[:, 0] % n) + n * ((g_mill[:, 1] % n) + n * (g_mill[:, 2] % n))
    ref = np.stack([sr.fwfft(drhoscf[f]).reshape(-1)[nl] for f in range(2)])
    ref[:, 1] += 1.0
    assert np.abs(col - ref).max() <= 1e-8 * np.abs(ref).max()
    # eps_GG' is 1 - v chi0 on the diagonal: real part above 1 at omega = 0
    assert col[0, 1].real > 1.0


def test_coulomb_columns_batched_equal_one_at_a_time(linter):
    w, _op, setup, b, omega = linter
    freq = np.array([0.0, 1.0j])
    gcut2 = 2 * 4 * 6.0
    pms = np.array([[1, 0, 0], [0, 1, 0], [-1, 1, 1]])
    g_mill = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [-1, 1, 1], [0, 0, 1]])
    print("coulomb columns: 3 perturbations in one batch", flush=True)
    eps = scf.coulomb_columns_direct(setup, pms, freq, g_mill, bg=b, gcut2=gcut2,
                                     threshold=1e-11)
    for p, pm in enumerate(pms):
        col, _it = scf.coulomb_column(setup, pm, freq, 1, g_mill, bg=b, gcut2=gcut2,
                                      threshold=1e-11)
        assert np.abs(eps[p] - col).max() <= 1e-8 * np.abs(col).max(), p


def test_coulomb_columns_streamed_beyond_the_slots(linter):
    """65 perturbations x 4 bands = 260 systems: more than the 256 slots, so
    the solve streams; the columns equal the ones computed one at a time."""
    w, _op, setup, b, omega = linter
    freq = np.array([0.0, 1.0j])
    gcut2 = 2 * 4 * 6.0
    r = np.arange(-2, 3)
    allm = np.array(np.meshgrid(r, r, r, indexing="ij")).reshape(3, -1).T
    allm = allm[np.any(allm != 0, axis=1)]
    pms = allm[np.argsort((allm ** 2).sum(1), kind="stable")][:65]
    assert len(pms) * w.nbnd > scf.SLOTS
    g_mill = np.vstack([[0, 0, 0], pms[:4]])
    print("coulomb columns: 65 perturbations streamed through 256 slots", flush=True)
    eps = scf.coulomb_columns_direct(setup, pms, freq, g_mill, bg=b, gcut2=gcut2,
                                     threshold=1e-11)
    for p in (0, 64):
        col, _it = scf.coulomb_column(setup, pms[p], freq, 1, g_mill, bg=b, gcut2=gcut2,
                                      threshold=1e-11)
        assert np.abs(eps[p] - col).max() <= 1e-8 * np.abs(col).max(), p
