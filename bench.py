"""Benchmark: batched multishift Sternheimer solves/s on one MI355X.

Workload (the 48^3 batched configuration): fcc a = 10.26 bohr, plane waves
|k+q+G|^2/2 <= 25 Ha with a seeded per-q cutoff scale (N_pw 1491..1706),
FFT box 48^3, 32 beta projectors + 16 valence states (alpha_pv P_v),
16 bands x 16 seeded q-points x 32 Gauss-Legendre imaginary shifts on
[0, 4 Ha] = 8192 solves of (H - eps_v - omega + alpha_pv P_v) dpsi = b,
b = -P_c psi_v (unit norm), multishift BiCGStab(4) in FP64.  A system stops
when its seed residual (the reference's test) and every shifted residual
estimate |1/(theta phi)| ||r_seed|| are below 1e-8 (--shift-test 0 keeps the
reference's seed-only test, which leaves some shifts above 1e-8).

A step is one batch of all 8192 systems solved from the right-hand sides
resident in HBM to the solutions resident in HBM.  By default the timed
steps' batches go through the solver as one queue (20 steps: 5120 systems),
1024 at a time (slots refilled as systems converge,
sgw_batch_solver_run_stream); the end of the queue, where the slow systems
finish with few others beside them, weighs less the longer the queue;
--stream 0 solves each batch in lock step.  After the timed steps every (system, shift) true
residual ||b - (A + sigma) x|| / ||b|| of every timed batch is recomputed in
FP64 on the device; only solves at <= 1e-8 count.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
Prints one JSON line.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

RES_TARGET = 1e-8
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "happly_traffic.json")
VALU_FILE = os.path.join(ROOT, "profiles", "happly_fp64_valu.json")
FP64_PEAK = 78.6  # TFLOP/s, MI355X FP64 (vector and matrix)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--threshold", type=float, default=1e-8)
    ap.add_argument("--lmax", type=int, default=4)
    ap.add_argument("--nq", type=int, default=16)
    ap.add_argument("--nbnd", type=int, default=16)
    ap.add_argument("--nshift", type=int, default=32)
    ap.add_argument("--shift-test", type=int, default=1,
                    help="1: a system stops when every shift's residual estimate passes the "
                         "threshold; 0: the reference's test on the seed residual alone")
    ap.add_argument("--slots", type=int, default=1024,
                    help="solver slots of the streamed solve (0: one batch, 256); more slots "
                         "put more vectors into every H-apply launch")
    ap.add_argument("--stream", type=int, default=1,
                    help="1: the timed steps' batches pass through the solver as one queue "
                         "(sgw_batch_solver_run_stream), slots refilled as systems finish; "
                         "0: one lock-step batched solve per step")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64",
                    help="mixed: FP32 local-potential FFTs once a system's residual estimate "
                         "is below --relax-threshold (vectors, recurrences and test in FP64)")
    ap.add_argument("--relax-threshold", type=float, default=1e-4)
    ap.add_argument("--inner-threshold", type=float, default=1e-5,
                    help="mixed workload: threshold of the FP32-FFT multishift stage")
    ap.add_argument("--mixed-cell", type=int, choices=(24, 48), default=24,
                    help="mixed workload: the 24^3 cell with chi0 columns at N_G = 500, or the "
                         "48^3 batch of the headline (dpsi errors only)")
    ap.add_argument("--cpu-baseline", type=int, default=1,
                    help="time the CPU oracle on a sample of the batch (0 to skip)")
    ap.add_argument("--kernel-timing", type=int, default=1,
                    help="HIP-event timing level in the timed steps: 0 off, 1 one event pair "
                         "per H-apply, 2 one per kernel launch")
    ap.add_argument("--breakdown", type=int, default=1,
                    help="after the timed steps, one untimed step with per-kernel events")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run one device copy of this size (MiB) for counter calibration")
    ap.add_argument("--workload", choices=("solve", "sigma", "response", "sigma_real", "mixed",
                                           "continuation"),
                    default="solve",
                    help="solve: the batched Sternheimer solve (headline); sigma: the Sigma_c "
                         "frequency convolution GEMM (N_G = 1500, 64 x 64 frequencies); "
                         "response: one dielectric column (direct solve_linter, 48^3); "
                         "sigma_real: the reference-form Sigma_c products (24^3, N_G 1500); "
                         "mixed: FP32 FFTs with FP64 refinement against FP64 (24^3, chi0 "
                         "columns at N_G 500); continuation: the analytic continuation "
                         "coefficients of W (robust Pade, AAA, AAA poles; N_G 500)")
    ap.add_argument("--cont-model", choices=("pade_robust", "aaa", "aaa_pole"),
                    default="pade_robust",
                    help="the model the continuation workload's value is quoted on (all three "
                         "are timed)")
    ap.add_argument("--cont-ng", type=int, default=500,
                    help="N_G of the continuation workload (N_G^2 elements)")
    ap.add_argument("--sigma-precision", choices=("fp64", "fp32s"), default="fp64",
                    help="fp32s: FP32 MFMA products, per-frequency FP32 sums, FP64 frequency sum")
    ap.add_argument("--sigma-batch", type=int, default=8,
                    help="output frequencies whose Green's-function stack is resident at once")
    return ap.parse_args()


def run_sigma(args):
    """Sigma_c(i w) = alpha sum_j w_j G(i w + i w'_j) W(i w'_j) for 64 output
    and 64 convolution frequencies at N_G = 1500 (8 n^3 flops per (i, j)).
    A step is the full convolution; the Green's-function stack is generated
    on the device a batch of output frequencies at a time (outside the timed
    GEMM time, inside the wall time)."""
    import numpy as np
    import torch

    from sternheimergw_amd import _lib, sigma

    L = _lib.lib()
    dev = torch.device("cuda:0")
    w = sigma.sigma_workload()
    n, ld, nb = w.n, w.n, max(1, min(args.sigma_batch, w.nout))
    fp32s = args.sigma_precision == "fp32s"
    prec = sigma.SIGMA_FP32_SPLIT if fp32s else sigma.SIGMA_FP64
    kname = "sigma_f32s" if fp32s else "sigma_f64"
    peak = 157.3 if fp32s else 78.6  # dense FP32 / FP64 MFMA, TFLOP/s
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
    d_nu, d_gvec, d_theta, d_cw, d_wt = t(w.nu), t(w.gvec), t(w.theta), t(w.cw), t(w.weight)
    d_w = torch.empty((w.nin, n, ld), dtype=torch.complex128, device=dev)
    d_g = torch.empty((nb, w.nin, n, ld), dtype=torch.complex128, device=dev)
    d_s = torch.zeros((w.nout, n, ld), dtype=torch.complex128, device=dev)
    sigma.synthetic_device(w, 0, 0, None, d_w, d_nu, d_gvec, d_theta, d_cw, ld)

    def batch(i0):
        ni = min(nb, w.nout - i0)
        sigma.synthetic_device(w, i0, ni, d_g, None, d_nu, d_gvec, d_theta, d_cw, ld)
        sigma.conv_device(d_g, d_w, d_wt, d_s[i0:i0 + ni], n, ld, ni, w.nin, w.alpha,
                          precision=prec)
        return ni

    for _ in range(args.warmup):
        batch(0)
    torch.cuda.synchronize()
    _lib.check(L.sgw_kernel_timing(1))
    t0 = time.perf_counter()
    last = 0
    for s in range(args.steps):
        for i0 in range(0, w.nout, nb):
            batch(i0)
            last = i0
        torch.cuda.synchronize()
        print(f"step {s}: {time.perf_counter() - t0:.2f} s", flush=True)
    elapsed = time.perf_counter() - t0
    ms, cnt = ctypes.c_double(), ctypes.c_int64()
    _lib.check(L.sgw_kernel_timing_query(kname.encode(), ctypes.byref(ms), ctypes.byref(cnt)))
    _lib.check(L.sgw_kernel_timing(0))
    flops = 8.0 * n ** 3 * w.nout * w.nin * args.steps
    conv_s = ms.value / 1e3
    # error against an FP64 CPU recomputation of sampled entries of the last batch
    rng = np.random.Generator(np.random.PCG64(5))
    worst = 0.0
    for _ in range(16):
        il = int(rng.integers(0, min(nb, w.nout - last)))
        a, b = int(rng.integers(0, n)), int(rng.integers(0, n))
        grow = d_g[il, :, :, a].cpu().numpy()
        wcol = d_w[:, b, :n].cpu().numpy()
        ref = w.alpha * np.sum(w.weight * np.sum(grow * wcol, axis=1))
        got = complex(d_s[last + il, b, a].item())
        worst = max(worst, abs(got - ref) / abs(ref))
    achieved = flops / conv_s / 1e12
    print(json.dumps({
        "metric": "Sigma_c frequency convolution (N_G=1500, 64x64 GL frequencies), complex GEMM "
                  "TFLOP/s",
        "value": round(achieved, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 products, c128 frequency sum" if fp32s else "c128",
        "data": "synthetic (seeded hash G stack generated on the device; Gaussian-kernel W)",
        "config": {"workload": ("sigma-conv-fp32split" if fp32s else "sigma-conv-fp64")
                               + ": N_G 1500, 64 output x 64 convolution frequencies, "
                               + ("FP32 MFMA" if fp32s else "FP64 MFMA"),
                   "n": n, "nout": w.nout, "nin": w.nin, "batch": nb,
                   "max_rel_err_sampled_vs_fp64_cpu": worst,
                   "wall_tflops_incl_generation": round(flops / elapsed / 1e12, 3)},
        "roofline": {"kernel": "k_" + kname, "bound": "mfma", "achieved": round(achieved, 3),
                     "peak": peak, "unit": "TFLOP/s", "frac": round(achieved / peak, 4),
                     "traffic": None, "launches": cnt.value,
                     "avg_launch_ms": round(ms.value / max(cnt.value, 1), 3),
                     "flops_per_launch": 8.0 * n ** 3 * nb * w.nin},
    }))


def _gl_freq_ry(nf, wmax_ha=4.0):
    """[0] + (nf - 1) Gauss-Legendre imaginary frequencies on [0, wmax] (Ry)."""
    from sternheimergw_amd import cell
    x, _w = cell.gauss_legendre(0.0, wmax_ha, nf - 1)
    return np.concatenate([[0.0], 1j * x * cell.HA2RY])


def run_response(args):
    """Columns of the dielectric matrix (coulomb.f90:115-170 around the direct
    solve_linter) on one k+q group of the 48^3 cell for the 64 shortest
    nonzero G: dvbare psi, -P_c, the multishift solve of 64 G x 16 bands over
    [freq, -freq(2:)] (16 frequencies, 31 shifts) streamed through 256
    slots, the +/-omega average, incdrhoscf, dv_of_drho and the gather at
    100 G'.  A step is the 64 columns."""
    import torch

    from oracle import oracle as orc
    from oracle import scf_oracle as sco
    from sternheimergw_amd import _lib, cell, operators, scf

    L = _lib.lib()
    t_setup = time.perf_counter()
    w = cell.workload_c2(nq=1, nbnd=16)
    op = operators.PwOperator.from_workload(w)
    setup = scf.LinterSetup(op, w, 0)
    _a, bg, omega_cell = cell.fcc_lattice(w.alat)
    freq = _gl_freq_ry(16)
    gcut2 = 2.0 * 4.0 * 25.0
    g = np.array([(i, j, k) for i in range(-3, 4) for j in range(-3, 4) for k in range(-3, 4)])
    g = g[np.argsort((g @ bg).__pow__(2).sum(1), kind="stable")][:100]
    npert = 64
    pms = g[1:1 + npert]
    pm = pms[0]
    setup_s = time.perf_counter() - t_setup
    stats = {}

    def step():
        return scf.coulomb_columns_direct(setup, pms, freq, g, bg=bg, gcut2=gcut2,
                                          threshold=args.threshold, lmax=args.lmax, stats=stats)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    _lib.check(L.sgw_kernel_timing(1))
    stats.clear()
    t0 = time.perf_counter()
    for s in range(args.steps):
        cols = step()
        col = cols[0]
        print(f"step {s}: eps_GG(q, 0) = {col[0, 1]:.6f}", flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms, cnt = ctypes.c_double(), ctypes.c_int64()
    _lib.check(L.sgw_kernel_timing_query(b"happly", ctypes.byref(ms), ctypes.byref(cnt)))
    _lib.check(L.sgw_kernel_timing(0))
    vec = stats.get("vector_happlies", 0)
    b_launch, _bv, _bs = pruned_bytes_per_launch(w, vec / max(cnt.value, 1))
    achieved = b_launch * cnt.value / (ms.value / 1e3) / 1e9 if ms.value > 0 else 0.0
    out = {
        "metric": "dielectric-matrix columns/s (direct solve_linter + coulomb, 48^3 FFT, "
                  "16 bands, 31 shifts, 1e-8)",
        "value": round(npert * args.steps / elapsed, 4),
        "unit": "columns/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "c128",
        "data": "synthetic (seeded fcc cell, V_loc, beta projectors; FP64 diagonalisation)",
        "config": {"workload": "response-48^3: 64 eps columns per step (64 G x 16 bands streamed "
                               "through 256 slots), 16 frequencies (31 shifts), G' = 100",
                   "setup_s": round(setup_s, 2), "eps_00": [col[0, 1].real, col[0, 1].imag]},
        "roofline": {"kernel": "H-apply (pw_becp_zinv+pw_reduce_plane+pw_zfwd+pw_assemble)",
                     "bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0,
                     "unit": "GB/s", "frac": round(achieved / 8000.0, 4), "traffic": None,
                     "algorithmic_bytes_per_launch": round(b_launch),
                     "vector_happlies": vec, "launches": cnt.value,
                     "avg_launch_ms": round(ms.value / max(cnt.value, 1), 4)},
    }
    if args.cpu_baseline:
        # the oracle's band solves dominate its solve_linter: time two bands at
        # the same threshold and scale to 16
        n = w.box[0]
        nlk, nlq = w.sphere_k.box_index(w.box), w.spheres[0].box_index(w.box)
        dvb = sco.invfft(np.eye(1, n ** 3, int((pm[0] % n) + n * ((pm[1] % n) + n * (pm[2] % n))),
                                dtype=np.complex128).reshape(n, n, n))
        psi_r = sco.wave_to_box(n, nlk, w.psi_k[:2])
        rhs = sco.orthogonalize(sco.dvscf_psi(n, psi_r, dvb[None], nlq)[0], w.evq[0][:, :])
        omegas, _zf = sco.coulomb_frequencies(freq)
        o = orc.PwOp(w, 0)
        tc = time.perf_counter()
        for v in range(2):
            orc.bicgstab(o, rhs[v], -(w.eps_v[v] + omegas), lmax=args.lmax,
                         threshold=args.threshold)
        tc = time.perf_counter() - tc
        out["cpu_baseline"] = {"value": round(1.0 / (tc / 2 * w.nbnd), 6), "unit": "columns/s",
                               "cores": 1, "kind": "port",
                               "sample": f"2 of 16 band solves (31 shifts) with the oracle on one "
                                         f"core, {tc:.1f} s, scaled to 16 bands"}
    print(json.dumps(out))


def run_sigma_real(args):
    """The reference form of Sigma_c (sigma.f90:653-745) on a 24^3 correlation
    box with N_G = 1500: W continued by Pade from 4 imaginary frequencies, 4
    Coulomb frequencies (8 Green's-function frequencies) and 2 Sigma
    frequencies, 16 sigma_prods per step (invfft6 of W, G W, fwfft6)."""
    import torch

    from oracle import green_oracle as go
    from oracle import sigma_real_oracle as sro
    from sternheimergw_amd import _lib, cell, green, sigma_real

    L = _lib.lib()
    dev = torch.device("cuda:0")
    n, ng = 24, 1500
    gv = cell.shortest_millers(ng)
    grid = green.CorrGrid(n, gv)
    omega_cell = 270.011
    gen = torch.Generator(device=dev)
    gen.manual_seed(7)
    ngreen, nsig = 8, 2
    d_green = torch.randn((ngreen, n ** 3, n ** 3), dtype=torch.complex128, device=dev,
                          generator=gen) * 1e-3
    solver = 1j * np.array([0.0, 0.5, 1.5, 4.0])
    d_w = torch.randn((solver.size, ng, ng), dtype=torch.complex128, device=dev, generator=gen)
    scr = sigma_real.ScreenedCoulomb(d_w, solver, _lib.PADE_APPROX, sigma_real.NO_SYMMETRY)
    coul = 1j * np.array([0.1, 0.6, 1.5, 3.0])
    weight = np.array([0.2, 0.3, 0.3, 0.2])
    sig_freq = np.array([-0.2, 0.3])

    def step():
        return sigma_real.sigma_correlation(grid, grid, omega_cell, 0.5, 0.05, coul, weight,
                                            sig_freq, d_green, scr)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    _lib.check(L.sgw_kernel_timing(1))
    t0 = time.perf_counter()
    for s in range(args.steps):
        step()
        torch.cuda.synchronize()
        print(f"step {s}: {time.perf_counter() - t0:.2f} s", flush=True)
    elapsed = time.perf_counter() - t0
    ms, cnt = ctypes.c_double(), ctypes.c_int64()
    _lib.check(L.sgw_kernel_timing_query(b"sigma_prod", ctypes.byref(ms), ctypes.byref(cnt)))
    _lib.check(L.sgw_kernel_timing(0))
    nr = n ** 3
    bytes_per = 16.0 * (4 * ng * nr + 7 * nr * nr)
    t_prod = ms.value / 1e3 / max(cnt.value, 1)
    achieved = bytes_per / t_prod / 1e9
    prods = ngreen * nsig * args.steps
    out = {
        "metric": "Sigma_c real-space products/s (sigma_prod: invfft6 W, G*W, fwfft6; 24^3 box, "
                  "N_G 1500)",
        "value": round(prods / elapsed, 3), "unit": "sigma_prod/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "c128",
        "data": "synthetic (random G(r,r') and W(G,G',i w) on the device)",
        "config": {"workload": "sigma-real-24^3: N_G 1500, Pade W (4 points), 8 Green x 2 Sigma "
                               "frequencies per step", "n": n, "num_g": ng},
        "roofline": {"kernel": "sigma_prod (box FFT passes + index maps + product)",
                     "bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0,
                     "unit": "GB/s", "frac": round(achieved / 8000.0, 4), "traffic": None,
                     "algorithmic_bytes_per_launch": bytes_per, "launches": cnt.value,
                     "avg_launch_ms": round(t_prod * 1e3, 3)},
    }
    if args.cpu_baseline:
        rng = np.random.default_rng(3)
        G = (rng.standard_normal((nr, nr)) + 1j * rng.standard_normal((nr, nr))) * 1e-3
        W = np.zeros((nr, nr), dtype=np.complex128)
        W[:ng, :ng] = rng.standard_normal((ng, ng))
        tc = time.perf_counter()
        sro.sigma_prod(n, n, go.box_index(gv, n), go.box_index(gv, n), omega_cell, 0.5, G, W)
        tc = time.perf_counter() - tc
        out["cpu_baseline"] = {"value": round(1.0 / tc, 5), "unit": "sigma_prod/s", "cores": 1,
                               "kind": "port",
                               "sample": f"one oracle sigma_prod (fft6 on the oracle FFT) at the "
                                         f"same size on one core, {tc:.1f} s"}
    print(json.dumps(out))


def run_mixed(args):
    """The mixed-precision configuration: the 24^3 cell of the plain FP64
    configuration (6 Ha, 4 bands, one q, 8 shifts i linspace(0, 2 Ha)) with the
    plane-wave perturbations e^{i(q+G)r} of the 500 shortest G, i.e. the
    columns of chi0(G, G'; i omega) at N_G = 500: 2000 systems x 8 shifts.
    A step solves them all (right-hand sides unit norm, threshold 1e-8 on
    every shift) three ways: FP64 (the reference's arithmetic), FP32 FFTs with
    FP64 refinement (sgw_sternheimer_solve_refined), and FP32 FFTs relaxed in
    once a residual estimate is below 1e-4 (vectors FP64, no refinement).
    Reported: solves/s of each, the ratio to FP64, and max/rms errors of
    dpsi and of chi0(G, G'; i omega) = sum_v b_v^G+ dpsi_v^G' against the FP64
    solutions, which are checked against the oracle on a sample."""
    import torch

    from oracle import oracle as orc
    from sternheimergw_amd import _lib, cell, mixed, operators
    from sternheimergw_amd.batch import BatchSolver

    dev = torch.device("cuda:0")
    t_setup = time.perf_counter()
    ng = 500
    big = args.mixed_cell == 48
    if big:  # the 48^3 batch of the headline (no perturbation columns: dpsi errors only)
        w = cell.workload_c2(nq=args.nq, nbnd=args.nbnd, nshift=args.nshift,
                             normalize_rhs=False)
    else:
        w = cell.workload_c1(keep_dense=False, pert_mill=cell.shortest_millers(ng),
                             normalize_rhs=False)
    bnorm = np.linalg.norm(w.rhs, axis=1)
    rhs = w.rhs / bnorm[:, None]
    op = operators.PwOperator.from_workload(w)
    nsys, nshift, ld = w.nsys, w.nshift, w.ld
    d_rhs = torch.from_numpy(rhs).to(dev)
    d_sig = torch.from_numpy(np.ascontiguousarray(w.sigma)).to(dev)
    d_g = torch.from_numpy(w.group).to(dev)
    thr = args.threshold
    setup_s = time.perf_counter() - t_setup

    def run(kind):
        x = torch.zeros((nsys, nshift, ld), dtype=torch.complex128, device=dev)
        extra = {}
        if kind == "fp64":
            sol = BatchSolver(op, nsys, nshift, ld, lmax=args.lmax, threshold=thr,
                              shift_test=True)
            f = lambda: sol.solve(d_g, d_rhs, d_sig, w.sigma, x)
        elif kind == "relaxed":
            sol = BatchSolver(op, nsys, nshift, ld, lmax=args.lmax, threshold=thr,
                              shift_test=True, relax_threshold=args.relax_threshold)
            f = lambda: sol.solve(d_g, d_rhs, d_sig, w.sigma, x)
        else:
            sol = BatchSolver(op, nsys, nshift, ld, lmax=args.lmax, threshold=thr,
                              shift_test=True)
            f = lambda: extra.update(mixed.solve_refined(sol, d_g, d_rhs, d_sig, w.sigma, x,
                                                         inner_threshold=args.inner_threshold,
                                                         max_refine=3)[1])
        for _ in range(args.warmup):
            f()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            f()
        torch.cuda.synchronize()
        el = (time.perf_counter() - t0) / args.steps
        d_res = torch.zeros((nsys, nshift), dtype=torch.float64, device=dev)
        _lib.check(_lib.lib().sgw_residual_norms_device(
            op.handle, nsys, nshift, ld, ctypes.c_void_p(d_g.data_ptr()),
            ctypes.c_void_p(d_rhs.data_ptr()), ctypes.c_void_p(d_sig.data_ptr()),
            ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(d_res.data_ptr()), None))
        res = d_res.cpu().numpy()
        print(f"{kind}: {el * 1e3:.1f} ms per step, max residual {res.max():.2e}", flush=True)
        return x.cpu().numpy(), el, res, extra

    out64, t64, res64, _ = run("fp64")
    outr, tr, resr, stats = run("refined")
    outx, tx, resx, _ = run("relaxed")
    nsolves = nsys * nshift

    def errors(x):
        d = np.linalg.norm(x - out64, axis=2) / np.linalg.norm(out64, axis=2)
        if big:
            return {"dpsi_max_rel": float(d.max()), "dpsi_rms_rel": float(np.sqrt((d ** 2).mean()))}
        # chi0[k][G][G'] = sum_v b_v^G+ dpsi_v^G'(k) with the unnormalised b
        xb = x * bnorm[:, None, None]
        x6 = out64 * bnorm[:, None, None]
        nb, npert = w.nbnd, w.npert
        b = w.rhs.reshape(1, nb, npert, ld)[0]
        chi = np.einsum("vgl,vhkl->kgh", b.conj(), xb.reshape(nb, npert, nshift, ld))
        chi6 = np.einsum("vgl,vhkl->kgh", b.conj(), x6.reshape(nb, npert, nshift, ld))
        dc = np.abs(chi - chi6)
        return {"dpsi_max_rel": float(d.max()), "dpsi_rms_rel": float(np.sqrt((d ** 2).mean())),
                "chi0_max_abs_over_max": float(dc.max() / np.abs(chi6).max()),
                "chi0_rms_rel": float(np.sqrt((dc ** 2).sum() / (np.abs(chi6) ** 2).sum()))}

    # the FP64 solutions against the oracle on a sample of systems
    rng = np.random.Generator(np.random.PCG64(3))
    sample = rng.choice(nsys, size=8, replace=False)
    worst = 0.0
    tc = time.perf_counter()
    if big:
        sample = sample[:2]
    for s_ in sample:
        o = orc.PwOp(w, int(w.group[s_]))
        n = int(w.npw[w.group[s_]])
        xo, eo, _ = orc.bicgstab(o, rhs[s_, :n], w.sigma[s_], lmax=args.lmax, threshold=1e-12)
        rel = np.linalg.norm(out64[s_, :, :n] - xo, axis=1) / np.linalg.norm(xo, axis=1)
        worst = max(worst, float(rel.max()))
    tc = time.perf_counter() - tc
    er, ex = errors(outr), errors(outx)
    good = lambda r: int((r <= thr).sum())
    out = {
        "metric": ("Mixed-precision Sternheimer solves/s at 1e-8 residual (48^3 FFT, 32 shifts): "
                   "FP32 FFT + FP64 refinement" if big else
                   "Mixed-precision Sternheimer solves/s at 1e-8 residual (24^3 FFT, 8 shifts, "
                   "chi0 columns at N_G=500): FP32 FFT + FP64 refinement"),
        "value": round(good(resr) / tr, 2),
        "unit": "solves/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(tr * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64 local-potential FFTs, c128 vectors, scalars, projectors and residuals",
        "data": "synthetic (seeded fcc cell, V_loc, beta projectors; FP64 diagonalisation)",
        "config": {"workload": ("batched-mixed-48^3: 16 bands x 16 q x 32 shifts, 25 Ha, "
                                "BiCGStab(4), lock-step, refinement to 1e-8 per shift" if big else
                                "mixed-24^3: 4 bands x 500 G perturbations x 8 shifts, 6 Ha, "
                                "BiCGStab(4), refinement to 1e-8 per shift"),
                   "nsys": nsys, "nshift": nshift, "threshold": thr,
                   "inner_threshold": args.inner_threshold, "setup_s": round(setup_s, 2),
                   "solves_per_step": nsolves},
        "fp64": {"solves_per_s": round(good(res64) / t64, 2), "ms_per_step": round(t64 * 1e3, 3),
                 "solves_meeting_target": good(res64), "max_residual": float(res64.max()),
                 "max_rel_diff_vs_oracle_sample": worst,
                 "oracle_sample": f"{len(sample)} systems, oracle at 1e-12, {tc:.1f} s"},
        "refined": {"solves_per_s": round(good(resr) / tr, 2), "ratio_to_fp64": round(t64 / tr, 3),
                    "solves_meeting_target": good(resr), "max_residual": float(resr.max()),
                    "errors_vs_fp64": er, "stats": stats},
        "relaxed": {"solves_per_s": round(good(resx) / tx, 2), "ratio_to_fp64": round(t64 / tx, 3),
                    "relax_threshold": args.relax_threshold,
                    "solves_meeting_target": good(resx), "max_residual": float(resx.max()),
                    "errors_vs_fp64": ex},
    }
    if args.cpu_baseline:
        out["cpu_baseline"] = {"value": round(len(sample) * nshift / tc, 3), "unit": "solves/s",
                               "cores": 1, "kind": "port",
                               "sample": f"{len(sample)} systems x {nshift} shifts with the oracle "
                                         f"(FP64, threshold 1e-12) on one core, {tc:.1f} s"}
    print(json.dumps(out))


def sphere_columns(sphere):
    """Number of (m1, m2) z-columns the plane-wave sphere occupies."""
    return len({(int(m[0]), int(m[1])) for m in sphere.mill})


def pruned_bytes_per_launch(w, nvec, sc=16, sr=8):
    """Algorithmic HBM bytes of one sphere-pruned H-apply launch over nvec
    vectors (the pipeline this library runs): per vector, the plane waves in
    and out (2 N_pw) and the z-stick intermediates between the z and the
    plane passes, written once and read once in each direction
    (4 n3 N_col, N_col the sphere's occupied columns); per launch, V(r) once
    and every group's projectors once.  N_pw and N_col are the means over
    the batch's systems."""
    npw_sys = w.npw[w.group].astype(float)
    ncol = np.array([sphere_columns(s) for s in w.spheres], dtype=float)
    per_vec = 2 * npw_sys.mean() * sc + 4 * w.box[2] * ncol[w.group].mean() * sc
    shared = int(np.prod(w.box)) * sr + w.nproj * float(w.npw.sum()) * sc
    return nvec * per_vec + shared, per_vec, shared


def _cpu_worker(job):
    """Solve the given systems with the oracle on one core; returns
    (solves, seconds, iterations, operator calls)."""
    w, systems, lmax, threshold, shift_test = job
    from oracle import oracle as orc
    t_used, it_sum, calls = 0.0, 0, 0
    for s in systems:
        g = int(w.group[s])
        o = orc.PwOp(w, g)
        n = int(w.npw[g])
        orc.reset_counters()
        t0 = time.perf_counter()
        _, _, it = orc.bicgstab(o, w.rhs[s, :n], w.sigma[s], lmax=lmax, threshold=threshold,
                                shift_test=bool(shift_test))
        t_used += time.perf_counter() - t0
        calls += orc.op_calls()
        it_sum += it
    return len(systems) * w.nshift, t_used, it_sum, calls


def cpu_baseline(w, threshold, lmax, gpu_iters, shift_test, per_core=2):
    """The oracle (C restatement of bicgstab.f90 with an FFT H-apply) on the
    host's cores, one independent system per process as the reference runs
    one system per MPI rank: 2 systems per core taken at spread quantiles of
    the GPU iteration counts, each solved completely (all shifts).  The rate
    is the solves over the slowest worker's solve time (process start-up
    excluded), scaled by the ratio of the sample's mean iteration count to
    the batch's so that it stands for the whole batch."""
    import multiprocessing

    import numpy as np
    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    nsample = per_core * cores
    order = np.argsort(gpu_iters, kind="stable")
    picks = [int(order[int(len(order) * (q + 0.5) / nsample)]) for q in range(nsample)]
    jobs = [(w, picks[i::cores], lmax, threshold, shift_test) for i in range(cores)]
    print(f"cpu baseline: {nsample} systems on {cores} processes", flush=True)
    # fresh interpreters: the children never touch the GPU this process uses
    with multiprocessing.get_context("spawn").Pool(cores) as pool:
        res = pool.map(_cpu_worker, jobs)
    solves = sum(r[0] for r in res)
    wall = max(r[1] for r in res)
    it_sum = sum(r[2] for r in res)
    calls = sum(r[3] for r in res)
    cpu_s = sum(r[1] for r in res)
    scale = (it_sum / nsample) / float(np.mean(gpu_iters))
    rate = solves / wall * scale
    return {
        "value": round(rate, 3),
        "unit": "solves/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"{nsample} systems x {w.nshift} shifts at spread iteration quantiles on "
                   f"{cores} processes ({it_sum / nsample:.1f} mean iterations vs "
                   f"{float(np.mean(gpu_iters)):.1f} for the batch; rate scaled by that ratio); "
                   f"slowest process {wall:.1f} s; one core alone "
                   f"{solves / cpu_s * scale:.3f} solves/s; {calls} H-applies at "
                   f"{cpu_s / max(calls, 1) * 1e6:.0f} us"),
    }


def run_continuation(args):
    """analytic_coeff (analytic.f90:50-208) on W(G, G') of N_G^2 elements:
    robust Pade from a 32-point circle of radius 1 (W a sum of 2-4 poles
    outside radius 2), AAA and AAA with poles from the 35 imaginary
    frequencies of the reference's GW example (W a sum of 1-4 poles off the
    axis plus noise at 1e-12 of the element's largest value, threshold 1e-10).  A step is one call per model;
    the value is the chosen model's elements/s."""
    import torch

    from oracle import aaa_oracle as aa
    from oracle import pade_robust_oracle as pro
    from sternheimergw_amd import _lib, sigma_real

    L = _lib.lib()
    dev = torch.device("cuda:0")
    ng = args.cont_ng
    ne = ng * ng
    rng = np.random.default_rng(11)
    gold = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests",
                                       "golden", "aaa_reference.json")))
    circle = np.exp(2j * np.pi * np.arange(32) / 32)
    axis = 1j * np.array(gold["realistic"]["z_imag"])

    def poles_on(z, outside):
        npole = 4
        p = (rng.uniform(2.0, 4.0, (npole, ne)) * np.exp(2j * np.pi * rng.random((npole, ne)))
             if outside else rng.standard_normal((npole, ne)) * 2.0
             + 1j * rng.standard_normal((npole, ne)) * 0.5
             + np.sign(rng.standard_normal((npole, ne))) * 0.5)
        keep = rng.random((npole, ne)) < 0.75
        keep[0] = True
        c = (rng.standard_normal((npole, ne)) + 1j * rng.standard_normal((npole, ne))) * keep
        out = np.zeros((z.size, ne), np.complex128)
        for i in range(npole):
            out += c[i][None, :] / (z[:, None] - p[i][None, :])
        return out.reshape(z.size, ng, ng)

    data = {"pade_robust": (circle, poles_on(circle, True), _lib.PADE_ROBUST, 0.0)}
    w_axis = poles_on(axis, False)
    scale = np.abs(w_axis).max(axis=0, keepdims=True)
    w_axis += 1e-12 * scale * (rng.standard_normal(w_axis.shape)
                               + 1j * rng.standard_normal(w_axis.shape))
    data["aaa"] = (axis, w_axis, _lib.AAA_APPROX, 1e-10)
    data["aaa_pole"] = (axis, w_axis, _lib.AAA_POLE, 1e-10)
    times = {}
    for name, (z, W, model, thres) in data.items():
        host = torch.from_numpy(np.ascontiguousarray(W)).to(dev)
        d_w = torch.empty_like(host)
        for _ in range(args.warmup):
            d_w.copy_(host)
            sigma_real.ScreenedCoulomb(d_w, z, model, 0, thres=thres)
        torch.cuda.synchronize()
        el = 0.0
        for s in range(args.steps):
            d_w.copy_(host)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            sigma_real.ScreenedCoulomb(d_w, z, model, 0, thres=thres)
            torch.cuda.synchronize()
            el += time.perf_counter() - t0
        times[name] = el / args.steps
        print(f"{name}: {ne / times[name]:.0f} elements/s ({times[name] * 1e3:.1f} ms per call)",
              flush=True)
    name = args.cont_model
    z, W, model, thres = data[name]
    t = times[name]
    nbytes = 2.0 * 16 * z.size * ne            # read and write every element's slices
    out = {
        "metric": f"W elements continued/s ({name}, analytic_coeff, N_G {ng})",
        "value": round(ne / t, 1), "unit": "elements/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(sum(times.values()) * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "c128",
        "data": "synthetic (W as sums of poles on the model's mesh)",
        "config": {"workload": f"continuation-{name}: N_G {ng} ({ne} elements), "
                               f"{z.size} frequencies", "num_g": ng, "nfreq": int(z.size)},
        "per_model_elements_per_s": {k: round(ne / v, 1) for k, v in times.items()},
        "roofline": {"kernel": {"pade_robust": "k_pade_robust_coeffs", "aaa": "k_aaa_coeffs",
                                "aaa_pole": "k_aaa_coeffs"}[name],
                     "bound": "hbm", "achieved": round(nbytes / t / 1e9, 3), "peak": 8000.0,
                     "unit": "GB/s", "frac": round(nbytes / t / 1e9 / 8000.0, 6),
                     "traffic": None, "algorithmic_bytes_per_launch": nbytes,
                     "note": "per-element SVD/QR and Aberth work in one wavefront (half a "
                             "wavefront for robust Pade) bounded by the group reductions' "
                             "latency, not by HBM; the fraction says so"},
    }
    if args.cpu_baseline:
        ns = {"pade_robust": 10000, "aaa": 5000, "aaa_pole": 300}[name]   # ~1.5-2 s each
        flat = W.reshape(z.size, ne)
        tc = time.perf_counter()
        for e in range(ns):
            if name == "pade_robust":
                pro.pade_coeff_robust(z, flat[:, e][None, :])
            else:
                aa.aaa_coeff(4 if name == "aaa" else 5, thres, z, flat[:, e])
        tc = time.perf_counter() - tc
        out["cpu_baseline"] = {"value": round(ns / tc, 2), "unit": "elements/s", "cores": 1,
                               "kind": "port",
                               "sample": f"{ns} elements through the oracle (numpy LAPACK SVD/QR/"
                                         f"eig per element) on one core, {tc:.2f} s"}
    print(json.dumps(out))


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "sigma":
        run_sigma(args)
        return
    if args.workload == "response":
        run_response(args)
        return
    if args.workload == "sigma_real":
        run_sigma_real(args)
        return
    if args.workload == "mixed":
        run_mixed(args)
        return
    if args.workload == "continuation":
        run_continuation(args)
        return
    import torch

    from sternheimergw_amd import _lib, cell, operators

    L = _lib.lib()
    dev = torch.device("cuda:0")
    t_setup = time.perf_counter()
    w = cell.workload_c2(nq=args.nq, nbnd=args.nbnd, nshift=args.nshift)
    op = operators.PwOperator.from_workload(w)
    nsys, nshift, ld = w.nsys, w.nshift, w.ld
    d_rhs = torch.from_numpy(w.rhs).to(dev)
    d_sigma = torch.from_numpy(np.ascontiguousarray(w.sigma)).to(dev)
    h_sigma = np.ascontiguousarray(w.sigma)
    d_group = torch.from_numpy(w.group).to(dev)
    d_x = torch.zeros((nsys, nshift, ld), dtype=torch.complex128, device=dev)
    relax = args.relax_threshold if args.precision == "mixed" else 0.0
    cfg = _lib.sgw_batch_config(nsys, nshift, ld, args.lmax, 10000, args.threshold,
                                args.shift_test, relax)
    solver = L.sgw_batch_solver_create(op.handle, ctypes.byref(cfg))
    if not solver:
        raise RuntimeError(L.sgw_last_error().decode())
    slots = args.slots if args.slots > 0 else nsys
    solver_q = solver
    if args.stream and slots != nsys:  # a streaming solver with its own slot count
        cfg_q = _lib.sgw_batch_config(slots, nshift, ld, args.lmax, 10000, args.threshold,
                                      args.shift_test, relax)
        solver_q = L.sgw_batch_solver_create(op.handle, ctypes.byref(cfg_q))
        if not solver_q:
            raise RuntimeError(L.sgw_last_error().decode())
    stream = torch.cuda.current_stream().cuda_stream
    ierr = np.zeros(nsys, dtype=np.int32)
    iters = np.zeros(nsys, dtype=np.int32)
    setup_s = time.perf_counter() - t_setup

    def step():
        st = L.sgw_batch_solver_run(solver, ctypes.c_void_p(d_group.data_ptr()),
                                    ctypes.c_void_p(d_rhs.data_ptr()),
                                    ctypes.c_void_p(d_sigma.data_ptr()),
                                    ctypes.c_void_p(h_sigma.ctypes.data),
                                    ctypes.c_void_p(d_x.data_ptr()),
                                    ctypes.c_void_p(ierr.ctypes.data),
                                    ctypes.c_void_p(iters.ctypes.data), ctypes.c_void_p(stream))
        _lib.check(st)

    # streaming: the steps' batches as one queue through the nsys slots
    # (each batch is the same 8192 solves; every solve is completed and
    # checked below)
    nq = nsys * max(args.steps, args.warmup, 1)
    if args.stream:
        reps = nq // nsys
        d_rhs_q = d_rhs.repeat(reps, 1).contiguous()
        d_sigma_q = d_sigma.repeat(reps, 1).contiguous()
        h_sigma_q = np.ascontiguousarray(np.tile(h_sigma, (reps, 1)))
        h_group_q = np.ascontiguousarray(np.tile(w.group, reps), dtype=np.int32)
        d_x_q = torch.zeros((nq, nshift, ld), dtype=torch.complex128, device=dev)
        ierr_q = np.zeros(nq, dtype=np.int32)
        iters_q = np.zeros(nq, dtype=np.int32)

    def stream_run(nbatch):
        n = nbatch * nsys
        st = L.sgw_batch_solver_run_stream(solver_q, n, ctypes.c_void_p(h_group_q.ctypes.data),
                                           ctypes.c_void_p(d_rhs_q.data_ptr()),
                                           ctypes.c_void_p(d_sigma_q.data_ptr()),
                                           ctypes.c_void_p(h_sigma_q.ctypes.data),
                                           ctypes.c_void_p(d_x_q.data_ptr()),
                                           ctypes.c_void_p(ierr_q.ctypes.data),
                                           ctypes.c_void_p(iters_q.ctypes.data),
                                           ctypes.c_void_p(stream))
        _lib.check(st)

    if args.stream:
        if args.warmup:
            stream_run(args.warmup)
            print(f"warmup: {args.warmup} batches streamed, max iterations "
                  f"{iters_q[:args.warmup * nsys].max()}", flush=True)
    else:
        for i in range(args.warmup):
            step()
            print(f"warmup {i}: max iterations {iters.max()}, ierr max {ierr.max()}", flush=True)
    torch.cuda.synchronize()
    _lib.check(L.sgw_kernel_timing(args.kernel_timing))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    total_vec_happly = 0
    if args.stream:
        stream_run(args.steps)
        its = iters_q[:args.steps * nsys]
        total_vec_happly = int(2 * args.lmax * its.sum())
        iters[:] = its[:nsys]
        ierr[:] = ierr_q[:nsys]
        print(f"streamed {args.steps} batches: iterations min/median/max {its.min()}/"
              f"{int(np.median(its))}/{its.max()}", flush=True)
    else:
        for i in range(args.steps):
            step()
            total_vec_happly += int(2 * args.lmax * iters.sum())
            print(f"step {i}: iterations min/median/max {iters.min()}/{int(np.median(iters))}/"
                  f"{iters.max()}", flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0

    def query(names):
        res = {}
        for name in names:
            ms, cnt = ctypes.c_double(), ctypes.c_int64()
            _lib.check(L.sgw_kernel_timing_query(name.encode(), ctypes.byref(ms),
                                                 ctypes.byref(cnt)))
            res[name] = (ms.value, cnt.value)
        return res

    happly = query(("happly",))["happly"]
    _lib.check(L.sgw_kernel_timing(0))
    # FP64 cubic boxes run the fused pairs; the separate stages otherwise
    happly_kernels = ("pw_becp_zinv", "pw_reduce_plane", "pw_zfwd", "pw_assemble")
    if os.environ.get("SGW_PW_FUSE") == "0":
        happly_kernels = ("pw_zinv", "pw_plane", "pw_zfwd", "pw_becp", "pw_assemble")
    solver_kernels = ("bicg_v1", "bicg_seed", "shift_cycle", "mr_mgs", "mr_seed")
    breakdown = None
    if args.breakdown:
        # per-kernel split from one extra, untimed step (events around every launch)
        _lib.check(L.sgw_kernel_timing(2))
        tb = time.perf_counter()
        if args.stream:
            stream_run(1)
        else:
            step()
        torch.cuda.synchronize()
        tb = time.perf_counter() - tb
        per = query(happly_kernels + solver_kernels)
        _lib.check(L.sgw_kernel_timing(0))
        breakdown = {k: {"share": round(v[0] / 1e3 / tb, 4),
                         "avg_us": round(v[0] * 1e3 / max(v[1], 1), 2), "launches": v[1]}
                     for k, v in per.items()}

    # true residuals of every (system, shift) in FP64: the last step's batch
    # (lock-step), or every streamed batch of the timed run; a solve counts
    # when it meets the requested threshold (1e-8 by default)
    target = min(args.threshold, RES_TARGET)
    d_res = torch.zeros((nsys, nshift), dtype=torch.float64, device=dev)
    batches = ([d_x_q[b * nsys:(b + 1) * nsys] for b in range(args.steps)] if args.stream
               else [d_x])
    good, res_max = 0, 0.0
    for xb in batches:
        _lib.check(L.sgw_residual_norms_device(op.handle, nsys, nshift, ld,
                                               ctypes.c_void_p(d_group.data_ptr()),
                                               ctypes.c_void_p(d_rhs.data_ptr()),
                                               ctypes.c_void_p(d_sigma.data_ptr()),
                                               ctypes.c_void_p(xb.data_ptr()),
                                               ctypes.c_void_p(d_res.data_ptr()),
                                               ctypes.c_void_p(stream)))
        res = d_res.cpu().numpy()
        good += int((res <= target).sum())
        res_max = max(res_max, float(res.max()))
    solves_per_s = good * (1 if args.stream else args.steps) / elapsed

    # roofline of the H-apply (one event pair around its kernels on the solver stream)
    t_h = happly[0] / 1e3
    launches = happly[1]
    vpl = total_vec_happly / max(launches, 1)
    b_launch, b_vec, b_shared = pruned_bytes_per_launch(w, vpl)
    achieved = b_launch / (t_h / max(launches, 1)) / 1e9 if t_h > 0 else 0.0

    traffic, traffic_note, fp64 = None, None, None
    key = [args.nq, args.nbnd, args.nshift, args.lmax, args.threshold, args.shift_test, relax,
           args.stream, args.steps if args.stream else 1, slots if args.stream else nsys]
    if os.path.exists(TRAFFIC_FILE):
        # HBM bytes per H-apply from two rocprofv3 --pmc runs of this command
        # (FETCH_SIZE, WRITE_SIZE; each scaled by a plain copy of known size)
        with open(TRAFFIC_FILE) as f:
            prof = json.load(f)
        if prof.get("workload") == key:
            traffic = prof["traffic_bytes_per_happly"]
            kms = prof.get("happly_avg_ms")
            traffic_note = {"source": os.path.relpath(TRAFFIC_FILE, ROOT),
                            "traffic_over_algorithmic": round(traffic / b_launch, 3),
                            "measured_hbm_GBps": (round(traffic / (t_h / max(launches, 1)) / 1e9, 1)
                                                  if t_h else None),
                            # the four kernels' summed durations from the kernel
                            # trace of the same command (no launch gaps or events)
                            "profiled_kernel_ms": kms,
                            "frac_over_kernel_time": (round(b_launch / (kms * 1e-3) / 1e9 / 8000.0, 4)
                                                      if kms else None)}
    if os.path.exists(VALU_FILE):
        # FP64 VALU instructions per H-apply from rocprofv3 --pmc runs of this
        # command (SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64 of the four kernels)
        with open(VALU_FILE) as f:
            prof = json.load(f)
        if prof.get("workload") == key:
            flops = prof["fp64_flops_per_happly"]
            rate = flops / (t_h / max(launches, 1)) / 1e12 if t_h else 0.0
            fp64 = {"source": os.path.relpath(VALU_FILE, ROOT),
                    "flops_per_launch": flops, "achieved_TFLOPs": round(rate, 2),
                    "peak_TFLOPs": FP64_PEAK, "frac": round(rate / FP64_PEAK, 4)}

    out = {
        "metric": ("Sternheimer solves/s at 1e-8 residual (48^3 FFT, 32 shifts); % HBM roofline"
                   if target == RES_TARGET else
                   f"Sternheimer solves/s at {target:g} residual (48^3 FFT, 32 shifts); "
                   "% HBM roofline"),
        "value": round(solves_per_s, 2),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128" if relax == 0 else "c128 (c64 local-potential FFTs once relaxed)",
        "data": "synthetic (seeded fcc cell, V_loc, beta projectors; FP64 diagonalisation)",
        "config": {
            "workload": ("batched-fp64-48^3" if relax == 0 else "batched-mixed-48^3")
                        + ": 16 bands x 16 q x 32 shifts, 25 Ha, BiCGStab(4)",
            "box": list(w.box), "nsys": nsys, "nshift": nshift,
            "npw_min_max": [int(w.npw.min()), int(w.npw.max())], "nproj": int(w.nproj),
            "threshold": args.threshold, "lmax": args.lmax, "shift_test": args.shift_test,
            "relax_threshold": relax,
            "solves_per_step": nsys * nshift, f"solves_meeting_{target:g}": good,
            "solves_checked": len(batches) * nsys * nshift,
            "max_true_residual": res_max,
            "iterations_max": int(iters.max()), "iterations_median": float(np.median(iters)),
            "execution": (f"streamed: the {args.steps} batches as one queue of {nq} systems "
                          f"through {slots} slots" if args.stream else "lock-step per batch"),
            "setup_s": round(setup_s, 2),
        },
        "roofline": {
            "kernel": "H-apply (" + "+".join(happly_kernels) + ")",
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": 8000.0,
            "unit": "GB/s",
            "frac": round(achieved / 8000.0, 4),
            "traffic": traffic,
            "algorithmic_bytes_per_launch": round(b_launch),
            "algorithmic_bytes_per_vector": round(b_vec),
            "algorithmic_bytes_shared_per_launch": round(b_shared),
            "traffic_detail": traffic_note,
            "fp64_valu": fp64,
            "vector_happlies": total_vec_happly,
            "launches": launches,
            "vectors_per_launch": round(vpl, 1),
            "avg_launch_ms": round(t_h * 1e3 / max(launches, 1), 4),
            "happly_time_share": round(t_h / elapsed, 4),
        },
    }
    if breakdown is not None:
        out["kernel_breakdown_untimed_step"] = breakdown
    if args.copy_calibration_mib:
        # a plain copy of known size: its FETCH_SIZE/WRITE_SIZE counters scale
        # the H-apply's counters into bytes
        nel = args.copy_calibration_mib * (1 << 20) // 8
        src = torch.ones(nel, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        out["copy_calibration_bytes"] = nel * 8
    if args.cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(w, args.threshold, args.lmax, iters.copy(),
                                           args.shift_test)
    print(json.dumps(out))
    if solver_q != solver:
        L.sgw_batch_solver_destroy(solver_q)
    L.sgw_batch_solver_destroy(solver)


if __name__ == "__main__":
    main()
