/*
 * sgw.h — C ABI of libsgw_hip.so, the MI355X (gfx950) implementation of the
 * Sternheimer linear-response solve of SternheimerGW.
 *
 * Plain C: pointers and sizes only.  Complex numbers are interleaved
 * (re, im) float64 pairs, layout-compatible with Fortran COMPLEX(dp),
 * C `double _Complex` and hipDoubleComplex.  Arrays are column-major in the
 * Fortran sense: a solution block xx(n, nshift) has xx(:, k) at xx + k*n.
 *
 * Return values: a non-negative value is the solver's error code `ierr`
 * with the reference's meaning (0 converged, 1 not converged within
 * max_iter, 2 NaN in the solution); a negative value is an API error
 * (SGW_E_*) and sgw_last_error() describes it.  Where the reference aborts
 * through `errore` (shape violations), this library returns SGW_E_SHAPE.
 *
 * Every entry point runs its arithmetic on the GPU; none falls back to the
 * CPU.  The library is not thread-safe per operator/solver object.
 */
#ifndef SGW_H
#define SGW_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } sgw_complex;

#define SGW_OK 0
#define SGW_E_ARG (-1)
#define SGW_E_SHAPE (-2)
#define SGW_E_HIP (-3)
#define SGW_E_ALLOC (-4)
#define SGW_E_STATE (-5)

/* Solver identifiers of select_solver (algo/linear_solver/src/select_solver.f90:39-45). */
#define SGW_BICGSTAB_MULTI 1
#define SGW_BICGSTAB_NO_MULTI 2
#define SGW_LINEAR_SOLVER 3

/* The operator procedure argument AA(sigma, xx, Ax): Ax = (A + sigma) xx.
 * Replaces the Fortran interface block at
 * algo/linear_solver/src/select_solver.f90:77-87 (same at bicgstab.f90:164-174,
 * linear_solver.f90:91-101).  x and Ax are host arrays of length n. */
typedef void (*sgw_op_fn)(const sgw_complex *sigma, const sgw_complex *x, sgw_complex *Ax,
                          int64_t n, void *ctx);

/* TYPE(select_solver_type), algo/linear_solver/src/select_solver.f90:48-62.
 * Defaults of the reference: max_iter 10000, threshold 1e-4, bicg_lmax 4;
 * `priority` must be given (the reference aborts when it is unallocated). */
typedef struct {
  const int32_t *priority;
  int32_t npriority;
  int32_t max_iter;
  double threshold;
  int32_t bicg_lmax;
} sgw_select_solver_config;

/* TYPE(bicgstab_type), algo/linear_solver/src/bicgstab.f90:44-55. */
typedef struct {
  int32_t lmax;
  double threshold;
  int32_t max_iter;
} sgw_bicgstab_config;

/* TYPE(linear_solver_config), algo/linear_solver/src/linear_solver.f90:42-56.
 * abs_threshold is set on output (the reference's INTENT(INOUT) config). */
typedef struct {
  int32_t recover_subspace;
  int32_t max_iter;
  double threshold;
  double abs_threshold;
} sgw_linear_solver_config;

/* ---------------------------------------------------------------------------
 * Drop-in solver entry points with a host operator callback.
 * bb[n], sigma[nshift], xx[n*nshift] are host arrays.
 * ------------------------------------------------------------------------- */

/* select_solver(config, AA, bb, sigma, xx, ierr),
 * algo/linear_solver/src/select_solver.f90:67-161. */
int sgw_select_solver(const sgw_select_solver_config *config, sgw_op_fn AA, void *ctx,
                      const sgw_complex *bb, int64_t n, const sgw_complex *sigma,
                      int32_t nshift, sgw_complex *xx);

/* bicgstab(config, AA, bb, sigma, xx, ierr), algo/linear_solver/src/bicgstab.f90:154-275.
 * iters (nullable) receives the outer iteration count at exit. */
int sgw_bicgstab(const sgw_bicgstab_config *config, sgw_op_fn AA, void *ctx,
                 const sgw_complex *bb, int64_t n, const sgw_complex *sigma, int32_t nshift,
                 sgw_complex *xx, int32_t *iters);

/* linear_solver(config, AA, bb, sigma, xx, ierr),
 * algo/linear_solver/src/linear_solver.f90:82-194. */
int sgw_linear_solver(sgw_linear_solver_config *config, sgw_op_fn AA, void *ctx,
                      const sgw_complex *bb, int64_t n, const sgw_complex *sigma,
                      int32_t nshift, sgw_complex *xx);

/* ---------------------------------------------------------------------------
 * Device-resident operators.  Replaces the host callback by an operator that
 * lives on the GPU, so the solvers never leave the device.
 * ------------------------------------------------------------------------- */
typedef struct sgw_operator sgw_operator;

/* Dense operator A (n x n, column-major, host array copied to the device):
 * Ax = MATMUL(A, x) + sigma x — the reference's unit-test operator
 * linear_problem_apply, algo/linear_solver/test/linear_solver.pf:100-113. */
sgw_operator *sgw_operator_dense(const sgw_complex *a, int64_t n);

/* Host callback wrapped as an operator (vectors cross PCIe on every call). */
sgw_operator *sgw_operator_callback(sgw_op_fn AA, void *ctx, int64_t n);

/* Plane-wave shifted Hamiltonian of linear_op (algo/linear_solver/src/linear_op.f90:46-206)
 * with QE's h_psi (kinetic + local via FFT + nonlocal projectors), in Rydberg:
 *   Ax = |k+q+G|^2 x + fwfft(V_loc(r) invfft(x)) + sum_p proj_p D_p <proj_p|x> + sigma x.
 * The projector list of group g holds the beta functions (D = the diagonal
 * nonlocal coefficients) followed, for the Coulomb operator, by the valence
 * states at k+q with D = alpha_pv, which is the alpha_pv P_v term of
 * linear_op.f90:131-138 (S = 1, norm-conserving).  One group per k+q point. */
typedef struct {
  int32_t n1, n2, n3;          /* FFT box                                          */
  const double *vloc;          /* V_loc(r), n1*n2*n3, index i1 + n1*(i2 + n2*i3)   */
  int32_t ngroup;              /* number of k+q points                             */
  const int32_t *npw;          /* plane waves per group                            */
  int64_t ld;                  /* padded vector length, >= max npw                 */
  const int32_t *mill;         /* [ngroup][ld][3] Miller indices (m1, m2, m3)       */
  const double *kin;           /* [ngroup][ld] |k+q+G|^2 in Ry                     */
  int32_t nproj;               /* projectors per group                             */
  const sgw_complex *proj;     /* [ngroup][nproj][ld]                              */
  const double *dcoef;         /* [ngroup][nproj] diagonal coefficients (Ry)       */
} sgw_pw_desc;
sgw_operator *sgw_operator_pw(const sgw_pw_desc *desc);

void sgw_operator_destroy(sgw_operator *op);

/* Precision of a plane-wave operator's local-potential FFTs in every later
 * application: 0 FP64 (default), 1 FP32 (input, output and the other terms
 * stay FP64).  No counterpart in the reference (FP64 only); used to measure
 * and test the FP32 pipeline that the mixed-precision solve relaxes to. */
int sgw_operator_set_fft_precision(sgw_operator *op, int32_t mode);

/* ax[v] = (A_{group[v]} + sigma[v]) x[v], v < nvec; host arrays with leading
 * dimension ldx (>= operator ld).  group may be NULL for one-group operators. */
int sgw_operator_apply(sgw_operator *op, int32_t nvec, const int32_t *group,
                       const sgw_complex *sigma, const sgw_complex *x, int64_t ldx,
                       sgw_complex *ax);

/* Same as sgw_operator_apply on device arrays, on `stream` (hipStream_t, NULL
 * = default stream); no synchronisation. */
int sgw_operator_apply_device(sgw_operator *op, int32_t nvec, const int32_t *d_group,
                              const sgw_complex *d_sigma, const sgw_complex *d_x, int64_t ldx,
                              sgw_complex *d_ax, void *stream);

/* Size of the fixed per-device work array of sgw_operator_apply_device_noalloc. */
#define SGW_NOALLOC_SCRATCH_BYTES (64u << 20)

/* sgw_operator_apply_device for plane-wave operators without any device
 * allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The work buffers (z-stick and
 * plane transforms, projector partial sums, vector lists) are carved from
 * one array of SGW_NOALLOC_SCRATCH_BYTES per device, which the library holds
 * until the process ends; a batch needing more returns SGW_E_SHAPE (at most
 * 1024 vectors; on the 48^3 box with N_pw ~ 1700 about 120 vectors fit).
 * Every call on a device shares that array: calls that overlap on one device
 * (other streams or threads) race.  The call waits for earlier work on
 * `stream` before it starts and returns with the application queued. */
int sgw_operator_apply_device_noalloc(sgw_operator *op, int32_t nvec, const int32_t *d_group,
                                      const sgw_complex *d_sigma, const sgw_complex *d_x,
                                      int64_t ldx, sgw_complex *d_ax, void *stream);

/* select_solver on a device operator (group selects the k+q point). */
int sgw_select_solver_op(const sgw_select_solver_config *config, sgw_operator *op,
                         int32_t group, const sgw_complex *bb, int64_t n,
                         const sgw_complex *sigma, int32_t nshift, sgw_complex *xx);

/* ---------------------------------------------------------------------------
 * Batched multishift Sternheimer solve: nsys independent systems
 *   (A_{group[s]} + sigma[s][k]) x[s][k] = rhs[s],  k < nshift,
 * each advanced by the multishift BiCGStab(l) recurrences of bicgstab.f90
 * with its own convergence test (||r_seed|| < threshold, bicgstab.f90:237,245),
 * all systems in every kernel launch.  This is the batched form of the band
 * loop of solve_linter (phys/coul/src/solve_linter.f90:367-374, sigma =
 * -(eps_v + omega)) and of the G' loop of green_function
 * (phys/green/src/green.f90:196-218, sigma = -omega).
 * ------------------------------------------------------------------------- */
typedef struct {
  int32_t nsys;        /* systems (bands x q, or G' columns)           */
  int32_t nshift;      /* shifts per system; shift 0 is the seed      */
  int64_t ld;          /* leading dimension of rhs and x (>= op ld)   */
  int32_t lmax;        /* BiCGStab(l) degree, 1..15                   */
  int32_t max_iter;
  double threshold;    /* absolute, on the seed residual              */
  int32_t shift_test;  /* 0: the reference's test on the seed residual
                          only; 1: also every shifted residual, through
                          |1/(theta phi)| ||r_seed|| (no counterpart in
                          the reference, which can leave a shift above
                          the threshold)                               */
  double relax_threshold; /* 0: FP64 throughout.  > 0: mixed precision
                          (plane-wave operators): once a system's
                          residual estimate is below relax_threshold its
                          local-potential FFTs run in FP32 (relaxed
                          matrix-vector products); vectors, recurrences
                          and the convergence test stay FP64             */
} sgw_batch_config;

typedef struct sgw_batch_solver sgw_batch_solver;

sgw_batch_solver *sgw_batch_solver_create(sgw_operator *op, const sgw_batch_config *cfg);
void sgw_batch_solver_destroy(sgw_batch_solver *solver);

/* Device arrays: d_group[nsys], d_rhs[nsys][ld], d_sigma[nsys][nshift],
 * d_x[nsys][nshift][ld]; h_sigma is a host copy of d_sigma.  h_ierr and
 * h_iters (host, nullable) receive per-system ierr and iteration counts.
 * Synchronises `stream` before returning. */
int sgw_batch_solver_run(sgw_batch_solver *solver, const int32_t *d_group,
                         const sgw_complex *d_rhs, const sgw_complex *d_sigma,
                         const sgw_complex *h_sigma, sgw_complex *d_x, int32_t *h_ierr,
                         int32_t *h_iters, void *stream);

/* A queue of nq systems through the solver's nsys slots (the callers' loops
 * over bands, q and G' of solve_linter.f90:366-374 and green.f90:196-218
 * as one stream).  Each operator group keeps a fixed share of the slots, and
 * a slot whose system has finished takes the next queued system of its
 * group, so the batch stays full until the queue runs dry.  Every system
 * runs exactly the iterations of sgw_batch_solver_run (a system's iterates do
 * not depend on the other slots).  h_group[nq] host; d_rhs[nq][ld],
 * d_sigma/h_sigma[nq][nshift], d_x[nq][nshift][ld]; h_ierr/h_iters[nq].
 * Device-resident operators.  With relax_threshold > 0 (plane-wave
 * operators) a slot's local-potential FFTs switch to FP32 once its residual
 * estimate is below it, and a refilled slot starts again in FP64.  A system
 * that reaches max_iter stops in the same iteration as in the lock-step
 * solve (ierr 1). */
int sgw_batch_solver_run_stream(sgw_batch_solver *solver, int32_t nq, const int32_t *h_group,
                                const sgw_complex *d_rhs, const sgw_complex *d_sigma,
                                const sgw_complex *h_sigma, sgw_complex *d_x, int32_t *h_ierr,
                                int32_t *h_iters, void *stream);

/* One-shot form on host arrays (copies in and out). */
int sgw_sternheimer_solve_batched(sgw_operator *op, const sgw_batch_config *cfg,
                                  const int32_t *group, const sgw_complex *rhs,
                                  const sgw_complex *sigma, sgw_complex *x, int32_t *ierr,
                                  int32_t *iters);

/* Mixed-precision batched solve with FP64 iterative refinement.  No
 * counterpart in the reference (FP64 throughout); the FP64 steps are those of
 * bicgstab.f90:154-275.  On `solver` (plane-wave operator) it
 *   1. solves the multishift systems with the local-potential FFTs in FP32
 *      down to `inner_threshold` (the solver's seed/shift test);
 *   2. forms every shift's residual r = b - (A + sigma) x in FP64;
 *   3. for each (system, shift) above the solver's threshold (relative to
 *      ||b||), solves the single-shift correction (A + sigma) d = r with FP32
 *      FFTs, streamed through `correction_slots` slots, and adds d to x;
 *      then back to 2, at most max_refine times.
 * The shifts need step 3 because FP32 products break the collinearity of the
 * shifted residuals that the multishift recurrences rely on.  Vectors,
 * recurrence scalars, the projector terms (including alpha_pv P_v) and the
 * residuals are FP64.  Arrays as in sgw_batch_solver_run.  h_ierr[nsys]
 * (nullable): 0 when every shift's FP64 residual meets the threshold;
 * h_stats[6] (nullable): FP32 H-applies of step 1 and of step 3, FP64
 * residual applications, refinement rounds, step-1 iterations, step-3
 * iterations; h_res[nsys*nshift] (nullable): final relative residuals. */
int sgw_sternheimer_solve_refined(sgw_batch_solver *solver, const int32_t *d_group,
                                  const sgw_complex *d_rhs, const sgw_complex *d_sigma,
                                  const sgw_complex *h_sigma, sgw_complex *d_x,
                                  double inner_threshold, int32_t max_refine,
                                  int32_t correction_slots, int32_t *h_ierr, int64_t *h_stats,
                                  double *h_res, void *stream);

/* Adjoint of the batched solve (no counterpart in the reference, which has no
 * derivatives).  For the objective
 *   L = sum_{s,k} weight[s][k] Re <c[s] | x[s][k]>,
 * with x the stored solutions of a forward sgw_batch_solver_run of `solver`
 * (rhs b, shifts sigma), it solves the backward systems
 *   (A_{group[s]} + conj(sigma[s][k])) y[s][k] = c[s]
 * with the same solver and configuration, and, for a plane-wave operator,
 * writes the gradient of L with respect to V_loc(r) on the FFT box:
 *   d_grad[i1 + n1*(i2 + n2*i3)] = -(1/N) sum_{s,k} weight[s][k] Re[conj(y(r)) x(r)],
 * x(r) = invfft(x(G)) (QE sign, no factor).  Device arrays: d_c[nsys][ld],
 * d_sigma[nsys][nshift] (h_sigma its host copy), d_weight[nsys][nshift] (real),
 * d_x and d_y [nsys][nshift][ld]; d_grad may be NULL (no gradient).
 * h_objective (nullable) receives L; h_ierr (nullable) the backward ierr
 * per system.  Returns the largest backward ierr or an API error. */
int sgw_sternheimer_adjoint(sgw_batch_solver *solver, const int32_t *d_group,
                            const sgw_complex *d_c, const sgw_complex *d_sigma,
                            const sgw_complex *h_sigma, const double *d_weight,
                            const sgw_complex *d_x, sgw_complex *d_y, double *d_grad,
                            double *h_objective, int32_t *h_ierr, void *stream);

/* Density response of the solutions of one k+q group, the accumulation that
 * follows the solve in solve_linter (phys/coul/src/solve_linter.f90:486-497,
 * QE incdrhoscf per frequency):
 *   drho[k](r) (+)= weight * sum_{v < nbnd} conj(psi_v(r)) dpsi[v][k](r),
 * psi_v(r), dpsi(r) = invfft of plane-wave coefficients (QE sign, no factor).
 * Device arrays: d_dpsi[nbnd][nshift][ld] in the order of `group`'s sphere
 * (the batched solver's output for the systems of one q), d_psi_r[nbnd][box]
 * and d_drho[nshift][box] with box index i1 + n1*(i2 + n2*i3).  QE's weight
 * is 2 wk / Omega.  Cubic 24/32/36/48 boxes. */
int sgw_drho_accumulate(sgw_operator *op, int32_t group, int32_t nbnd, int32_t nshift,
                        const sgw_complex *d_dpsi, int64_t ld, const sgw_complex *d_psi_r,
                        double weight, sgw_complex *d_drho, int32_t accumulate, void *stream);

/* True relative residuals ||rhs[s] - (A + sigma[s][k]) x[s][k]|| / ||rhs[s]||
 * recomputed in FP64 on the device for every (system, shift); device arrays,
 * result d_res[nsys][nshift]. */
int sgw_residual_norms_device(sgw_operator *op, int32_t nsys, int32_t nshift, int64_t ld,
                              const int32_t *d_group, const sgw_complex *d_rhs,
                              const sgw_complex *d_sigma, const sgw_complex *d_x,
                              double *d_res, void *stream);

/* ---------------------------------------------------------------------------
 * The self-consistent response around the solves (solve_linter.f90:279-583).
 * Boxes: cubic n^3, n in {16, 18, 24, 27, 32, 36, 48}, [nvec][i3][i2][i1] (QE order
 * i1 + n1*(i2 + n2*i3)); device arrays; asynchronous on `stream` unless
 * stated.  `d_nl[i]` is the box index of plane wave i of a sphere (QE's
 * dffts%nl(igk_k(i))).
 * ------------------------------------------------------------------------- */

/* QE fwfft/invfft of nvec boxes in place: sign +1 = invfft (sum e^{+iGr}, no
 * factor), sign -1 = fwfft (sum e^{-iGr} / n^3).  FFTXlib (external). */
int sgw_box_fft(int32_t n, int32_t nvec, int32_t sign, sgw_complex *d_box, void *stream);

/* psi(r) = invfft of the sphere coefficients d_psi[nvec][ld] (i < npw)
 * scattered through d_nl: QE cft_wave(+1) / the evc -> aux step of
 * dvqpsi_us.f90:99-108 and solve_linter.f90:394. */
int sgw_wave_to_box(int32_t n, int32_t nvec, int32_t npw, const int32_t *d_nl,
                    const sgw_complex *d_psi, int64_t ld, sgw_complex *d_box, void *stream);

/* psi[i] (+)= fwfft(box)[d_nl[i]] for i < npw (QE cft_wave(-1), which adds);
 * d_box is overwritten by its transform; with accumulate = 0 the padding
 * npw <= i < ld is zeroed. */
int sgw_box_to_wave(int32_t n, int32_t nvec, int32_t npw, const int32_t *d_nl,
                    sgw_complex *d_box, sgw_complex *d_psi, int64_t ld, int32_t accumulate,
                    void *stream);

/* dV psi on the k+q sphere for every frequency and band (dvqpsi_us.f90:99-132
 * for the bare potential; solve_linter.f90:393-398 for dvscf):
 *   dvpsi[f][v] = base[v] + gather_{k+q}( fwfft( psi_r[v](r) * dvscf[f](r) ) ),
 * base[v] (sphere order of k+q, ld) may be NULL (then no base term).
 * psi_r[nbnd][box] = invfft of the unperturbed states at k; dvscf[nfreq][box].
 * Synchronises `stream` (scratch of at most 512 boxes). */
int sgw_dvscf_psi(int32_t n, int32_t nbnd, int32_t nfreq, const sgw_complex *d_psi_r,
                  const sgw_complex *d_dvscf, int32_t npwq, const int32_t *d_nlq,
                  const sgw_complex *d_base, sgw_complex *d_dvpsi, int64_t ld, void *stream);

/* QE orthogonalize (LR_Modules, called at solve_linter.f90:337, 411) for
 * insulators with S = 1: dvpsi <- -(1 - |evq><evq|) dvpsi = -P_c^+ dvpsi for
 * nvec vectors; evq[nocc][ld] the occupied states at k+q, nocc <= 64. */
int sgw_orthogonalize(int32_t nvec, int32_t npw, int32_t nocc, const sgw_complex *d_evq,
                      sgw_complex *d_dvpsi, int64_t ld, void *stream);

/* The +/-omega average of solve_linter.f90:462-480 with the wg/wk rescale of
 * :373 folded into `scale`: out[v][f] = scale * x(+omega_f) for f = 0 when
 * zero_freq, else scale/2 (x(+omega_f) + x(-omega_f)), out[nbnd][nf][ld].
 * layout 0: x[nbnd][nshift][ld] over the shift list of :237-252
 * ([freq, -freq(2:)] when zero_freq, nshift = 2nf-1; else [freq, -freq]);
 * layout 1: x[nbnd][nf][2][ld], shift 0 = +omega_f, shift 1 = -omega_f. */
int sgw_average_pm(int32_t nbnd, int32_t nf, int32_t zero_freq, int32_t layout,
                   const sgw_complex *d_x, int64_t ld, double scale, sgw_complex *d_out,
                   void *stream);

/* Hartree response potential of QE's dv_of_drho with lrpa (called at
 * solve_linter.f90:556; lrpa set at main/src/gwq_readin.f90:223), in place:
 *   dv(r) = invfft[ e2 4 pi / |q+G|^2 fwfft(drho)(G) ],  e2 = 2 (Ry),
 * over |G|^2 <= gcut2 (QE's density sphere, bohr^-2), q + G = 0 dropped.
 * zero_mean = 1 also drops G = 0 (the meandvb < eps10 branch,
 * solve_linter.f90:545-549).  bg: b1, b2, b3 Cartesian (bohr^-1, 2 pi
 * included), xq Cartesian (bohr^-1).  Miller index of box index i: i for
 * i <= n/2, i - n above. */
int sgw_dv_of_drho(int32_t n, int32_t nvec, const double *bg, const double *xq, double gcut2,
                   int32_t zero_mean, sgw_complex *d_box, void *stream);

/* Modified Broyden mixing of complex potentials, mix_potential_c
 * (phys/coul/src/mix_pot_c.f90:25-190; D. D. Johnson, PRB 38, 12807), same
 * arguments: on return vout = vout - vin, vin = the mixed potential,
 * *dr2 = (||vout - vin|| / ndim)^2, *conv = dr2 < tr2 (vin unchanged then).
 * The history of earlier iterations (the reference's SAVE arrays) lives in
 * the sgw_broyden state: n_iter <= 8 stored iterations of ndim elements.
 * Synchronises `stream`. */
typedef struct sgw_broyden sgw_broyden;
sgw_broyden *sgw_broyden_create(int64_t ndim, int32_t n_iter);
void sgw_broyden_destroy(sgw_broyden *b);
int sgw_mix_potential_c(sgw_broyden *b, int64_t ndim, sgw_complex *d_vout, sgw_complex *d_vin,
                        double alphamix, double *dr2, double tr2, int32_t iter, int32_t n_iter,
                        int32_t *conv, void *stream);

/* ---------------------------------------------------------------------------
 * Green's function of the self-energy: assembly (green.f90:196-218) and the
 * 6-D Fourier transforms (data/fft/src/fft6.f90).  A two-point function is a
 * Fortran matrix f(nr, nr_p) per frequency (first index fastest), nr = n^3,
 * nr_p = n_p^3 (the correlation boxes corr_fft and corr_par_fft, cubic, n in
 * {16, 18, 24, 27, 32, 36, 48}); in reciprocal space only the (num_g,
 * num_g_p) top-left block is defined.  d_nl / d_nl_p: box index of each G
 * (QE's dfft%nl, 0-based).  nmat matrices back to back; synchronises.
 * ------------------------------------------------------------------------- */

/* invfft6 (fft6.f90:215-345):
 *   f(r, r') = 1/omega sum_{G,G'} e^{-iGr} f(G, G') e^{+iG'r'}. */
int sgw_invfft6(int32_t n, int32_t n_p, int32_t num_g, const int32_t *d_nl, int32_t num_g_p,
                const int32_t *d_nl_p, double omega, int32_t nmat, sgw_complex *d_f,
                void *stream);

/* fwfft6 (fft6.f90:84-181):
 *   f(G, G') = omega / (nr nr_p) sum_{r,r'} e^{+iGr} f(r, r') e^{-iG'r'}. */
int sgw_fwfft6(int32_t n, int32_t n_p, int32_t num_g, const int32_t *d_nl, int32_t num_g_p,
               const int32_t *d_nl_p, double omega, int32_t nmat, sgw_complex *d_f,
               void *stream);

/* green(nr, nr_p, nfreq) from the batched solutions of green_function's
 * columns (green.f90:184-186, 196-218):
 *   green(r, c, f) = x[colsys[c]][f][rowmap[r]]  for r < nrow, c < ncol with
 *   both maps >= 0, and zero elsewhere.
 * rowmap[r]: 0-based position of correlation G r in the k-point's sphere, -1
 * when the reference's mask (0 < map < num_g, :212) excludes it; colsys[c]:
 * the batched system that solved column c (-1 when green.f90:200 skips it).
 * x[nsys][nshift][ld] with shift f = frequency f (nfreq <= nshift). */
int sgw_green_assemble(int32_t nfreq, int32_t n, int32_t n_p, int32_t nrow,
                       const int32_t *d_rowmap, int32_t ncol, const int32_t *d_colsys,
                       const sgw_complex *d_x, int32_t nshift, int64_t ld, sgw_complex *d_green,
                       void *stream);

/* ---------------------------------------------------------------------------
 * The reference form of Sigma_c: analytic continuation of W
 * (algo/analytic/src/analytic.f90) and the real-space product sigma_prod
 * (phys/corr/src/sigma.f90:417-523).  W arrays are Fortran
 * scrcoul(num_g, num_g, nfreq) (first index fastest), device memory.
 * ------------------------------------------------------------------------- */
#define SGW_GODBY_NEEDS 1   /* analytic.f90:33, godby_needs.f90 */
#define SGW_PADE_APPROX 2   /* analytic.f90:36, pade.f90 (Vidberg-Serene) */
#define SGW_PADE_ROBUST 3   /* analytic.f90:39, pade_robust.f90 (SVD-based) */
#define SGW_AAA_APPROX 4    /* analytic.f90:42, vendor/analytic/src/aaa.f90 */
#define SGW_AAA_POLE 5      /* analytic.f90:45, AAA with pole_correction */

/* analytic_coeff (analytic.f90:50-208) in place: W on the frequency mesh
 * becomes the model's coefficients; thres is its threshold argument (tr2_gw,
 * sigma.f90:333), used by the AAA models only.  Godby-Needs: nfreq = 2, h_z[1] = i
 * omega_p (freq%solver(2)).  Pade: the mesh of freqbins_symm, h_z[nfreq]
 * (nfreq <= 64), and h_src[i] = the slice of d_w that holds W(h_z[i]) (the
 * even-symmetry copy of freqbins_symm; NULL = i).  Robust Pade
 * (pade_coeff_robust, pade_robust.f90:89-160): h_z = the solver frequencies,
 * 10 <= nfreq <= 64 equally spaced points R e^{2 pi i f / nfreq} on a circle
 * (checked: |z_f| = R = Re z_0 within 1e-12), h_src = NULL; each element's
 * slices become [m, n, a_0..a_m, b_0..b_n, 0...] of the approximant
 * (sum a_i z^i) / (sum b_i z^i).  AAA (3 <= nfreq <= 64, the mesh and
 * h_src of freqbins_symm as for Pade): SGW_AAA_APPROX keeps at most
 * nfreq / 3 support points and stores [positions, values, weights] in
 * thirds of nfreq / 3 slices (analytic.f90:174-193); SGW_AAA_POLE keeps the
 * poles whose residue exceeds thres as [poles, residues] in halves of
 * nfreq / 2 (pole_correction, analytic.f90:389-427).  Synchronises.  Errors
 * (SGW_E_STATE, the element's slices NaN): more relevant poles than
 * nfreq / 2 (the reference's errore), or an SGW_AAA_POLE element that is
 * not converged at nfreq / 2 support points (the Loewner submatrix would
 * have more columns than rows and no unique null vector). */
int sgw_analytic_coeff(int32_t model, double thres, int32_t num_g, int32_t nfreq,
                       const sgw_complex *h_z, const int32_t *h_src, sgw_complex *d_w,
                       void *stream);

/* analytic_eval (analytic.f90:211-348): out(nr, nr_p) = 0, then
 * out(r, c) = W_model(freq; coeff(gmap[r], gmap_p[c], :)) for r < num_g,
 * c < num_g_p (gmap: gmapsym, gmap_p: gmapsym(fft_map), 0-based).  freq is
 * the symmetrised output frequency (freq_in%symmetrize).  Robust Pade
 * (pade_eval_robust, pade_robust.f90:37-86) reads the degrees from the first
 * two slices (h_z may be NULL); an element whose degrees do not fit the
 * nfreq slices evaluates to NaN.  AAA (aaa_approx_eval, aaa_pole_eval,
 * analytic.f90:351-459; h_z may be NULL): the barycentric form of the
 * support points whose weight exceeds 1e-12 (a support point's value within
 * 1e-14 of it), or the sum of residue / (freq - pole). */
int sgw_analytic_eval(int32_t model, int32_t num_gc, int32_t nfreq, const sgw_complex *h_z,
                      const sgw_complex *d_coeff, sgw_complex freq, int32_t n, int32_t n_p,
                      int32_t num_g, int32_t num_g_p, const int32_t *d_gmap,
                      const int32_t *d_gmap_p, sgw_complex *d_out, void *stream);

/* sigma_prod (sigma.f90:417-523) on one W: d_array (nr, nr_p) holds W(G, G')
 * in its (num_g, num_g_p) block on input and alpha fwfft6[G(r,r') W(r,r')]
 * there on output; d_green is G(r, r') (nr, nr_p).  With d_sigma (ld_sigma >=
 * num_g) the block is also added to Sigma (sigma.f90:741).  Synchronises. */
int sgw_sigma_prod(int32_t n, int32_t n_p, int32_t num_g, const int32_t *d_nl, int32_t num_g_p,
                   const int32_t *d_nl_p, double omega, sgw_complex alpha,
                   const sgw_complex *d_green, sgw_complex *d_array, sgw_complex *d_sigma,
                   int64_t ld_sigma, void *stream);

/* ---------------------------------------------------------------------------
 * Sigma_c frequency convolution (phys/corr/src/sigma.f90:705-745, the
 * accumulation sigma(:,:,isigma) += alpha * w * [G (x) W] over the Green's
 * function frequencies), restated as a batched complex GEMM on the device:
 *   Sigma[i] (+)= alpha * sum_j weight[j] * G[i][j] * W[j],
 * i < nout, j < nin, every matrix n x n complex, column-major, leading
 * dimension ld; G[i][j] at d_g + i*g_stride_i + j*g_stride_j, W[j] at
 * d_w + j*w_stride, Sigma[i] at d_sigma + i*s_stride (element offsets).
 * The reference forms this product in real space (sigma_prod,
 * sigma.f90:417-523); here it is the G'' contraction in reciprocal space.
 * ------------------------------------------------------------------------- */
#define SGW_SIGMA_FP64 0   /* FP64 operands, FP64 MFMA, FP64 accumulation */
#define SGW_SIGMA_FP32_SPLIT 1  /* operands rounded to FP32, FP32 MFMA; each
                                   frequency's sum FP32, the sum over
                                   frequencies FP64 (no counterpart in the
                                   reference, which is FP64) */

typedef struct {
  int32_t n;
  int64_t ld;
  int32_t nout, nin;
  int64_t g_stride_i, g_stride_j, w_stride, s_stride;
  int32_t precision;   /* SGW_SIGMA_*                                    */
  int32_t accumulate;  /* 1: Sigma += ..., 0: Sigma = ...                */
} sgw_sigma_conv_desc;

/* Device arrays; d_weight[nin] (real) on the device.  Asynchronous on `stream`. */
int sgw_sigma_conv_device(const sgw_sigma_conv_desc *desc, const sgw_complex *d_g,
                          const sgw_complex *d_w, const double *d_weight, sgw_complex alpha,
                          sgw_complex *d_sigma, void *stream);

/* Seeded synthetic inputs of the convolution benchmark, generated on the
 * device (the full G stack at n = 1500, 64 x 64 frequencies is 147 GB):
 * G[i][j], i0 <= i < i0 + ni (written at d_g + (i - i0)*g_stride_i + ...),
 * from a counter-based hash scaled by 1 / ((1 + nu[i*nin + j]) sqrt(n));
 * W[j](a, b) = cw[j] exp(-|g_a - g_b|^2 / (2 length^2)) e^{i(theta_a - theta_b)}
 * / ((1 + |g_a|^2)(1 + |g_b|^2)), Hermitian positive semidefinite.
 * Either output may be NULL. */
int sgw_sigma_synthetic_device(int32_t n, int64_t ld, int32_t i0, int32_t ni, int32_t nout,
                               int32_t nin, uint64_t seed, const double *d_nu,
                               sgw_complex *d_g, int64_t g_stride_i, int64_t g_stride_j,
                               const double *d_gvec, const double *d_theta, const double *d_cw,
                               double length, sgw_complex *d_w, int64_t w_stride, void *stream);

/* Kernel timing with HIP events recorded on each kernel's launch stream.
 * sgw_kernel_timing(level) sets the level and resets the totals: 0 off,
 * 1 one event pair per operator application ("happly"), 2 also one pair per
 * kernel launch ("pw_zinv", "pw_plane", "pw_zfwd", "pw_becp", "pw_assemble",
 * "bicg_v1", "bicg_seed" (the two seed halves of a BiCG step, each with its
 * dot product), "shift_cycle", "mr_mgs", "mr_seed").
 * Query returns the summed duration (ms) and launch count of one name. */
int sgw_kernel_timing(int enable);
int sgw_kernel_timing_query(const char *name, double *total_ms, int64_t *launches);

/* Launch-shape autotuning.  No counterpart in the reference.  The split plane
 * transforms of cubic boxes up to 48 (three planes on two wavefronts, one
 * plane per wavefront, one plane on two, two planes on two) and the final assembly of
 * an application (16, 32 or 64 plane waves per block) have several launch
 * shapes with the same results; the first application of each size class (batch size
 * rounded up to a power of two) times every shape once and keeps the fastest.
 * SGW_AUTOTUNE=0 in the environment keeps the first shape.  Returns the
 * candidate index chosen for `key` ("plane<box>/M<m>/b<band>/n<class>",
 * "assemble/n<class>/g<groups>/p<projectors>"), -1 if none yet. */
int sgw_autotune_choice(const char *key);

const char *sgw_last_error(void);
const char *sgw_version(void);

#ifdef __cplusplus
}
#endif
#endif
