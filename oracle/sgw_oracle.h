/*
 * CPU FP64 oracle for the Sternheimer linear-response solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only the test
 * suite, the smoke check and the CPU-baseline leg of the benchmark may load
 * it.  The product (libsgw_hip.so) never links, loads or calls it.
 *
 * It restates, in plain C, the algorithms of SternheimerGW's solver path:
 *   oracle_norm           data/algebra/src/norm.f90:73-107        (ZLANGE 'F')
 *   oracle_gram_schmidt   data/algebra/src/gram_schmidt.f90:35-134
 *   oracle_bicgstab       algo/linear_solver/src/bicgstab.f90:154-999
 *   oracle_linear_solver  algo/linear_solver/src/linear_solver.f90:82-503
 *   oracle_select_solver  algo/linear_solver/src/select_solver.f90:67-202
 *   oracle_pw_apply       algo/linear_solver/src/linear_op.f90:46-206 composed
 *                         with the kinetic + local + nonlocal Hamiltonian that
 *                         Quantum ESPRESSO's h_psi applies (external, not in tree)
 *   oracle_fft3d          QE invfft/fwfft sign and 1/N conventions
 *
 * All complex data are interleaved (re, im) float64, vectors contiguous.
 */
#ifndef SGW_ORACLE_H
#define SGW_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } ozc;

/* Operator callback: Ax = (A + sigma) x for a vector of length n. */
typedef void (*oracle_op_fn)(const ozc *sigma, const ozc *x, ozc *Ax,
                             int64_t n, void *ctx);

double oracle_norm(const ozc *v, int64_t n);
ozc oracle_zdotu(const ozc *x, const ozc *y, int64_t n);
ozc oracle_zdotc(const ozc *x, const ozc *y, int64_t n);

void oracle_gram_schmidt(int first, ozc *basis, int64_t n, int nbasis, ozc *vector);

/* Returns ierr (0 ok, 1 max_iter, 2 NaN).  iters_out (nullable) receives the
 * number of outer iterations performed (the loop counter at exit). */
int oracle_bicgstab(int lmax, double threshold, int max_iter,
                    oracle_op_fn op, void *ctx, const ozc *b, int64_t n,
                    const ozc *sigma, int nshift, ozc *x, int *iters_out);

/* oracle_bicgstab with shift_test = 1: a solve also waits for every shifted
 * residual estimate |1/(theta phi)| ||r_seed|| to pass (an extension; 0 is
 * the reference's test).  A shifted system whose estimate is below half the
 * threshold at the end of an iteration keeps its vectors from then on. */
int oracle_bicgstab_ex(int lmax, double threshold, int max_iter,
                       oracle_op_fn op, void *ctx, const ozc *b, int64_t n,
                       const ozc *sigma, int nshift, ozc *x, int *iters_out, int shift_test);

int oracle_linear_solver(int max_iter, double threshold,
                         oracle_op_fn op, void *ctx, const ozc *b, int64_t n,
                         const ozc *sigma, int nshift, ozc *x, int *basis_out);

int oracle_select_solver(const int *priority, int npriority, int max_iter,
                         double threshold, int bicg_lmax,
                         oracle_op_fn op, void *ctx, const ozc *b, int64_t n,
                         const ozc *sigma, int nshift, ozc *x);

/* Dense operator: ctx -> oracle_dense_ctx, A column-major n x n. */
typedef struct { const ozc *a; int64_t n; } oracle_dense_ctx;
void oracle_dense_op(const ozc *sigma, const ozc *x, ozc *Ax, int64_t n, void *ctx);

/* Plane-wave operator (Rydberg units), ctx -> oracle_pw_ctx. */
typedef struct {
  int n1, n2, n3;          /* FFT box; linear index i1 + n1*(i2 + n2*i3)      */
  const double *vloc;     /* real V(r) on the box                            */
  int64_t npw;            /* number of plane waves of this sphere            */
  const int32_t *nl;      /* box index of each plane wave                    */
  const double *kin;      /* |k+q+G|^2 (Ry)                                  */
  int nproj;              /* nonlocal projectors                             */
  const ozc *beta;        /* nproj x npw, projector p contiguous             */
  const double *dcoef;    /* diagonal D (Ry), nproj                          */
  int nocc;               /* valence states for alpha_pv P_v (0 = none)      */
  const ozc *evq;         /* nocc x npw, state v contiguous                  */
  double alpha_pv;        /* projector shift (Ry)                            */
} oracle_pw_ctx;
void oracle_pw_op(const ozc *sigma, const ozc *x, ozc *Ax, int64_t n, void *ctx);

/* In-place 3D FFT of a box (x fastest).  sign=+1: G->r, no scaling (invfft);
 * sign=-1: r->G scaled by 1/(n1 n2 n3) (fwfft). */
void oracle_fft3d(ozc *box, int n1, int n2, int n3, int sign);

/* Counters for the CPU baseline. */
int64_t oracle_op_calls(void);
void oracle_reset_counters(void);

#ifdef __cplusplus
}
#endif
#endif
