/*
 * CPU FP64 oracle for the Sternheimer linear-response solve.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  See sgw_oracle.h for the list of
 * reference routines restated here.  Every routine follows the reference's
 * update order and dot-product flavour (ZDOTU where the reference uses
 * ZDOTU, ZDOTC where it uses ZDOTC) so that iterates can be compared.
 */
#include "sgw_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* complex helpers (explicit formulas, no C99 complex semantics)       */
/* ------------------------------------------------------------------ */
static inline ozc zc(double re, double im) { ozc z = {re, im}; return z; }
static inline ozc zadd(ozc a, ozc b) { return zc(a.re + b.re, a.im + b.im); }
static inline ozc zsub(ozc a, ozc b) { return zc(a.re - b.re, a.im - b.im); }
static inline ozc zmul(ozc a, ozc b) {
  return zc(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline ozc zneg(ozc a) { return zc(-a.re, -a.im); }
static inline ozc zconj(ozc a) { return zc(a.re, -a.im); }
/* Smith's algorithm for complex division. */
static inline ozc zdiv(ozc a, ozc b) {
  if (fabs(b.re) >= fabs(b.im)) {
    double r = b.im / b.re, d = b.re + b.im * r;
    return zc((a.re + a.im * r) / d, (a.im - a.re * r) / d);
  } else {
    double r = b.re / b.im, d = b.im + b.re * r;
    return zc((a.re * r + a.im) / d, (a.im * r - a.re) / d);
  }
}
static const ozc ONE = {1.0, 0.0};

/* y = a*x + y  (ZAXPY) */
static void zaxpy(int64_t n, ozc a, const ozc *x, ozc *y) {
  for (int64_t i = 0; i < n; ++i) y[i] = zadd(y[i], zmul(a, x[i]));
}
/* x = a*x  (ZSCAL) */
static void zscal(int64_t n, ozc a, ozc *x) {
  for (int64_t i = 0; i < n; ++i) x[i] = zmul(a, x[i]);
}

ozc oracle_zdotu(const ozc *x, const ozc *y, int64_t n) {
  ozc s = zc(0, 0);
  for (int64_t i = 0; i < n; ++i) s = zadd(s, zmul(x[i], y[i]));
  return s;
}
ozc oracle_zdotc(const ozc *x, const ozc *y, int64_t n) {
  ozc s = zc(0, 0);
  for (int64_t i = 0; i < n; ++i) s = zadd(s, zmul(zconj(x[i]), y[i]));
  return s;
}

/* ZLANGE('F', 1, n, v, 1) / DNRM2(2n, v): scaled sum of squares over the
 * real and imaginary parts (norm.f90:96-105; linear_solver.f90:373). */
double oracle_norm(const ozc *v, int64_t n) {
  double scale = 0.0, ssq = 1.0;
  for (int64_t i = 0; i < n; ++i) {
    double part[2] = {v[i].re, v[i].im};
    for (int k = 0; k < 2; ++k) {
      if (part[k] != 0.0 || isnan(part[k])) {
        double a = fabs(part[k]);
        if (scale < a || isnan(a)) {
          double r = scale / a;
          ssq = 1.0 + ssq * r * r;
          scale = a;
        } else {
          double r = a / scale;
          ssq += r * r;
        }
      }
    }
  }
  return scale * sqrt(ssq);
}

static int any_nan(const ozc *v, int64_t n) {
  for (int64_t i = 0; i < n; ++i)
    if (isnan(v[i].re) || isnan(v[i].im)) return 1;
  return 0;
}

static int64_t g_op_calls = 0;
int64_t oracle_op_calls(void) { return g_op_calls; }
void oracle_reset_counters(void) { g_op_calls = 0; }

static void apply_op(oracle_op_fn op, void *ctx, ozc sigma, const ozc *x, ozc *ax,
                     int64_t n) {
  ++g_op_calls;
  op(&sigma, x, ax, n, ctx);
}

/* ------------------------------------------------------------------ */
/* gram_schmidt.f90:35-134                                             */
/* basis, vector: n x nbasis column-major; first is 1-based.           */
/* ------------------------------------------------------------------ */
void oracle_gram_schmidt(int first, ozc *basis, int64_t n, int nbasis, ozc *vector) {
  if (n < nbasis) {
    fprintf(stderr, "gram_schmidt: number of vectors larger than dimension\n");
    abort();
  }
  /* 1. orthogonalize the new vectors to the existing basis (:93-106) */
  for (int ib = first; ib <= nbasis; ++ib) {
    ozc *bi = basis + (int64_t)(ib - 1) * n;
    for (int jb = 1; jb <= first - 1; ++jb) {
      const ozc *bj = basis + (int64_t)(jb - 1) * n;
      ozc nrm = oracle_zdotc(bj, bi, n);
      zaxpy(n, zneg(nrm), bj, bi);
      if (vector)
        zaxpy(n, zneg(nrm), vector + (int64_t)(jb - 1) * n, vector + (int64_t)(ib - 1) * n);
    }
  }
  /* 2.-4. normalize and orthogonalize the following ones (:111-132) */
  for (int ib = first; ib <= nbasis; ++ib) {
    ozc *bi = basis + (int64_t)(ib - 1) * n;
    ozc nrm = zdiv(ONE, zc(oracle_norm(bi, n), 0.0));
    zscal(n, nrm, bi);
    if (vector) zscal(n, nrm, vector + (int64_t)(ib - 1) * n);
    for (int jb = ib + 1; jb <= nbasis; ++jb) {
      ozc *bj = basis + (int64_t)(jb - 1) * n;
      nrm = oracle_zdotc(bi, bj, n);
      zaxpy(n, zneg(nrm), bi, bj);
      if (vector)
        zaxpy(n, zneg(nrm), vector + (int64_t)(ib - 1) * n, vector + (int64_t)(jb - 1) * n);
    }
  }
}

/* ------------------------------------------------------------------ */
/* bicgstab.f90 — multishift BiCGStab(l), Frommer Algorithms 2-5       */
/* ------------------------------------------------------------------ */
typedef struct {
  ozc *uu, *rr;   /* n x (l+1) */
  ozc *xx, *tr0;  /* n */
  ozc sigma, rho, rho_old, alpha, alpha_old, beta, omega;
  ozc *gamma, *gamma_p, *gamma_pp; /* l, 1-based storage [1..l] -> [0..l-1] */
} seed_t;

typedef struct {
  ozc *uu;        /* n x (l+1) */
  ozc *xx;
  ozc sigma, inv_phi_old, inv_phi, inv_phi_new, inv_theta, alpha, beta;
  ozc *mu;        /* l(l+1)/2, 1-based index ij -> [ij-1] */
  ozc *gamma, *gamma_p, *gamma_pp;
  int frozen;     /* shift test: done at the end of an earlier iteration, vectors kept */
} shift_t;

#define U(s, i) ((s)->uu + (int64_t)(i) * n)
#define R(s, i) ((s)->rr + (int64_t)(i) * n)

/* init_seed, bicgstab.f90:304-350 */
static void init_seed(int l, const ozc *b, int64_t n, ozc sigma, seed_t *s) {
  s->uu = calloc((size_t)n * (l + 1), sizeof(ozc));
  s->rr = calloc((size_t)n * (l + 1), sizeof(ozc));
  s->xx = calloc((size_t)n, sizeof(ozc));
  s->tr0 = calloc((size_t)n, sizeof(ozc));
  s->gamma = calloc(l, sizeof(ozc));
  s->gamma_p = calloc(l, sizeof(ozc));
  s->gamma_pp = calloc(l, sizeof(ozc));
  memcpy(s->rr, b, sizeof(ozc) * n);
  memcpy(s->tr0, b, sizeof(ozc) * n);
  s->sigma = sigma;
  s->rho_old = ONE;
  s->alpha_old = ONE;
  s->alpha = zc(0, 0);
  s->omega = ONE;
}

/* init_shift, bicgstab.f90:370-490 */
static void init_shift(int l, int64_t n, const ozc *sigma, int nshift, shift_t *sh) {
  int mu_size = l * (l + 1) / 2;
  double *binomial = calloc(mu_size, sizeof(double));
  ozc *sigma_pow = calloc(l, sizeof(ozc));
  for (int jj = 0; jj <= l - 1; ++jj) {
    int offset = jj * (jj + 1) / 2 + 1;
    for (int ii = 0; ii <= jj; ++ii) {
      int ij = offset + ii;
      if (ii == 0) binomial[ij - 1] = 1;
      else binomial[ij - 1] = (binomial[ij - 2] * (jj - ii + 1)) / ii;
    }
  }
  for (int k = 0; k < nshift - 1; ++k) {
    shift_t *a = &sh[k];
    a->uu = calloc((size_t)n * (l + 1), sizeof(ozc));
    a->xx = calloc((size_t)n, sizeof(ozc));
    a->gamma = calloc(l, sizeof(ozc));
    a->gamma_p = calloc(l, sizeof(ozc));
    a->gamma_pp = calloc(l, sizeof(ozc));
    a->mu = calloc(mu_size, sizeof(ozc));
    a->inv_phi_old = ONE;
    a->inv_phi = ONE;
    a->inv_theta = ONE;
    a->sigma = zsub(sigma[k + 1], sigma[0]);
    sigma_pow[0] = ONE;
    for (int ii = 1; ii <= l - 1; ++ii) sigma_pow[ii] = zmul(a->sigma, sigma_pow[ii - 1]);
    for (int jj = 0; jj <= l - 1; ++jj) {
      int offset = jj * (jj + 1) / 2 + 1;
      for (int ii = 0; ii <= jj; ++ii) {
        int ij = offset + ii;
        ozc bin = zc(binomial[ij - 1], 0.0);
        a->mu[ij - 1] = zmul(bin, sigma_pow[jj - ii]);
      }
    }
  }
  free(binomial);
  free(sigma_pow);
}

/* bicg_part, bicgstab.f90:520-702 (Frommer Algorithm 3) */
static void bicg_part(int l, oracle_op_fn op, void *ctx, int64_t n, seed_t *s,
                      shift_t *sh, int nsh) {
  s->rho_old = zmul(zneg(s->omega), s->rho_old);                       /* L3 */
  for (int jj = 0; jj <= l - 1; ++jj) {
    s->rho = oracle_zdotu(R(s, jj), s->tr0, n);                        /* L6 */
    s->beta = zdiv(zmul(s->alpha, s->rho), s->rho_old);
    s->rho_old = s->rho;
    for (int ii = 0; ii <= jj; ++ii) {                                 /* L7-8 */
      zscal(n, zneg(s->beta), U(s, ii));
      zaxpy(n, ONE, R(s, ii), U(s, ii));
    }
    apply_op(op, ctx, s->sigma, U(s, jj), U(s, jj + 1), n);            /* L10 */
    s->alpha = zdiv(s->rho, oracle_zdotu(U(s, jj + 1), s->tr0, n));    /* L11 */
    for (int k = 0; k < nsh; ++k) {
      shift_t *a = &sh[k];
      /* L13: 1/phi_new */
      ozc t = zsub(zdiv(a->inv_phi, a->inv_phi_old), ONE);
      ozc den = zadd(zadd(ONE, zmul(s->alpha, a->sigma)),
                     zmul(zdiv(zmul(s->alpha, s->beta), s->alpha_old), t));
      a->inv_phi_new = zdiv(a->inv_phi, den);
      ozc q = zdiv(a->inv_phi, a->inv_phi_old);
      a->beta = zmul(zmul(q, q), s->beta);
      a->alpha = zmul(zdiv(a->inv_phi_new, a->inv_phi), s->alpha);
      ozc factor = zmul(a->inv_theta, a->inv_phi);
      if (!a->frozen) {
        for (int ii = 0; ii <= jj; ++ii) {                             /* L14-15 */
          zscal(n, zneg(a->beta), U(a, ii));
          zaxpy(n, factor, R(s, ii), U(a, ii));
        }
        zaxpy(n, a->alpha, U(a, 0), a->xx);                            /* L17 */
      }
      a->inv_phi_old = a->inv_phi;                                     /* L18 */
      a->inv_phi = a->inv_phi_new;
      if (!a->frozen) {
        memcpy(U(a, jj + 1), R(s, jj), sizeof(ozc) * n);               /* L19 */
        zscal(n, factor, U(a, jj + 1));
      }
    }
    s->alpha_old = s->alpha;                                           /* L18 */
    for (int ii = 0; ii <= jj; ++ii)                                   /* L21-22 */
      zaxpy(n, zneg(s->alpha), U(s, ii + 1), R(s, ii));
    apply_op(op, ctx, s->sigma, R(s, jj), R(s, jj + 1), n);            /* L24 */
    zaxpy(n, s->alpha, U(s, 0), s->xx);                                /* L25 */
    for (int k = 0; k < nsh; ++k) {                                    /* L26 */
      shift_t *a = &sh[k];
      if (a->frozen) continue;
      ozc factor = zneg(zmul(a->inv_theta, a->inv_phi));
      zaxpy(n, factor, R(s, jj), U(a, jj + 1));
      zscal(n, zdiv(ONE, a->alpha), U(a, jj + 1));
      zaxpy(n, zneg(a->sigma), U(a, jj), U(a, jj + 1));
    }
  }
}

/* horner_scheme, bicgstab.f90:941-999 (Frommer Algorithm 5) */
static void horner_scheme(int l, const ozc *seed_gamma, ozc sigma, ozc *sg, ozc *psi) {
  /* arrays are 1-based in the reference: g(j) -> [j-1] */
  sg[l - 1] = zneg(seed_gamma[l - 1]);
  for (int jj = l - 1; jj >= 1; --jj)
    sg[jj - 1] = zsub(zmul(zneg(sigma), sg[jj]), seed_gamma[jj - 1]);
  *psi = zadd(zmul(zneg(sigma), sg[0]), ONE);
  for (int ii = 1; ii <= l - 1; ++ii)
    for (int jj = l - 1; jj >= ii; --jj)
      sg[jj - 1] = zadd(zmul(zneg(sigma), sg[jj]), sg[jj - 1]);
  for (int jj = 1; jj <= l; ++jj) sg[jj - 1] = zdiv(zneg(sg[jj - 1]), *psi);
}

/* mr_part, bicgstab.f90:708-931 (Frommer Algorithm 4) */
static void mr_part(int l, int64_t n, seed_t *s, shift_t *sh, int nsh) {
  ozc *nu = calloc(l, sizeof(ozc));                       /* nu(1..l) -> [0..l-1] */
  ozc *tau = calloc((l + 1) * (l + 2) / 2, sizeof(ozc));  /* tau(ij) -> [ij-1]   */
#define NU(j) nu[(j) - 1]
#define TAU(ij) tau[(ij) - 1]
#define G(arr, j) (arr)[(j) - 1]
  for (int jj = 1; jj <= l; ++jj) {                                    /* L1 */
    int offset = jj * (jj + 1) / 2 + 1;
    for (int ii = 1; ii <= jj - 1; ++ii) {                             /* L2-3 */
      int ij = offset + ii;
      TAU(ij) = zdiv(oracle_zdotu(R(s, jj), R(s, ii), n), NU(ii));
      zaxpy(n, zneg(TAU(ij)), R(s, ii), R(s, jj));
    }
    NU(jj) = oracle_zdotu(R(s, jj), R(s, jj), n);                      /* L5 */
    G(s->gamma_p, jj) = zdiv(oracle_zdotu(R(s, 0), R(s, jj), n), NU(jj));
  }
  G(s->gamma, l) = G(s->gamma_p, l);                                   /* L7 */
  s->omega = G(s->gamma, l);
  for (int jj = l - 1; jj >= 1; --jj) {                                /* L8-9 */
    G(s->gamma, jj) = G(s->gamma_p, jj);
    for (int ii = jj + 1; ii <= l; ++ii) {
      int ij = ii * (ii + 1) / 2 + jj + 1;
      G(s->gamma, jj) = zsub(G(s->gamma, jj), zmul(TAU(ij), G(s->gamma, ii)));
    }
  }
  for (int jj = 1; jj <= l - 1; ++jj) {                                /* L11-12 */
    G(s->gamma_pp, jj) = G(s->gamma, jj + 1);
    for (int ii = jj + 1; ii <= l - 1; ++ii) {
      int ij = ii * (ii + 1) / 2 + jj + 1;
      G(s->gamma_pp, jj) = zadd(G(s->gamma_pp, jj), zmul(TAU(ij), G(s->gamma, ii + 1)));
    }
  }
  zaxpy(n, G(s->gamma, 1), R(s, 0), s->xx);                            /* L14 */
  zaxpy(n, zneg(G(s->gamma, l)), U(s, l), U(s, 0));
  for (int jj = 1; jj <= l - 1; ++jj) {                                /* L15-17 */
    zaxpy(n, zneg(G(s->gamma, jj)), U(s, jj), U(s, 0));
    zaxpy(n, G(s->gamma_pp, jj), R(s, jj), s->xx);
  }
  for (int k = 0; k < nsh; ++k) {
    shift_t *a = &sh[k];
    ozc psi;
    horner_scheme(l, s->gamma, a->sigma, a->gamma, &psi);              /* L20 */
    ozc inv_xi = zmul(a->inv_theta, a->inv_phi);                       /* L21 */
    a->inv_theta = zdiv(a->inv_theta, psi);
    for (int jj = 1; jj <= l; ++jj) {                                  /* L22-23 */
      G(a->gamma_p, jj) = zc(0, 0);
      for (int ii = jj; ii <= l; ++ii) {
        int ij = (ii - 1) * ii / 2 + jj;
        G(a->gamma_p, jj) = zadd(G(a->gamma_p, jj), zmul(a->mu[ij - 1], G(a->gamma, ii)));
      }
    }
    for (int jj = 1; jj <= l - 1; ++jj) {                              /* L25-26 */
      G(a->gamma_pp, jj) = G(a->gamma_p, jj + 1);
      for (int ii = jj + 1; ii <= l - 1; ++ii) {
        int ij = ii * (ii + 1) / 2 + jj + 1;
        G(a->gamma_pp, jj) = zadd(G(a->gamma_pp, jj), zmul(TAU(ij), G(a->gamma_p, ii + 1)));
      }
    }
    if (a->frozen) continue;
    ozc factor = zmul(G(a->gamma_p, 1), inv_xi);                       /* L28 */
    zaxpy(n, factor, R(s, 0), a->xx);
    zaxpy(n, zneg(G(s->gamma, l)), U(a, l), U(a, 0));
    for (int jj = 1; jj <= l - 1; ++jj) {                              /* L29-31 */
      zaxpy(n, zneg(G(s->gamma, jj)), U(a, jj), U(a, 0));
      factor = zmul(G(a->gamma_pp, jj), inv_xi);
      zaxpy(n, factor, R(s, jj), a->xx);
    }
    zscal(n, zdiv(ONE, psi), U(a, 0));                                 /* L34 */
  }
  for (int jj = 1; jj <= l; ++jj)                                      /* L32 (delayed) */
    zaxpy(n, zneg(G(s->gamma_p, jj)), R(s, jj), R(s, 0));
#undef NU
#undef TAU
#undef G
  free(tau);
  free(nu);
}

/* Convergence test of bicgstab.f90:296-299 on the seed residual; with
 * shift_test (an extension, no counterpart in the reference) the largest
 * shifted residual estimate |1/(theta phi)| ||r_seed|| must pass as well.
 * At the end of an iteration (end_of_iter) a shifted system whose estimate
 * is below half the threshold is done: its vectors are kept from then on and
 * it leaves the test (the half leaves room for the gap between the estimate
 * and the true residual). */
static int converged(int64_t n, const seed_t *s, shift_t *sh, int nsh, double threshold,
                     int shift_test, int end_of_iter) {
  double nr = oracle_norm(s->rr, n), zmax = 1.0;
  if (shift_test)
    for (int k = 0; k < nsh; ++k) {
      if (sh[k].frozen) continue;
      ozc z = zmul(sh[k].inv_theta, sh[k].inv_phi);
      double a = sqrt(z.re * z.re + z.im * z.im);
      if (end_of_iter && a * nr < 0.5 * threshold) sh[k].frozen = 1;
      else if (a > zmax) zmax = a;
    }
  return nr * zmax < threshold;
}

/* bicgstab, bicgstab.f90:154-275 (Frommer Algorithm 2) */
int oracle_bicgstab(int lmax, double threshold, int max_iter, oracle_op_fn op,
                    void *ctx, const ozc *b, int64_t n, const ozc *sigma,
                    int nshift, ozc *x, int *iters_out) {
  return oracle_bicgstab_ex(lmax, threshold, max_iter, op, ctx, b, n, sigma, nshift, x,
                            iters_out, 0);
}

int oracle_bicgstab_ex(int lmax, double threshold, int max_iter, oracle_op_fn op,
                       void *ctx, const ozc *b, int64_t n, const ozc *sigma,
                       int nshift, ozc *x, int *iters_out, int shift_test) {
  int l = lmax, ierr = 0, iter;
  seed_t s;
  int nsh = nshift - 1;
  shift_t *sh = calloc(nsh > 0 ? nsh : 1, sizeof(shift_t));
  init_seed(l, b, n, sigma[0], &s);
  init_shift(l, n, sigma, nshift, sh);
  for (iter = 1; iter <= max_iter; ++iter) {
    bicg_part(l, op, ctx, n, &s, sh, nsh);
    if (converged(n, &s, sh, nsh, threshold, shift_test, 0)) break;
    mr_part(l, n, &s, sh, nsh);
    if (converged(n, &s, sh, nsh, threshold, shift_test, 1)) break;
  }
  if (iter > max_iter) {
    printf("WARNING: BiCGstab algorithm did not converge in %d iterations.\n", max_iter);
    ierr = 1;
  }
  memcpy(x, s.xx, sizeof(ozc) * n);
  for (int k = 0; k < nsh; ++k) memcpy(x + (int64_t)(k + 1) * n, sh[k].xx, sizeof(ozc) * n);
  if (any_nan(x, n * (int64_t)nshift)) {
    printf(" WARNING: BiCGstab algorithm resulted in one or more NaN\n");
    ierr = 2;
  }
  if (iters_out) *iters_out = iter;
  free(s.uu); free(s.rr); free(s.xx); free(s.tr0);
  free(s.gamma); free(s.gamma_p); free(s.gamma_pp);
  for (int k = 0; k < nsh; ++k) {
    free(sh[k].uu); free(sh[k].xx); free(sh[k].mu);
    free(sh[k].gamma); free(sh[k].gamma_p); free(sh[k].gamma_pp);
  }
  free(sh);
  return ierr;
}
#undef U
#undef R

/* ------------------------------------------------------------------ */
/* linear_solver.f90:82-503 — subspace solver reused across shifts      */
/* ------------------------------------------------------------------ */
int oracle_linear_solver(int max_iter, double threshold, oracle_op_fn op, void *ctx,
                         const ozc *b, int64_t n, const ozc *sigma, int nshift,
                         ozc *x, int *basis_out) {
  int ierr = 0;
  double abs_threshold = threshold * oracle_norm(b, n);                /* :200-215 */
  int cap = 16, nb = 0;
  ozc *vv = malloc(sizeof(ozc) * n * cap), *ww = malloc(sizeof(ozc) * n * cap);
  ozc sub_sigma = zc(0, 0);                                            /* :241-244 */
  ozc *res = malloc(sizeof(ozc) * n), *newv = malloc(sizeof(ozc) * n);
  for (int ishift = 0; ishift < nshift; ++ishift) {
    /* orthogonal_subspace :272-300 */
    ozc diff = zsub(sigma[ishift], sub_sigma);
    zaxpy(n * (int64_t)nb, diff, vv, ww);
    oracle_gram_schmidt(1, ww, n, nb, vv);
    sub_sigma = sigma[ishift];
    int conv = 0;
    for (int iter = 1; iter <= max_iter; ++iter) {
      /* residual :312-383 */
      memcpy(res, b, sizeof(ozc) * n);
      for (int ib = 0; ib < nb; ++ib) {
        const ozc *w = ww + (int64_t)ib * n;
        ozc overlap = oracle_zdotc(w, b, n);
        for (int64_t i = 0; i < n; ++i) res[i] = zsub(res[i], zmul(overlap, w[i]));
      }
      double nrm = oracle_norm(res, n);
      conv = nrm < abs_threshold;
      if (!conv && n == nb) {
        printf("WARNING: SternheimerGW linear solver right-hand side cannot be mapped by complete basis\n");
        ierr = 3;
      } else {
        ierr = 0;
      }
      if (conv || ierr != 0) break;
      apply_op(op, ctx, sigma[ishift], res, newv, n);                  /* :168 */
      /* expand_subspace :391-440 */
      if (nb + 1 > cap) {
        cap *= 2;
        vv = realloc(vv, sizeof(ozc) * n * cap);
        ww = realloc(ww, sizeof(ozc) * n * cap);
      }
      memcpy(ww + (int64_t)nb * n, newv, sizeof(ozc) * n);
      memcpy(vv + (int64_t)nb * n, res, sizeof(ozc) * n);
      ++nb;
      oracle_gram_schmidt(nb, ww, n, nb, vv);
    }
    if (!conv) {                                                       /* :176-180 */
      printf("WARNING: SternheimerGW linear solver did not converge\n");
      ierr = 1;
      break;
    }
    /* obtain_result :453-503 */
    ozc *xs = x + (int64_t)ishift * n;
    for (int64_t i = 0; i < n; ++i) xs[i] = zc(0, 0);
    for (int ib = 0; ib < nb; ++ib) {
      ozc overlap = oracle_zdotc(ww + (int64_t)ib * n, b, n);
      zaxpy(n, overlap, vv + (int64_t)ib * n, xs);
    }
    if (any_nan(xs, n)) {                                              /* :186-190 */
      printf("WARNING: SternheimerGW linear solver produced at least one NaN\n");
      ierr = 2;
      break;
    }
  }
  if (basis_out) *basis_out = nb;
  free(vv); free(ww); free(res); free(newv);
  return ierr;
}

/* ------------------------------------------------------------------ */
/* select_solver.f90:67-202                                            */
/* ------------------------------------------------------------------ */
int oracle_select_solver(const int *priority, int npriority, int max_iter,
                         double threshold, int bicg_lmax, oracle_op_fn op, void *ctx,
                         const ozc *b, int64_t n, const ozc *sigma, int nshift,
                         ozc *x) {
  if (npriority <= 0) {
    fprintf(stderr, "select_solver: priority of the solvers not specified\n");
    abort();
  }
  int ierr = 1;
  for (int is = 0; is < npriority; ++is) {
    if (is > 0) printf("First choice of solver did not converge, try a different one\n");
    switch (priority[is]) {
      case 1:
        ierr = oracle_bicgstab(bicg_lmax, threshold, max_iter, op, ctx, b, n, sigma,
                               nshift, x, NULL);
        break;
      case 2:
        /* select_solver.f90:144-146 repeats the full multishift solve once
         * per shift; every repetition yields the same x and ierr. */
        for (int k = 0; k < nshift; ++k)
          ierr = oracle_bicgstab(bicg_lmax, threshold, max_iter, op, ctx, b, n,
                                 sigma, nshift, x, NULL);
        break;
      case 3:
        ierr = oracle_linear_solver(max_iter, threshold, op, ctx, b, n, sigma, nshift,
                                    x, NULL);
        break;
      default:
        break;
    }
    if (ierr == 0) break;
  }
  return ierr;
}

/* ------------------------------------------------------------------ */
/* operators                                                           */
/* ------------------------------------------------------------------ */
/* Ax = MATMUL(A, x) + sigma*x   (linear_solver.pf:100-113) */
void oracle_dense_op(const ozc *sigma, const ozc *x, ozc *ax, int64_t n, void *ctx) {
  const oracle_dense_ctx *c = (const oracle_dense_ctx *)ctx;
  for (int64_t i = 0; i < n; ++i) ax[i] = zc(0, 0);
  for (int64_t j = 0; j < n; ++j) {
    const ozc *col = c->a + j * c->n;
    ozc xj = x[j];
    for (int64_t i = 0; i < n; ++i) ax[i] = zadd(ax[i], zmul(col[i], xj));
  }
  for (int64_t i = 0; i < n; ++i) ax[i] = zadd(ax[i], zmul(*sigma, x[i]));
}

/* ---- mixed-radix FFT (Cooley-Tukey, decimation in time, recursive) ----
 * Twiddles come from a table of the N-th roots of unity of the top-level
 * length: W_n^e = W_N^(e*N/n). */
typedef struct { int n; ozc *w; } fft_plan;

static ozc twid(const fft_plan *pl, long e, int n, int sign) {
  long idx = (e * (pl->n / n)) % pl->n;
  ozc w = pl->w[idx];
  return sign > 0 ? w : zconj(w);
}

static void fft1d_rec(ozc *out, const ozc *in, int n, int64_t stride, int sign,
                      ozc *tmp, const fft_plan *pl) {
  if (n == 1) { out[0] = in[0]; return; }
  int p = 0;
  if (n % 4 == 0) p = 4;
  else if (n % 2 == 0) p = 2;
  else if (n % 3 == 0) p = 3;
  else if (n % 5 == 0) p = 5;
  if (p == 0) { /* direct DFT for other prime lengths */
    for (int k = 0; k < n; ++k) {
      ozc s = zc(0, 0);
      for (int j = 0; j < n; ++j)
        s = zadd(s, zmul(in[(int64_t)j * stride], twid(pl, ((long)j * k) % n, n, sign)));
      out[k] = s;
    }
    return;
  }
  int m = n / p;
  for (int r = 0; r < p; ++r)
    fft1d_rec(out + r * m, in + (int64_t)r * stride, m, stride * p, sign, tmp, pl);
  for (int k = 0; k < m; ++k) {
    ozc v[5];
    for (int r = 0; r < p; ++r) v[r] = zmul(out[r * m + k], twid(pl, (long)r * k, n, sign));
    for (int q = 0; q < p; ++q) {
      ozc s = v[0];
      for (int r = 1; r < p; ++r) s = zadd(s, zmul(v[r], twid(pl, (long)((r * q) % p), p, sign)));
      tmp[q * m + k] = s;
    }
  }
  memcpy(out, tmp, sizeof(ozc) * n);
}

static void plan_init(fft_plan *pl, int n) {
  pl->n = n;
  pl->w = malloc(sizeof(ozc) * n);
  for (int k = 0; k < n; ++k) {
    double ang = 2.0 * M_PI * (double)k / n;
    pl->w[k] = zc(cos(ang), sin(ang));
  }
}

static void fft_line(ozc *line, int n, int sign, ozc *buf, ozc *tmp, const fft_plan *pl) {
  memcpy(buf, line, sizeof(ozc) * n);
  fft1d_rec(line, buf, n, 1, sign, tmp, pl);
}

void oracle_fft3d(ozc *box, int n1, int n2, int n3, int sign) {
  int nmax = n1 > n2 ? n1 : n2;
  nmax = nmax > n3 ? nmax : n3;
  ozc *line = malloc(sizeof(ozc) * nmax), *buf = malloc(sizeof(ozc) * nmax),
      *tmp = malloc(sizeof(ozc) * nmax);
  fft_plan p1, p2, p3;
  plan_init(&p1, n1); plan_init(&p2, n2); plan_init(&p3, n3);
  /* along x */
  for (int64_t o = 0; o < (int64_t)n2 * n3; ++o) fft_line(box + o * n1, n1, sign, buf, tmp, &p1);
  /* along y */
  for (int i3 = 0; i3 < n3; ++i3)
    for (int i1 = 0; i1 < n1; ++i1) {
      ozc *base = box + (int64_t)i3 * n1 * n2 + i1;
      for (int i2 = 0; i2 < n2; ++i2) line[i2] = base[(int64_t)i2 * n1];
      fft_line(line, n2, sign, buf, tmp, &p2);
      for (int i2 = 0; i2 < n2; ++i2) base[(int64_t)i2 * n1] = line[i2];
    }
  /* along z */
  for (int64_t o = 0; o < (int64_t)n1 * n2; ++o) {
    ozc *base = box + o;
    for (int i3 = 0; i3 < n3; ++i3) line[i3] = base[(int64_t)i3 * n1 * n2];
    fft_line(line, n3, sign, buf, tmp, &p3);
    for (int i3 = 0; i3 < n3; ++i3) base[(int64_t)i3 * n1 * n2] = line[i3];
  }
  if (sign < 0) {
    double inv = 1.0 / ((double)n1 * n2 * n3);
    for (int64_t i = 0; i < (int64_t)n1 * n2 * n3; ++i) { box[i].re *= inv; box[i].im *= inv; }
  }
  free(p1.w); free(p2.w); free(p3.w);
  free(line); free(buf); free(tmp);
}

/* (H + sigma + alpha_pv P_v) x with H = |k+q+G|^2 + V_loc + sum_p |b_p> D_p <b_p|
 * in the order h_psi -> +sigma*S psi -> +alpha_pv P_v S psi (linear_op.f90:116-138),
 * S = 1 (norm-conserving). */
void oracle_pw_op(const ozc *sigma, const ozc *x, ozc *ax, int64_t n, void *ctx) {
  const oracle_pw_ctx *c = (const oracle_pw_ctx *)ctx;
  int64_t nbox = (int64_t)c->n1 * c->n2 * c->n3;
  ozc *box = calloc((size_t)nbox, sizeof(ozc));
  /* kinetic */
  for (int64_t g = 0; g < n; ++g) ax[g] = zc(c->kin[g] * x[g].re, c->kin[g] * x[g].im);
  /* local: scatter, invfft, x V(r), fwfft, gather (QE vloc_psi_k) */
  for (int64_t g = 0; g < n; ++g) box[c->nl[g]] = x[g];
  oracle_fft3d(box, c->n1, c->n2, c->n3, +1);
  for (int64_t i = 0; i < nbox; ++i) { box[i].re *= c->vloc[i]; box[i].im *= c->vloc[i]; }
  oracle_fft3d(box, c->n1, c->n2, c->n3, -1);
  for (int64_t g = 0; g < n; ++g) ax[g] = zadd(ax[g], box[c->nl[g]]);
  /* nonlocal: becp = beta^H x ; ax += beta (D becp)  (QE calbec + add_vuspsi) */
  for (int p = 0; p < c->nproj; ++p) {
    const ozc *bp = c->beta + (int64_t)p * n;
    ozc becp = oracle_zdotc(bp, x, n);
    ozc db = zc(c->dcoef[p] * becp.re, c->dcoef[p] * becp.im);
    zaxpy(n, db, bp, ax);
  }
  /* shift: ax += sigma * x  (linear_op.f90:122-126) */
  zaxpy(n, *sigma, x, ax);
  /* valence projector: ax += alpha_pv evq (evq^H x)  (linear_op.f90:131-138, 195-200) */
  if (fabs(c->alpha_pv) > 1e-14 && c->nocc > 0) {
    ozc *proj = malloc(sizeof(ozc) * c->nocc);
    for (int v = 0; v < c->nocc; ++v) {
      ozc pr = oracle_zdotc(c->evq + (int64_t)v * n, x, n);
      proj[v] = zc(c->alpha_pv * pr.re, c->alpha_pv * pr.im);
    }
    for (int64_t g = 0; g < n; ++g) {
      ozc acc = zc(0, 0);
      for (int v = 0; v < c->nocc; ++v) acc = zadd(acc, zmul(c->evq[(int64_t)v * n + g], proj[v]));
      ax[g] = zadd(ax[g], acc);
    }
    free(proj);
  }
  free(box);
}
