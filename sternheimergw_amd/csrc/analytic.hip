// The reference form of the correlation self-energy on the device:
//  * analytic continuation of W (algo/analytic/src/analytic.f90:50-348) with
//    the Godby-Needs plasmon-pole model (godby_needs.f90) and the N-point
//    Pade approximant of Vidberg and Serene (pade.f90:41-147), one thread per
//    (G, G') element, and the robust Pade approximant (pade_robust.f90), one
//    group of 32 lanes per element;
//  * sigma_prod (phys/corr/src/sigma.f90:417-523): W(G, G') -> W(r, r') by
//    invfft6, the product G(r, r') W(r, r'), fwfft6 back, times alpha, and
//    optionally added into Sigma (sigma.f90:741).
#include <algorithm>
#include <cmath>
#include <vector>

#include "aaa.h"
#include "boxfft.h"
#include "fft6.h"
#include "ktimer.h"
#include "sgw_internal.h"

namespace sgw {
namespace {

constexpr int kEB = 256;
constexpr int kMaxPade = 64;  // frequencies of the symmetrised mesh
constexpr double kEps8 = 1e-8, kEps24 = 1e-24;

unsigned blocks(int64_t count) { return (unsigned)((count + kEB - 1) / kEB); }

__device__ inline double cabs2(double2 a) { return a.x * a.x + a.y * a.y; }

// principal square root
__device__ inline double2 csqrt_p(double2 z) {
  const double r = sqrt(cabs2(z));
  const double re = sqrt(0.5 * (r + z.x));
  const double im = copysign(sqrt(0.5 * fmax(r - z.x, 0.0)), z.y);
  return mk(re, im);
}

__device__ inline double2 protect(double2 x) {
  return sqrt(cabs2(x)) > kEps24 ? x : mk(kEps24, 0.0);
}

// godby_needs_coeffs (godby_needs.f90): W0 = W(0), W1 = W(i omega_p) of every
// element become the pole strength and frequency, in place
__global__ void k_gn_coeffs(int64_t ne, double omega_p, double2 *__restrict__ w0,
                            double2 *__restrict__ w1) {
  const int64_t e = (int64_t)blockIdx.x * kEB + threadIdx.x;
  if (e >= ne) return;
  const double2 a = w0[e], b = w1[e];
  double2 work = csub(a, b);
  bool zero = fabs(work.x) < kEps8;
  if (!zero) {
    work = cdiv(b, work);
    zero = work.x < kEps8;
  }
  if (!zero) {
    const double2 c2 = csqrt_p(work);
    const double2 p = mk(c2.x * omega_p, c2.y * omega_p);
    const double2 h = cmul(mk(0.5 * a.x, 0.5 * a.y), p);
    w1[e] = p;
    w0[e] = h;
  } else {
    w0[e] = mk(0.0, 0.0);
    w1[e] = mk(0.0, 0.0);
  }
}

struct Mesh {
  double2 z[kMaxPade];
  int src[kMaxPade];  // input slice of each symmetrised point
};

// pade_coeff (Vidberg-Serene recursion) per element; u_i = w[src[i]][e],
// coefficients a_p written to slice p.  The rows g(p, .) of the recursion
// are kept one at a time in g[].
__global__ void k_pade_coeffs(int64_t ne, int N, Mesh m, double2 *__restrict__ w) {
  const int64_t e = (int64_t)blockIdx.x * kEB + threadIdx.x;
  if (e >= ne) return;
  double2 g[kMaxPade];
  for (int i = 0; i < N; ++i) g[i] = protect(w[(int64_t)m.src[i] * ne + e]);
  w[e] = g[0];
  for (int p = 1; p < N; ++p) {
    const double2 gpp = g[p - 1];
    for (int i = p; i < N; ++i) {
      const double2 t1 = cdiv(gpp, g[i]);
      const double2 t2 = cdiv(g[i], g[i]);
      g[i] = protect(cdiv(csub(t1, t2), csub(m.z[i], m.z[p - 1])));
    }
    w[(int64_t)p * ne + e] = g[p];
  }
}

// ---------------------------------------------------------------------------
// Robust Pade approximation (pade_robust.f90:164-535, the SVD algorithm of
// P. Gonnet, S. Guettel and L. N. Trefethen, SIAM Rev. 55 (2013) 101-117)
// of every element from its values on a circle of N points.  One group of 32
// lanes (half a wavefront) per element; lane l holds row l of the group's
// matrix, which lives in LDS column by column (column stride 32).  Numerical
// rank from one-sided Jacobi singular values (the reference calls ZGESDD);
// null vectors from Householder QR in the form of ZGEQRF + ZUNGQR (the
// reference's qr wrapper, lapack.f90:150-225).

constexpr int kPrLanes = 32;                            // lanes per element
constexpr int kPrGroups = 4;                            // most elements per block
constexpr int kPrLdsMax = 65536;                        // LDS bytes per block
constexpr int kPrSweeps = 40;                           // Jacobi sweep limit
constexpr double kPrTol = 1e-14;                        // eps14, the default tolerance
constexpr double kSqrtEps = 1.4901161193847656e-08;     // sqrt(epsilon(1.0_dp))
constexpr double kDblEps = 2.220446049250313e-16;

// LDS per element: f then a (64), Taylor coefficients (64), b (32), tau (32),
// the matrix (32 x nmax)
__host__ __device__ inline int pr_group_lds(int nmax) { return 192 + kPrLanes * nmax; }

// sums and extrema over the 32 lanes of a group; the butterfly gives every
// lane the same bits (each stage adds the same two operands)
__device__ inline double gsum(double v) {
#pragma unroll
  for (int o = kPrLanes / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kPrLanes);
  return v;
}
__device__ inline double gmax(double v) {
#pragma unroll
  for (int o = kPrLanes / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kPrLanes));
  return v;
}
__device__ inline int gmaxi(int v) {
#pragma unroll
  for (int o = kPrLanes / 2; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, kPrLanes));
  return v;
}
__device__ inline int gmini(int v) {
#pragma unroll
  for (int o = kPrLanes / 2; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, kPrLanes));
  return v;
}
__device__ inline double cabsd(double2 a) { return hypot(a.x, a.y); }

// row l of (C D)^T: A(l, i) = C(i, l) d_l = c(m+1+i-l) d_l for i < n, l <= n,
// C = rows m+1..m+n of the Toeplitz matrix of c (pade_robust.f90:347-352)
__device__ void pr_build(double2 *A, const double2 *cc, int m, int n, int l, double d) {
  __syncwarp();
  for (int i = 0; i < n; ++i) {
    const int k = m + 1 + i - l;
    const double2 v = (l <= n && k >= 0) ? cc[k] : mk(0.0, 0.0);
    A[i * kPrLanes + l] = mk(v.x * d, v.y * d);
  }
  __syncwarp();
}

// number of singular values of A (rows x cols, rows <= 32) above tol:
// one-sided (Hestenes) Jacobi orthogonalises the columns, whose norms are
// then the singular values.  Each lane touches only its own row of A.
__device__ int pr_rank(double2 *A, int rows, int cols, int l, double tol) {
  const double ctol = sqrt((double)rows) * kDblEps;
  for (int sweep = 0; sweep < kPrSweeps; ++sweep) {
    int rot = 0;
    for (int p = 0; p + 1 < cols; ++p)
      for (int q = p + 1; q < cols; ++q) {
        const double2 ap = l < rows ? A[p * kPrLanes + l] : mk(0.0, 0.0);
        const double2 aq = l < rows ? A[q * kPrLanes + l] : mk(0.0, 0.0);
        const double al = gsum(cabs2(ap)), be = gsum(cabs2(aq));
        double2 gm = mk(0.0, 0.0);
        cmacc(gm, ap, aq);                                   // conj(a_p) a_q
        gm = mk(gsum(gm.x), gsum(gm.y));
        const double g = hypot(gm.x, gm.y);
        if (!(g > ctol * sqrt(al) * sqrt(be))) continue;    // uniform over the group
        ++rot;
        // a_q~ = a_q conj(g)/|g| makes a_p^H a_q~ = |g| real; then the real
        // rotation with t^2 + 2 zeta t - 1 = 0, |t| <= 1
        const double zeta = (be - al) / (2.0 * g);
        const double az = fabs(zeta);
        const double t = az > 1e150 ? 0.5 / zeta : copysign(1.0, zeta) / (az + sqrt(1.0 + zeta * zeta));
        const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
        const double2 ph = mk(gm.x / g, -gm.y / g);
        const double2 aqt = cmul(aq, ph);
        if (l < rows) {
          A[p * kPrLanes + l] = mk(cs * ap.x - sn * aqt.x, cs * ap.y - sn * aqt.y);
          A[q * kPrLanes + l] = mk(sn * ap.x + cs * aqt.x, sn * ap.y + cs * aqt.y);
        }
      }
    if (rot == 0) break;
  }
  int rho = 0;
  for (int i = 0; i < cols; ++i) {
    const double2 a = l < rows ? A[i * kPrLanes + l] : mk(0.0, 0.0);
    if (sqrt(gsum(cabs2(a))) > tol) ++rho;
  }
  return rho;
}

// Householder QR of A (rows x cols, rows = cols + 1 <= 31) in place, in the
// conventions of ZGEQR2/ZLARFG: H_k = I - tau_k v_k v_k^H with v_k(k) = 1,
// beta = -sign(Re alpha) ||(alpha, x)||; R on and above the diagonal, v_k
// below it, tau_k in tau[k]
__device__ void pr_qr(double2 *A, double2 *tau, int rows, int cols, int l) {
  for (int k = 0; k < cols; ++k) {
    __syncwarp();
    const double2 alpha = A[k * kPrLanes + k];
    const bool below = l > k && l < rows;
    const double2 x = below ? A[k * kPrLanes + l] : mk(0.0, 0.0);
    const double xn2 = gsum(cabs2(x));
    double2 tk = mk(0.0, 0.0);
    if (!(xn2 == 0.0 && alpha.y == 0.0)) {
      const double beta = -copysign(sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2), alpha.x);
      tk = mk((beta - alpha.x) / beta, -alpha.y / beta);
      const double2 sc = cdiv(mk(1.0, 0.0), mk(alpha.x - beta, alpha.y));
      const double2 v = l == k ? mk(1.0, 0.0) : (below ? cmul(x, sc) : mk(0.0, 0.0));
      if (below) A[k * kPrLanes + l] = v;
      if (l == k) A[k * kPrLanes + k] = mk(beta, 0.0);
      // H_k^H = I - conj(tau_k) v v^H on the columns right of k
      const double2 ct = cconj(tk);
      for (int j = k + 1; j < cols; ++j) {
        const bool act = l >= k && l < rows;
        const double2 a = act ? A[j * kPrLanes + l] : mk(0.0, 0.0);
        double2 s = mk(0.0, 0.0);
        cmacc(s, v, a);
        s = mk(gsum(s.x), gsum(s.y));
        if (act) A[j * kPrLanes + l] = csub(a, cmul(ct, cmul(v, s)));
      }
    }
    if (l == 0) tau[k] = tk;
  }
  __syncwarp();
}

// row l of Q e_{rows-1}, Q = H_0 H_1 ... H_{cols-1} (the last column of the
// full Q of ZUNGQR)
__device__ double2 pr_qlast(const double2 *A, const double2 *tau, int rows, int cols, int l) {
  double2 y = l == rows - 1 ? mk(1.0, 0.0) : mk(0.0, 0.0);
  for (int k = cols - 1; k >= 0; --k) {
    const double2 v = l == k ? mk(1.0, 0.0)
                             : ((l > k && l < rows) ? A[k * kPrLanes + l] : mk(0.0, 0.0));
    double2 s = mk(0.0, 0.0);
    cmacc(s, v, y);
    s = mk(gsum(s.x), gsum(s.y));
    y = csub(y, cmul(tau[k], cmul(v, s)));
  }
  return y;
}

// pade_coeff_robust (pade_robust.f90:89-160) per element: w[f * ne + e] holds
// the values on the circle z_f = R e^{2 pi i f / N} and receives
// [m, n, a_0..a_m, b_0..b_n, 0, ...] (the reference leaves whatever was there
// beyond the coefficients; pade_eval_robust never reads it)
__global__ __launch_bounds__(kPrLanes *kPrGroups) void k_pade_robust_coeffs(
    int64_t ne, int N, double radius, int nmax, double2 *__restrict__ w) {
  extern __shared__ double2 pr_lds[];
  __shared__ double2 tw[kMaxPade];
  for (int t = threadIdx.x; t < N; t += blockDim.x) {
    double s, c;
    sincospi(-2.0 * t / N, &s, &c);
    tw[t] = mk(c, s);
  }
  __syncthreads();
  const int grp = threadIdx.x / kPrLanes, l = threadIdx.x % kPrLanes;
  const int64_t e = (int64_t)blockIdx.x * (blockDim.x / kPrLanes) + grp;
  if (e >= ne) return;
  double2 *fa = pr_lds + (size_t)grp * pr_group_lds(nmax);   // f, later a
  double2 *cc = fa + kMaxPade;
  double2 *bv = cc + kMaxPade;
  double2 *tau = bv + kPrLanes;
  double2 *A = tau + kPrLanes;
  for (int j = l; j < N; j += kPrLanes) fa[j] = w[(int64_t)j * ne + e];
  __syncwarp();

  int m = N / 2 - 2, n = m;            // the (n, n) start, pade_coeff_robust :146-147
  const int nd = m + n + 1;
  // pade_derivative (:445-535): c_k = FFT(f)_k / N / R^k, k < nd (<= N)
  double2 ck[2];
  double nrm = 0.0;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int k = l + h * kPrLanes;
    double2 acc = mk(0.0, 0.0);
    if (k < nd) {
      for (int j = 0; j < N; ++j) cmac(acc, fa[j], tw[(j * k) % N]);
      acc = mk(acc.x * (1.0 / N), acc.y * (1.0 / N));
      if (fabs(radius - 1.0) > kPrTol) {
        double r = 1.0;
        for (int i = 0; i < k; ++i) r = r / radius;
        acc = mk(acc.x * r, acc.y * r);
      }
    }
    ck[h] = acc;
    nrm += cabs2(acc);
  }
  {
    const double tol = kPrTol * sqrt(gsum(nrm));   // near-zero coefficients dropped
    double im = 0.0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (cabsd(ck[h]) < tol) ck[h] = mk(0.0, 0.0);
      im = fmax(im, fabs(ck[h].y));
    }
    if (gmax(im) < tol) {                           // real functions made real
      ck[0].y = 0.0;
      ck[1].y = 0.0;
    }
  }
  cc[l] = ck[0];
  cc[l + kPrLanes] = ck[1];
  const double abs_tol = kPrTol * sqrt(gsum(cabs2(ck[0]) + cabs2(ck[1])));   // :276
  double cmax = 0.0, cmax_m = 0.0;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int k = l + h * kPrLanes;
    cmax = fmax(cmax, cabsd(ck[h]));
    if (k <= m) cmax_m = fmax(cmax_m, cabsd(ck[h]));
  }
  cmax = gmax(cmax);
  cmax_m = gmax(cmax_m);
  __syncwarp();

  double2 a = mk(0.0, 0.0), b = mk(0.0, 0.0);      // row l of a and of b
  const bool zero = cmax_m <= kPrTol * cmax;
  if (zero) {                                       // zero function, returned as is (:287-306)
    m = 0;
    n = 0;
    b = l == 0 ? mk(1.0, 0.0) : b;
  } else {
    // diagonal hopping (:318-363): while rank(C) < n, shrink m and n
    while (n > 0) {
      pr_build(A, cc, m, n, l, 1.0);
      const int rho = pr_rank(A, n + 1, n, l, abs_tol);
      if (rho == n) break;
      m -= n - rho;
      n = rho;
    }
    if (n == 0) {                                   // (:329-343)
      a = l <= m ? cc[l] : a;
      b = l == 0 ? mk(1.0, 0.0) : b;
    } else {
      // null vector of C (a unit vector, up to phase the one of ZGESDD) gives
      // the weights D = |v| + sqrt(eps); then b = D q_{n+1} of the QR of
      // (C D)^T, normalised (:366-386)
      pr_build(A, cc, m, n, l, 1.0);
      pr_qr(A, tau, n + 1, n, l);
      double2 y = pr_qlast(A, tau, n + 1, n, l);
      const double d = l <= n ? cabsd(y) + kSqrtEps : 0.0;
      pr_build(A, cc, m, n, l, d);
      pr_qr(A, tau, n + 1, n, l);
      y = pr_qlast(A, tau, n + 1, n, l);
      b = mk(d * y.x, d * y.y);
      const double bn = sqrt(gsum(cabs2(b)));
      b = mk(b.x / bn, b.y / bn);
      bv[l] = b;
      __syncwarp();
      // a = Z(0:m, 0:n) b, Z(i, j) = c(i - j) for i >= j (:389)
      if (l <= m)
        for (int j = 0; j <= min(l, n); ++j) cmac(a, cc[l - j], bv[j]);
      // leading zeros of b (and a) dropped (:392-408)
      int lam = gmini((l <= n && cabsd(b) > kPrTol) ? l : kPrLanes);
      if (lam == kPrLanes) lam = 0;
      if (lam > 0) {
        fa[l] = a;
        __syncwarp();
        a = l + lam <= m ? fa[l + lam] : mk(0.0, 0.0);
        b = l + lam <= n ? bv[l + lam] : mk(0.0, 0.0);
        m -= lam;
        n -= lam;
        __syncwarp();
      }
      // trailing zeros of b dropped (:411-420)
      const int last = gmaxi((l <= n && cabsd(b) > kPrTol) ? l : -1);
      if (last != n) {
        n = last;
        if (l > n) b = mk(0.0, 0.0);
      }
    }
  }
  // trailing zeros of a dropped (:425-434), then a and b over b_0 (:437-438)
  const int last_a = zero ? m : gmaxi((l <= m && cabsd(a) > abs_tol) ? l : -1);
  if (last_a != m) {
    m = last_a;
    if (l > m) a = mk(0.0, 0.0);
  }
  bv[l] = b;
  __syncwarp();
  const double2 b0 = bv[0];
  a = cdiv(a, b0);
  b = cdiv(b, b0);
  __syncwarp();
  fa[l] = a;
  bv[l] = b;
  __syncwarp();
  // slices: 0 -> m, 1 -> n, 2..2+m -> a, 3+m..3+m+n -> b, the rest 0
  for (int s = l; s < N; s += kPrLanes) {
    double2 v = mk(0.0, 0.0);
    if (s == 0) v = mk((double)m, 0.0);
    else if (s == 1) v = mk((double)n, 0.0);
    else if (s - 2 <= m) v = fa[s - 2];
    else if (s - 3 - m <= n) v = bv[s - 3 - m];
    w[(int64_t)s * ne + e] = v;
  }
}

// analytic_eval (analytic.f90:276-345): out(r, c) for r < num_g, c < num_g_p
// from the coefficients of element (gmap[r], gmap_p[c]); out zeroed before
__global__ void k_analytic_eval(int model, int N, Mesh m, double2 freq, int num_gc,
                                const double2 *__restrict__ coeff, int num_g, int num_g_p,
                                const int32_t *__restrict__ gmap,
                                const int32_t *__restrict__ gmap_p, int64_t nr,
                                double2 *__restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * kEB + threadIdx.x;
  if (e >= (int64_t)num_g * num_g_p) return;
  const int r = (int)(e % num_g), c = (int)(e / num_g);
  const int64_t ne = (int64_t)num_gc * num_gc;
  const int64_t el = (int64_t)gmap_p[c] * num_gc + gmap[r];
  double2 res = mk(0.0, 0.0);
  if (model == 1) {  // godby_needs_model
    const double2 c1 = coeff[el], c2 = coeff[ne + el];
    if (sqrt(cabs2(c1)) > kEps8) {
      const double2 one = mk(1.0, 0.0);
      res = cmul(c1, cadd(cdiv(one, cadd(c2, freq)), cdiv(one, csub(c2, freq))));
    }
  } else if (model == SGW_AAA_APPROX) {  // aaa_approx_eval (analytic.f90:351-386)
    const int mmax = N / 3;
    int mm = 0;
    for (int c = 0; c < mmax; ++c)
      if (sqrt(cabs2(coeff[(int64_t)(2 * mmax + c) * ne + el])) > 1e-12) ++mm;
    // aaa_evaluate (aaa.f90:71-90): the barycentric form, or a support
    // point's value within 1e-14 of it
    double2 num = mk(0.0, 0.0), den = mk(0.0, 0.0);
    double dmin = INFINITY;
    int imin = 0;
    for (int c = 0; c < mm; ++c) {
      const double2 pc = coeff[(int64_t)c * ne + el];
      const double2 vc = coeff[(int64_t)(mmax + c) * ne + el];
      const double2 wc = coeff[(int64_t)(2 * mmax + c) * ne + el];
      double2 dz = csub(freq, pc);
      const double dist = sqrt(cabs2(dz));
      if (dist <= 1e-14) dz = mk(1e-14, 0.0);
      const double2 cc = cdiv(mk(1.0, 0.0), dz);
      num = cadd(num, cmul(cc, cmul(wc, vc)));
      den = cadd(den, cmul(cc, wc));
      if (dist < dmin) {
        dmin = dist;
        imin = c;
      }
    }
    res = (mm > 0 && dmin < 1e-14) ? coeff[(int64_t)(mmax + imin) * ne + el] : cdiv(num, den);
  } else if (model == SGW_AAA_POLE) {  // aaa_pole_eval (analytic.f90:430-459)
    const int half = N / 2;
    int npole = 0;
    for (int i = half; i < N; ++i)
      if (sqrt(cabs2(coeff[(int64_t)i * ne + el])) > 0.0) ++npole;
    for (int i = 0; i < npole && half + i < N; ++i)
      res = cadd(res, cdiv(coeff[(int64_t)(half + i) * ne + el],
                           csub(freq, coeff[(int64_t)i * ne + el])));
  } else if (model == SGW_PADE_ROBUST) {  // pade_eval_robust (pade_robust.f90:37-86)
    const int m = (int)lround(sqrt(cabs2(coeff[el])));
    const int n = (int)lround(sqrt(cabs2(coeff[ne + el])));
    if (4 + m + n > N) {  // degrees that do not fit the slices
      res = mk(NAN, NAN);
    } else {
      double2 num = mk(0.0, 0.0), den = mk(0.0, 0.0);
      for (int i = 2 + m; i >= 2; --i) num = cadd(coeff[(int64_t)i * ne + el], cmul(num, freq));
      for (int i = 3 + m + n; i >= 3 + m; --i)
        den = cadd(coeff[(int64_t)i * ne + el], cmul(den, freq));
      res = cdiv(num, den);
    }
  } else {  // pade_eval
    double2 a0 = mk(0.0, 0.0), a1 = coeff[el], b0 = mk(1.0, 0.0), b1 = mk(1.0, 0.0);
    for (int i = 1; i < N; ++i) {
      const double2 t = cmul(csub(freq, m.z[i - 1]), coeff[(int64_t)i * ne + el]);
      const double2 a2 = cadd(a1, cmul(t, a0)), b2 = cadd(b1, cmul(t, b0));
      a0 = a1; a1 = a2; b0 = b1; b1 = b2;
    }
    res = cdiv(a1, b1);
  }
  out[(int64_t)c * nr + r] = res;
}

__global__ void k_fill0(int64_t count, double2 *__restrict__ d) {
  const int64_t e = (int64_t)blockIdx.x * kEB + threadIdx.x;
  if (e < count) d[e] = mk(0.0, 0.0);
}

// array(r, c) *= alpha for r < num_g, c < num_g_p; sigma(r, c) += array(r, c)
__global__ void k_alpha_block(int num_g, int num_g_p, int64_t nr, double2 alpha,
                              double2 *__restrict__ a, double2 *__restrict__ sigma,
                              int64_t lds) {
  const int64_t e = (int64_t)blockIdx.x * kEB + threadIdx.x;
  if (e >= (int64_t)num_g * num_g_p) return;
  const int64_t r = e % num_g, c = e / num_g;
  const double2 v = cmul(alpha, a[c * nr + r]);
  a[c * nr + r] = v;
  if (sigma) sigma[c * lds + r] = cadd(sigma[c * lds + r], v);
}

template <typename F> int guarded(F &&fn) {
  try {
    return fn();
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::exception &e) {
    set_error(e.what());
    return SGW_E_STATE;
  }
}

Mesh make_mesh(int N, const sgw_complex *h_z, const int32_t *h_src) {
  SGW_REQUIRE(N >= 1 && N <= kMaxPade, SGW_E_SHAPE, "analytic: 1 <= frequencies <= 64");
  Mesh m{};
  for (int i = 0; i < N; ++i) {
    m.z[i] = make_double2(h_z[i].re, h_z[i].im);
    m.src[i] = h_src ? h_src[i] : i;
  }
  return m;
}

}  // namespace
}  // namespace sgw

using namespace sgw;

extern "C" int sgw_analytic_coeff(int32_t model, double thres, int32_t num_g, int32_t nfreq,
                                  const sgw_complex *h_z, const int32_t *h_src,
                                  sgw_complex *d_w, void *stream) {
  return guarded([&] {
    SGW_REQUIRE(d_w && h_z && num_g >= 1, SGW_E_ARG, "analytic_coeff: null argument");
    hipStream_t st = (hipStream_t)stream;
    const int64_t ne = (int64_t)num_g * num_g;
    double2 *w = reinterpret_cast<double2 *>(d_w);
    KScope ks("analytic_coeff", st);
    if (model == SGW_GODBY_NEEDS) {
      SGW_REQUIRE(nfreq == 2, SGW_E_SHAPE, "godby_needs_coeffs: must provide exactly 2 frequencies");
      const double omega_p = h_z[1].im;
      SGW_REQUIRE(omega_p >= 0.0, SGW_E_ARG, "godby_needs_coeffs: plasmon frequency must be positive");
      k_gn_coeffs<<<blocks(ne), kEB, 0, st>>>(ne, omega_p, w, w + ne);
    } else if (model == SGW_PADE_APPROX) {
      const Mesh m = make_mesh(nfreq, h_z, h_src);
      for (int i = 0; i < nfreq; ++i)
        SGW_REQUIRE(m.src[i] >= 0 && m.src[i] < nfreq, SGW_E_ARG,
                    "pade_coeff: source slice out of range");
      k_pade_coeffs<<<blocks(ne), kEB, 0, st>>>(ne, nfreq, m, w);
    } else if (model == SGW_PADE_ROBUST) {
      // pade_coeff_robust (pade_robust.f90:89-160): the solver frequencies
      // themselves, on a circle, first point (R, 0)
      SGW_REQUIRE(h_src == nullptr, SGW_E_ARG, "pade_coeff_robust: no symmetrised mesh (h_src = NULL)");
      SGW_REQUIRE(nfreq >= 10, SGW_E_SHAPE, "pade_coeff_robust: use at least 10 frequencies to form the circle");
      SGW_REQUIRE(nfreq <= kMaxPade, SGW_E_SHAPE, "pade_coeff_robust: at most 64 frequencies");
      const double radius = h_z[0].re;
      for (int i = 0; i < nfreq; ++i)
        SGW_REQUIRE(std::fabs(std::hypot(h_z[i].re, h_z[i].im) - radius) <= 1e-12, SGW_E_ARG,
                    "pade_coeff_robust: frequencies must span circle in the complex plane");
      SGW_REQUIRE(radius > 0.0, SGW_E_ARG, "pade_robust: radius in the complex plane must be > 0");
      const int nmax = nfreq / 2 - 2;
      const int per = (int)sizeof(double2) * pr_group_lds(nmax);   // <= 18432 bytes
      const int gpb = std::min(kPrGroups, kPrLdsMax / per);
      k_pade_robust_coeffs<<<(unsigned)((ne + gpb - 1) / gpb), kPrLanes * gpb, (size_t)gpb * per,
                             st>>>(ne, nfreq, radius, nmax, w);
    } else if (model == SGW_AAA_APPROX || model == SGW_AAA_POLE) {
      // aaa_generate per element on the symmetrised mesh (analytic.f90:162-199)
      const Mesh m = make_mesh(nfreq, h_z, h_src);
      for (int i = 0; i < nfreq; ++i)
        SGW_REQUIRE(m.src[i] >= 0 && m.src[i] < nfreq, SGW_E_ARG, "aaa: source slice out of range");
      SGW_REQUIRE(thres >= 0.0, SGW_E_ARG, "aaa: threshold must be >= 0");
      aaa_coeffs(model, thres, ne, nfreq, m.z, m.src, w, st);
    } else {
      throw Error{SGW_E_ARG, "analytic_coeff: model must be SGW_GODBY_NEEDS, SGW_PADE_APPROX, "
                             "SGW_PADE_ROBUST, SGW_AAA_APPROX or SGW_AAA_POLE"};
    }
    SGW_HIP(hipGetLastError());
    return 0;
  });
}

extern "C" int sgw_analytic_eval(int32_t model, int32_t num_gc, int32_t nfreq,
                                 const sgw_complex *h_z, const sgw_complex *d_coeff,
                                 sgw_complex freq, int32_t n, int32_t n_p, int32_t num_g,
                                 int32_t num_g_p, const int32_t *d_gmap,
                                 const int32_t *d_gmap_p, sgw_complex *d_out, void *stream) {
  return guarded([&] {
    const bool self_contained =
        model == SGW_PADE_ROBUST || model == SGW_AAA_APPROX || model == SGW_AAA_POLE;
    SGW_REQUIRE(d_coeff && d_gmap && d_gmap_p && d_out && (h_z || self_contained), SGW_E_ARG,
                "analytic_eval: null argument");
    SGW_REQUIRE(model >= SGW_GODBY_NEEDS && model <= SGW_AAA_POLE, SGW_E_ARG,
                "analytic_eval: model must be SGW_GODBY_NEEDS, SGW_PADE_APPROX, SGW_PADE_ROBUST, "
                "SGW_AAA_APPROX or SGW_AAA_POLE");
    SGW_REQUIRE(model != SGW_AAA_APPROX || (nfreq >= 3 && nfreq <= kMaxPade), SGW_E_SHAPE,
                "analytic_eval: AAA coefficients need 3..64 slices");
    SGW_REQUIRE(model != SGW_AAA_POLE || (nfreq >= 2 && nfreq <= kMaxPade), SGW_E_SHAPE,
                "analytic_eval: AAA pole coefficients need 2..64 slices");
    SGW_REQUIRE(model != SGW_PADE_ROBUST || (nfreq >= 4 && nfreq <= kMaxPade), SGW_E_SHAPE,
                "analytic_eval: robust Pade coefficients need 4..64 slices");
    SGW_REQUIRE(model != SGW_GODBY_NEEDS || nfreq == 2, SGW_E_SHAPE,
                "analytic_eval: Godby-Needs uses 2 coefficients");
    const int64_t nr = (int64_t)n * n * n, nr_p = (int64_t)n_p * n_p * n_p;
    SGW_REQUIRE(num_g >= 1 && num_g <= nr && num_g_p >= 1 && num_g_p <= nr_p, SGW_E_SHAPE,
                "analytic_eval: need num_g <= n^3 and num_g_p <= n_p^3");
    const Mesh m = self_contained ? Mesh{} : make_mesh(nfreq, h_z, nullptr);
    hipStream_t st = (hipStream_t)stream;
    double2 *out = reinterpret_cast<double2 *>(d_out);
    KScope ks("analytic_eval", st);
    k_fill0<<<blocks(nr * nr_p), kEB, 0, st>>>(nr * nr_p, out);
    k_analytic_eval<<<blocks((int64_t)num_g * num_g_p), kEB, 0, st>>>(
        model, nfreq, m, make_double2(freq.re, freq.im), num_gc,
        reinterpret_cast<const double2 *>(d_coeff), num_g, num_g_p, d_gmap, d_gmap_p, nr, out);
    SGW_HIP(hipGetLastError());
    return 0;
  });
}

extern "C" int sgw_sigma_prod(int32_t n, int32_t n_p, int32_t num_g, const int32_t *d_nl,
                              int32_t num_g_p, const int32_t *d_nl_p, double omega,
                              sgw_complex alpha, const sgw_complex *d_green, sgw_complex *d_array,
                              sgw_complex *d_sigma, int64_t ld_sigma, void *stream) {
  return guarded([&] {
    SGW_REQUIRE(d_green && d_array, SGW_E_ARG, "sigma_prod: null argument");
    SGW_REQUIRE(!d_sigma || ld_sigma >= num_g, SGW_E_SHAPE, "sigma_prod: ld_sigma < num_g");
    Fft6Args a{n, n_p, num_g, num_g_p, d_nl, d_nl_p, omega, 1,
               reinterpret_cast<double2 *>(d_array)};
    fft6_check(a, "sigma_prod");
    hipStream_t st = (hipStream_t)stream;
    const int64_t nr = (int64_t)n * n * n, nr_p = (int64_t)n_p * n_p * n_p;
    DevBuf<double2> S((size_t)num_g * nr_p);
    KScope ks("sigma_prod", st);
    // invfft6 (sigma.f90:487), the product with G (:499) and fwfft6 (:510).
    // Along r: the first xy pass builds its planes from the rows S[g][r']
    // (no zero-filled matrix in HBM), the z passes of both transforms and
    // the product are one kernel, and the last xy pass keeps only the rows'
    // positions.  The plane lists: the G vectors of each z plane of the box.
    std::vector<int32_t> nl(num_g);
    SGW_HIP(hipMemcpyAsync(nl.data(), d_nl, sizeof(int32_t) * num_g, hipMemcpyDeviceToHost, st));
    SGW_HIP(hipStreamSynchronize(st));
    for (int g = 0; g < num_g; ++g)
      SGW_REQUIRE(nl[g] >= 0 && nl[g] < nr, SGW_E_ARG, "sigma_prod: box index out of range");
    std::vector<int32_t> start(n + 1, 0), gl(num_g), pos(num_g);
    for (int g = 0; g < num_g; ++g) ++start[nl[g] / (n * n) + 1];
    for (int z = 0; z < n; ++z) start[z + 1] += start[z];
    std::vector<int32_t> fill(start.begin(), start.end() - 1);
    for (int g = 0; g < num_g; ++g) {
      const int k = fill[nl[g] / (n * n)]++;
      gl[k] = g;
      pos[k] = nl[g] % (n * n);
    }
    DevBuf<int32_t> d_start(n + 1), d_gl(num_g), d_pos(num_g);
    d_start.upload(start.data(), n + 1, st);
    d_gl.upload(gl.data(), num_g, st);
    d_pos.upload(pos.data(), num_g, st);
    invfft6_rows(st, a, S.p, 0);
    box_xy_scatter(st, n, nr_p, S.p, d_start.p, d_gl.p, d_pos.p, a.f, -1);
    box_z_product(st, n, nr_p, a.f, reinterpret_cast<const double2 *>(d_green));
    box_xy_gather(st, n, nr_p, a.f, d_start.p, d_gl.p, d_pos.p, 1.0 / (double)nr, S.p, +1);
    fwfft6_rows(st, a, S.p, 0);
    k_alpha_block<<<blocks((int64_t)num_g * num_g_p), kEB, 0, st>>>(
        num_g, num_g_p, nr, make_double2(alpha.re, alpha.im), a.f,
        reinterpret_cast<double2 *>(d_sigma), ld_sigma);                  // :519-521, :741
    SGW_HIP(hipGetLastError());
    SGW_HIP(hipStreamSynchronize(st));  // scratch freed on return
    return 0;
  });
}
