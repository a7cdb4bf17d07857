// AAA continuation of W on the device: aaa_generate of the vendored AAA
// library (vendor/analytic/src/aaa.f90:53-302; Nakatsukasa, Sete and
// Trefethen, SIAM J. Sci. Comput. 40 (2018) A1494) for every (G, G')
// element, and the pole-residue form of analytic.f90:389-427.
//
// One wavefront per element, lane i on mesh point i (at most 64 points).
// Per greedy step the Loewner submatrix (rows: the points that are not
// support points, columns: the support points in mesh order) sits in LDS
// column by column; its last right singular vector comes from one-sided
// (Hestenes) Jacobi with the rotations accumulated in V (the reference calls
// ZGESDD and takes the conjugated last row of V^H, which is that vector up
// to a phase; the phase cancels in the barycentric form).  Poles: the m - 1
// zeros of the barycentric denominator sum_j w_j / (z - z_j), i.e. the
// finite eigenvalues of the reference's arrowhead pencil (ZGGEV there),
// found together by Aberth-Ehrlich iteration, one lane per zero.
#include <algorithm>
#include <cmath>

#include "aaa.h"
#include "sgw_internal.h"

namespace sgw {
namespace {

constexpr int kAaaMax = 64;                             // mesh points
constexpr int kVs = 32;                                 // column stride of V
constexpr double kEps6 = 1e-6, kEps14 = 1e-14;
constexpr double kDblEpsA = 2.220446049250313e-16;
constexpr int kSweeps = 40, kAberthIters = 500;

__device__ inline double wsum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ inline double wmax(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline double2 wsum2(double2 v) { return mk(wsum(v.x), wsum(v.y)); }
__device__ inline double cabs2a(double2 a) { return a.x * a.x + a.y * a.y; }
__device__ inline double cabsa(double2 a) { return hypot(a.x, a.y); }

// (value, index) of the largest value over the wavefront, the lowest index
// among equal ones (MAXLOC); every lane gets the same pair
__device__ inline int wargmax(double v, int i) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double v2 = __shfl_xor(v, o, 64);
    const int i2 = __shfl_xor(i, o, 64);
    if (v2 > v || (v2 == v && i2 < i)) {
      v = v2;
      i = i2;
    }
  }
  return i;
}

__device__ inline uint64_t below(int l) { return l == 0 ? 0ull : (~0ull >> (64 - l)); }

// 1 / (x - y), eps14 for |x - y| <= eps14 (construct_Cauchy_matrix, aaa.f90:308-327)
__device__ inline double2 cauchy(double2 x, double2 y) {
  double2 d = csub(x, y);
  if (cabsa(d) <= kEps14) d = mk(kEps14, 0.0);
  return cdiv(mk(1.0, 0.0), d);
}

// aaa_evaluate (aaa.f90:71-90) at x from m support points in LDS
__device__ double2 aaa_eval_at(double2 x, int m, const double2 *pos, const double2 *val,
                               const double2 *wgt) {
  double2 num = mk(0.0, 0.0), den = mk(0.0, 0.0);
  double dmin = INFINITY;
  int imin = 0;
  for (int c = 0; c < m; ++c) {
    const double2 cc = cauchy(x, pos[c]);
    num = cadd(num, cmul(cc, cmul(wgt[c], val[c])));
    den = cadd(den, cmul(cc, wgt[c]));
    const double dist = cabsa(csub(x, pos[c]));
    if (dist < dmin) {
      dmin = dist;
      imin = c;
    }
  }
  return dmin < kEps14 ? val[imin] : cdiv(num, den);
}

struct AaaMesh {
  double2 z[kAaaMax];
  int src[kAaaMax];
};

// LDS per wavefront (double2 units): A (64 x mcap), V (32 x mcap), pos, val,
// wgt, lam (64 each)
__host__ __device__ inline int aaa_lds(int mcap) { return (kAaaMax + kVs) * mcap + 4 * kAaaMax; }

// model 4: slices [pos (mmax), val (mmax), wgt (mmax), 0...], mmax = max_point;
// model 5: [poles, 0..., residues, 0...] in halves of N / 2.
// status[0] = 1: a model-5 element needed more than mcap support points;
// status[1] = 1: more relevant poles than N / 2 (pole_correction's error)
__global__ __launch_bounds__(64) void k_aaa_coeffs(int64_t ne, int M, AaaMesh mesh, double thres,
                                                  int model, int max_point, int mcap,
                                                  double2 *__restrict__ w,
                                                  int32_t *__restrict__ status) {
  extern __shared__ double2 aaa_smem[];
  const int64_t e = blockIdx.x;
  const int l = threadIdx.x;
  double2 *A = aaa_smem;
  double2 *V = A + kAaaMax * mcap;
  double2 *pos = V + kVs * mcap;
  double2 *val = pos + kAaaMax;
  double2 *wgt = val + kAaaMax;
  double2 *lam = wgt + kAaaMax;
  const bool pt = l < M;
  const double2 zl = pt ? mesh.z[l] : mk(0.0, 0.0);
  const double2 fl = pt ? w[(int64_t)mesh.src[l] * ne + e] : mk(0.0, 0.0);
  // setup_work_type (:119-148)
  const double thres_abs = thres * wmax(pt ? cabsa(fl) : 0.0);
  const double2 sumf = wsum2(fl);
  double2 fit = mk(sumf.x / M, sumf.y / M);
  bool sup = false;
  int m = 0;
  const int mp = max_point;
  for (;;) {
    // update_support_point (:178-187)
    const int j = wargmax(pt ? cabsa(csub(fl, fit)) : -1.0, l);
    if (l == j) {
      fit = fl;
      sup = true;
    }
    ++m;
    const uint64_t smask = __ballot(sup) & (M == 64 ? ~0ull : (1ull << M) - 1);
    const uint64_t rmask = ~smask & (M == 64 ? ~0ull : (1ull << M) - 1);
    const int nrows = M - m;
    if (m > mcap) {  // the device keeps at most mcap support points
      if (l == 0) status[0] = 1;
      break;
    }
    __syncwarp();
    if (sup) {
      const int c = __popcll(smask & below(l));
      pos[c] = zl;
      val[c] = fl;
    }
    __syncwarp();
    // determine_weight (:205-225): the Loewner submatrix, rows = the other
    // points in mesh order, columns = the support points in mesh order
    const int r = __popcll(rmask & below(l));
    if (pt && !sup)
      for (int c = 0; c < m; ++c) A[c * kAaaMax + r] = cdiv(csub(fl, val[c]), csub(zl, pos[c]));
    if (l < m)
      for (int c = 0; c < m; ++c) V[c * kVs + l] = mk(l == c ? 1.0 : 0.0, 0.0);
    __syncwarp();
    int cmin = m - 1;  // an empty submatrix: V = I, the weight is e_m
    if (nrows > 0) {
      const double ctol = sqrt((double)nrows) * kDblEpsA;
      for (int sweep = 0; sweep < kSweeps; ++sweep) {
        int rot = 0;
        for (int p = 0; p + 1 < m; ++p)
          for (int q = p + 1; q < m; ++q) {
            const double2 ap = l < nrows ? A[p * kAaaMax + l] : mk(0.0, 0.0);
            const double2 aq = l < nrows ? A[q * kAaaMax + l] : mk(0.0, 0.0);
            const double al = wsum(cabs2a(ap)), be = wsum(cabs2a(aq));
            double2 gm = mk(0.0, 0.0);
            cmacc(gm, ap, aq);
            gm = wsum2(gm);
            const double g = hypot(gm.x, gm.y);
            if (!(g > ctol * sqrt(al) * sqrt(be))) continue;
            ++rot;
            const double zeta = (be - al) / (2.0 * g);
            const double az = fabs(zeta);
            const double t =
                az > 1e150 ? 0.5 / zeta : copysign(1.0, zeta) / (az + sqrt(1.0 + zeta * zeta));
            const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
            const double2 ph = mk(gm.x / g, -gm.y / g);
            if (l < nrows) {
              const double2 aqt = cmul(aq, ph);
              A[p * kAaaMax + l] = mk(cs * ap.x - sn * aqt.x, cs * ap.y - sn * aqt.y);
              A[q * kAaaMax + l] = mk(sn * ap.x + cs * aqt.x, sn * ap.y + cs * aqt.y);
            }
            if (l < m) {
              const double2 vp = V[p * kVs + l], vqt = cmul(V[q * kVs + l], ph);
              V[p * kVs + l] = mk(cs * vp.x - sn * vqt.x, cs * vp.y - sn * vqt.y);
              V[q * kVs + l] = mk(sn * vp.x + cs * vqt.x, sn * vp.y + cs * vqt.y);
            }
          }
        if (rot == 0) break;
      }
      // the right singular vector of the smallest singular value
      double smin = INFINITY;
      for (int c = 0; c < m; ++c) {
        const double2 a = l < nrows ? A[c * kAaaMax + l] : mk(0.0, 0.0);
        const double s = wsum(cabs2a(a));
        if (s <= smin) {  // the last of equal ones, as ZGESDD's descending order
          smin = s;
          cmin = c;
        }
      }
    }
    __syncwarp();
    if (l < m) wgt[l] = V[cmin * kVs + l];
    __syncwarp();
    // update_fit (:228-258)
    if (pt && !sup) {
      double2 num = mk(0.0, 0.0), den = mk(0.0, 0.0);
      for (int c = 0; c < m; ++c) {
        const double2 cc = cauchy(zl, pos[c]);
        num = cadd(num, cmul(cc, cmul(wgt[c], val[c])));
        den = cadd(den, cmul(cc, wgt[c]));
      }
      fit = cdiv(num, den);
    }
    // converged (:276-282) or max_step_reached (:285-292)
    const double err = wmax(pt ? (cabsa(csub(fit, fl)) <= thres_abs ? 0.0 : 1.0) : 0.0);
    if (err == 0.0 || m >= mp) break;
  }
  __syncwarp();
  if (m > mcap) {  // flagged: the coefficients are left as NaN
    for (int s = l; s < M; s += 64) w[(int64_t)s * ne + e] = mk(NAN, NAN);
    return;
  }
  if (model == 4) {
    const int mmax = max_point;
    for (int s = l; s < M; s += 64) {
      double2 v = mk(0.0, 0.0);
      if (s < m) v = pos[s];
      else if (s >= mmax && s < mmax + m) v = val[s - mmax];
      else if (s >= 2 * mmax && s < 2 * mmax + m) v = wgt[s - 2 * mmax];
      w[(int64_t)s * ne + e] = v;
    }
    return;
  }
  // model 5: poles (zeros of the denominator) by Aberth-Ehrlich, then the
  // residues of calculate_residual (:458-500) and pole_correction's filter
  const int nr = m - 1;
  double2 cen = wsum2(l < m ? pos[l] : mk(0.0, 0.0));
  cen = mk(cen.x / m, cen.y / m);
  double rad = wmax(l < m ? cabsa(csub(pos[l], cen)) : 0.0);
  if (!(rad > 0.0)) rad = 1.0;
  double2 lk = mk(0.0, 0.0);
  if (l < nr) {
    double s, c;
    sincospi(2.0 * l / nr + 0.23, &s, &c);
    lk = mk(cen.x + 1.5 * rad * c, cen.y + 1.5 * rad * s);
  }
  bool done = l >= nr;
  double prev = INFINITY;
  for (int it = 0; it < kAberthIters; ++it) {
    __syncwarp();
    lam[l] = lk;
    __syncwarp();
    if (!done) {
      double2 sd = mk(0.0, 0.0), sd1 = mk(0.0, 0.0), sq = mk(0.0, 0.0);
      for (int c = 0; c < m; ++c) {
        const double2 inv = cdiv(mk(1.0, 0.0), csub(lk, pos[c]));
        const double2 t = cmul(wgt[c], inv);
        sd = cadd(sd, t);
        sd1 = cadd(sd1, cmul(t, inv));
        sq = cadd(sq, inv);
      }
      // p'/p = q'/q + D'/D with D = p / q, q = prod (z - z_j)
      const double2 ratio = csub(sq, cdiv(sd1, sd));
      const double2 nk = cdiv(mk(1.0, 0.0), ratio);
      double2 ak = mk(0.0, 0.0);
      for (int jj = 0; jj < nr; ++jj)
        if (jj != l) ak = cadd(ak, cdiv(mk(1.0, 0.0), csub(lk, lam[jj])));
      const double2 dl = cdiv(nk, csub(mk(1.0, 0.0), cmul(nk, ak)));
      const bool fin = isfinite(dl.x) && isfinite(dl.y);
      if (fin) lk = csub(lk, dl);
      // converged, or the corrections stopped shrinking at the rounding level
      const double a = fin ? cabsa(dl) : 0.0, al = cabsa(lk);
      done = a <= 1e-14 * al || (a >= prev && a <= 1e-10 * al);
      prev = a;
    }
    if (__ballot(!done) == 0ull) break;
  }
  // finite zeros only (the pencil's infinite eigenvalues are not zeros of
  // the denominator; a zero run off to infinity means sum w_j = 0)
  const bool finite_pole = l < nr && cabsa(lk) < 1e12 * (rad + cabsa(cen));
  double2 res = mk(0.0, 0.0);
  if (finite_pole) {
    const double2 sh[4] = {mk(kEps6, 0.0), mk(0.0, kEps6), mk(-kEps6, 0.0), mk(0.0, -kEps6)};
#pragma unroll
    for (int k = 0; k < 4; ++k)
      res = cadd(res, cmul(aaa_eval_at(cadd(lk, sh[k]), m, pos, val, wgt), sh[k]));
    res = mk(res.x / 4.0, res.y / 4.0);
  }
  const bool keep = finite_pole && cabsa(res) > thres;
  const uint64_t kmask = __ballot(keep);
  const int cnt = __popcll(kmask), half = M / 2;
  if (cnt > half) {
    if (l == 0) status[1] = 1;
    for (int s = l; s < M; s += 64) w[(int64_t)s * ne + e] = mk(NAN, NAN);
    return;
  }
  __syncwarp();
  if (keep) {
    const int k = __popcll(kmask & below(l));
    pos[k] = lk;   // the support points are no longer needed
    val[k] = res;
  }
  __syncwarp();
  for (int s = l; s < M; s += 64) {
    double2 v = mk(0.0, 0.0);
    if (s < cnt) v = pos[s];
    else if (s >= half && s < half + cnt) v = val[s - half];
    w[(int64_t)s * ne + e] = v;
  }
}

}  // namespace

void aaa_coeffs(int model, double thres, int64_t ne, int nfreq, const double2 *z,
                const int32_t *src, double2 *d_w, hipStream_t st) {
  SGW_REQUIRE(model == 4 || model == 5, SGW_E_ARG, "aaa: model 4 or 5");
  SGW_REQUIRE(nfreq >= 3 && nfreq <= kAaaMax, SGW_E_SHAPE, "aaa: 3 <= frequencies <= 64");
  AaaMesh mesh{};
  for (int i = 0; i < nfreq; ++i) {
    mesh.z[i] = z[i];
    mesh.src[i] = src ? src[i] : i;
  }
  // model 4: mmax = N / 3 support points at most (analytic.f90:174-175);
  // model 5: max_point = N, of which the device keeps N / 2 (beyond that the
  // Loewner submatrix has more columns than rows and its null vector is not
  // unique)
  const int max_point = model == 4 ? nfreq / 3 : nfreq;
  const int mcap = model == 4 ? nfreq / 3 : nfreq / 2;
  const size_t lds = sizeof(double2) * (size_t)aaa_lds(mcap);
  DevBuf<int32_t> status(2);
  SGW_HIP(hipMemsetAsync(status.p, 0, 2 * sizeof(int32_t), st));
  k_aaa_coeffs<<<(unsigned)ne, 64, lds, st>>>(ne, nfreq, mesh, thres, model, max_point, mcap,
                                              d_w, status.p);
  SGW_HIP(hipGetLastError());
  int32_t h[2] = {0, 0};
  SGW_HIP(hipMemcpyAsync(h, status.p, sizeof(h), hipMemcpyDeviceToHost, st));
  SGW_HIP(hipStreamSynchronize(st));
  SGW_REQUIRE(h[0] == 0, SGW_E_STATE,
              "aaa: an element needs more than N/2 support points (its coefficients are NaN)");
  SGW_REQUIRE(h[1] == 0, SGW_E_STATE,
              "aaa: two many relevant poles, try reducing the coulomb threshold or increasing "
              "the number of frequencies (the element's coefficients are NaN)");
}

}  // namespace sgw
