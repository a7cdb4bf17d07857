// Batched multishift BiCGStab(l) on the GPU.
//
// Algorithm: Frommer, Computing 70, 87 (2003), as implemented by SternheimerGW
// algo/linear_solver/src/bicgstab.f90:154-999 (Algorithms 2-5).  theta, phi
// and xi are carried as inverses (bicgstab.f90:30-32).  The seed system is the
// first shift; every other shift is carried as sigma_k - sigma_1 (:458).
//
// Batching: S independent systems (band x q) advance in lock step.  Each
// system freezes as soon as the reference's convergence test
// (||r_seed|| < threshold, :237,:245) passes, so every system performs exactly
// the iterations the per-system reference would.  All vector state lives in
// HBM; the scalar recurrences (rho, alpha, omega, phi, theta, mu, gamma...)
// run on the device in FP64, one thread per system.
//
// Layout (ld = padded vector length, zero padding is preserved by all
// updates):  R,U: [S][l+1][ld]   X,RT: [S][ld]
//            SU: [S][nsh][ld] (u_0 of each shifted system)   SX: [S][nsh][ld]
#include "sgw_internal.h"
#include "bicgstab.h"
#include "ktimer.h"
#include "operators.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>

namespace sgw {

namespace {

constexpr int kBlock = 256;
// slots from which a BiCG seed half step runs as one launch (one block per
// system) rather than two (dot product, then the update over a grid of
// tiles): the one-launch form is faster at 1024 slots, slower at 256
constexpr int kFuseHalves = 512;
constexpr int kShiftGroups = 4;  // default shifted-system groups per k_shift_cycle block

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------
__device__ inline double2 wave_sum(double2 v) {
  for (int off = 32; off > 0; off >>= 1) {
    v.x += __shfl_xor(v.x, off, 64);
    v.y += __shfl_xor(v.y, off, 64);
  }
  return v;
}

// Block-wide sum (deterministic); every thread returns the total.
template <int NT> __device__ inline double2 block_sum(double2 v, double2 *lds) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  double2 t = mk(0, 0);
  if (threadIdx.x < 64) {
    t = threadIdx.x < NT / 64 ? lds[threadIdx.x] : mk(0, 0);
    t = wave_sum(t);
    if (threadIdx.x == 0) lds[0] = t;
  }
  __syncthreads();
  t = lds[0];
  return t;
}

// sum_i a[i] * b[i]  (ZDOTU) over n entries by one kBlock block; every thread
// returns the total
__device__ inline double2 block_dotu(const double2 *__restrict__ pa,
                                     const double2 *__restrict__ pb, int64_t n, double2 *lds) {
  double2 acc = mk(0, 0);
#pragma unroll 8
  for (int64_t i = threadIdx.x; i < n; i += kBlock) cmac(acc, pa[i], pb[i]);  // loads in flight
  return block_sum<kBlock>(acc, lds);
}

// out[s] = sqrt(sum_i |a[s][i]|^2)
__global__ __launch_bounds__(kBlock) void k_norm(int nsys, const int32_t *__restrict__ active,
                                                 const double2 *__restrict__ a, int64_t sa,
                                                 int64_t n, double *__restrict__ out) {
  __shared__ double2 lds[kBlock / 64];
  const int s = blockIdx.x;
  if (!active[s]) return;
  const double2 *pa = a + s * sa;
  double2 acc = mk(0, 0);
#pragma unroll 8
  for (int64_t i = threadIdx.x; i < n; i += kBlock) {
    double2 v = pa[i];
    acc.x = fma(v.x, v.x, acc.x);
    acc.x = fma(v.y, v.y, acc.x);
  }
  acc = block_sum<kBlock>(acc, lds);
  if (threadIdx.x == 0) out[s] = sqrt(acc.x);
}

// ---------------------------------------------------------------------------
// scalar recurrences (one thread per system)
// ---------------------------------------------------------------------------
struct Scal {
  // seed, per system
  double2 *rho, *rho_old, *alpha, *alpha_old, *beta, *omega, *dot, *alpha_new;
  double2 *gamma, *gamma_p, *gamma_pp;  // [S][l]
  // shifts, per (system, shift): index s*nsh + k
  double2 *sigma, *inv_phi_old, *inv_phi, *inv_phi_new, *inv_theta, *alpha_s, *beta_s;
  double2 *fa, *fb, *ia;       // factors for L15/L19, L26 and 1/alpha^sigma
  double2 *mu;                 // [S*nsh][l(l+1)/2]
  double2 *coef;               // [S*nsh][l+2]: gamma'_1/xi, gamma''_j/xi (j=1..l-1), 1/psi
  // with the per-shift test: a shifted system whose residual estimate
  // |zeta_k| ||r_seed|| fell below half the threshold at the end of an
  // iteration keeps its x^sigma from then on (k_shift_cycle skips it)
  int32_t *frozen;             // [S*nsh]
};

// bicgstab.f90:595-599  rho = (r_j, r0~); beta = alpha rho / rho_old; rho_old = rho.
// `begin` (first step of an iteration) first applies bicgstab.f90:587,
// rho_old = -omega rho_old.  `commit` completes L18 of the previous step
// (alpha_old = alpha = alpha_new), which the shifted-scalar update leaves to
// this step so that its per-shift threads can all read the old alpha.
// `fresh` (first step of an iteration only): a system in its first iteration
// has no previous step to complete; the flag is consumed here.  `count`
// (first step only) counts the iterations each system has begun.
__device__ inline void bicg_s1_body(int s, int begin, int commit, Scal sc,
                                    int32_t *__restrict__ fresh, int32_t *__restrict__ count,
                                    double2 rho) {
  if (begin) sc.rho_old[s] = cmul(cneg(sc.omega[s]), sc.rho_old[s]);
  if (count) count[s] += 1;
  if (fresh) {
    if (fresh[s]) commit = 0;
    fresh[s] = 0;
  }
  if (commit) {
    sc.alpha[s] = sc.alpha_new[s];
    sc.alpha_old[s] = sc.alpha_new[s];
  }
  sc.rho[s] = rho;
  sc.beta[s] = cdiv(cmul(sc.alpha[s], rho), sc.rho_old[s]);
  sc.rho_old[s] = rho;
}

// rho = (r_j, r0~) by block s, then the scalar step above by its thread 0
__global__ __launch_bounds__(kBlock) void k_dotu_s1(const int32_t *__restrict__ active,
                                                    const double2 *__restrict__ a, int64_t sa,
                                                    const double2 *__restrict__ b, int64_t n,
                                                    int begin, int commit, Scal sc,
                                                    int32_t *__restrict__ fresh,
                                                    int32_t *__restrict__ count) {
  __shared__ double2 lds[kBlock / 64];
  const int s = blockIdx.x;
  if (!active[s]) return;
  const double2 acc = block_dotu(a + s * sa, b + s * n, n, lds);
  if (threadIdx.x == 0) {
    sc.dot[s] = acc;
    bicg_s1_body(s, begin, commit, sc, fresh, count, acc);
  }
}

// bicgstab.f90:614-670  alpha = rho / (u_{j+1}, r0~) and the shifted scalars
// of shifted system k; `dot` = (u_{j+1}, r0~).  The per-step factors go to slot
// (s*nsh + k)*l + j: the shifted vectors are updated once per iteration, for
// all l steps (k_shift_cycle).
__device__ inline void bicg_s2_body(int s, int k, int nsh, int l, int j, Scal sc, double2 dot) {
  const double2 one = mk(1, 0);
  const double2 alpha = cdiv(sc.rho[s], dot);
  if (k == 0) sc.alpha_new[s] = alpha;
  if (nsh == 0) return;
  const double2 beta = sc.beta[s], alpha_old = sc.alpha_old[s];
  const double2 ab = cdiv(cmul(alpha, beta), alpha_old);
  const int i = s * nsh + k;
  double2 ipo = sc.inv_phi_old[i], ip = sc.inv_phi[i];
  double2 q = cdiv(ip, ipo);
  double2 den = cadd(cadd(one, cmul(alpha, sc.sigma[i])), cmul(ab, csub(q, one)));
  double2 ipn = cdiv(ip, den);                                     // L13
  double2 beta_s = cmul(cmul(q, q), beta);
  double2 alpha_s = cmul(cdiv(ipn, ip), alpha);
  double2 th = sc.inv_theta[i];
  const int pj = i * l + j;
  sc.fa[pj] = cmul(th, ip);                                        // 1/(theta phi) before L18
  sc.beta_s[pj] = beta_s;
  sc.alpha_s[pj] = alpha_s;
  sc.inv_phi_new[i] = ipn;
  sc.inv_phi_old[i] = ip;                                          // L18
  sc.inv_phi[i] = ipn;
  sc.fb[pj] = cneg(cmul(th, ipn));                                 // L26 factor
  sc.ia[pj] = cdiv(one, alpha_s);
}

// (u_{j+1}, r0~) by block s, then the shifted scalars of system s, one thread
// per shifted system (the k = 0 thread also stores the seed alpha)
__global__ __launch_bounds__(kBlock) void k_dotu_s2(const int32_t *__restrict__ active,
                                                    const double2 *__restrict__ a, int64_t sa,
                                                    const double2 *__restrict__ b, int64_t n,
                                                    int nsh, int l, int j, Scal sc) {
  __shared__ double2 lds[kBlock / 64];
  const int s = blockIdx.x;
  if (!active[s]) return;
  const double2 acc = block_dotu(a + s * sa, b + s * n, n, lds);
  if (threadIdx.x == 0) sc.dot[s] = acc;
  for (int k = threadIdx.x; k < max(nsh, 1); k += kBlock) bicg_s2_body(s, k, nsh, l, j, sc, acc);
}

// ---------------------------------------------------------------------------
// element-wise vector kernels; grid = (ceil(ld/kBlock), S)
// ---------------------------------------------------------------------------
// residual snapshot slots: r_i at the start of BiCG step j (i <= j), then
// r_{l-1} after the last step
__host__ __device__ constexpr int snap_index(int j, int i) { return j * (j + 1) / 2 + i; }
__host__ __device__ constexpr int snap_final(int l) { return l * (l + 1) / 2; }

struct Vecs {
  double2 *R, *U, *X, *RT, *SU, *SX, *RS;
  int64_t ld;
  int l, nsh, nsnap;
  __device__ double2 *rs(int s, int q) const { return RS + ((int64_t)s * nsnap + q) * ld; }
  __device__ double2 *r(int s, int i) const { return R + ((int64_t)s * (l + 1) + i) * ld; }
  __device__ double2 *u(int s, int i) const { return U + ((int64_t)s * (l + 1) + i) * ld; }
  __device__ double2 *x(int s) const { return X + (int64_t)s * ld; }
  // u_0^sigma of shifted system k (u_1..u_l^sigma are rebuilt each iteration)
  __device__ double2 *su(int s, int k) const { return SU + ((int64_t)s * nsh + k) * ld; }
  __device__ double2 *sx(int s, int k) const { return SX + ((int64_t)s * nsh + k) * ld; }
};

// L7-L8: u_i = r_i - beta u_i, i = 0..j
__global__ __launch_bounds__(kBlock) void k_bicg_v1(int j, const int32_t *__restrict__ active,
                                                    Vecs v, Scal sc) {
  const int s = blockIdx.y;
  if (!active[s]) return;
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= v.ld) return;
  const double2 mb = cneg(sc.beta[s]);
  for (int i = 0; i <= j; ++i) {
    double2 *pu = v.u(s, i) + e;
    double2 t = cmul(mb, *pu);
    *pu = cadd(t, v.r(s, i)[e]);
  }
}

// Seed half of step j after the matrix-vector product: L21-L22 (seed
// residuals r_i -= alpha u_{i+1}) and L25 (x += alpha u_0).  The residuals
// r_0..r_j before the update are kept as snapshots for the shifted systems
// (their L15/L19 use them); after the last step, r_{l-1} after its update
// (used by L26 of that step) is kept as well.
__global__ __launch_bounds__(kBlock) void k_bicg_seed(int j, const int32_t *__restrict__ active,
                                                      Vecs v, Scal sc) {
  const int s = blockIdx.y;
  if (!active[s]) return;
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= v.ld) return;
  const double2 alpha = sc.alpha_new[s];
  const double2 malpha = cneg(alpha);
  for (int i = 0; i <= j; ++i) {
    double2 t = v.r(s, i)[e];
    v.rs(s, snap_index(j, i))[e] = t;
    cmac(t, malpha, v.u(s, i + 1)[e]);
    v.r(s, i)[e] = t;
    if (j == v.l - 1 && i == j) v.rs(s, snap_final(v.l))[e] = t;
  }
  double2 *px = v.x(s) + e;
  double2 t = *px;
  cmac(t, alpha, v.u(s, 0)[e]);
  *px = t;
}

// The seed halves of BiCG step j around its H-apply, one block per system
// (bicgstab.f90:594-608, and 613-616 with 672-686; the shifted scalars
// bicgstab.f90:617-670):
//   k_bicg_half1: rho = (r_j, r0~) (ZDOTU), the scalar step of bicg_s1_body
//                 by thread 0, then L7-L8, u_i = r_i - beta u_i for i <= j;
//   k_bicg_half2: (u_{j+1}, r0~), alpha = rho / (u_{j+1}, r0~) and the
//                 shifted scalars (one thread per shifted system), then
//                 L21-L22 (r_i -= alpha u_{i+1}, the residuals before the
//                 update kept as snapshots for the shifted systems; after
//                 the last step r_{l-1} after it as well) and L25
//                 (x += alpha u_0).
// A block reads its own system's vectors only, so the dot product and the
// update that depends on it share one launch.
__global__ __launch_bounds__(kBlock) void k_bicg_half1(int j, const int32_t *__restrict__ active,
                                                       Vecs v, int begin, int commit, Scal sc,
                                                       int32_t *__restrict__ fresh,
                                                       int32_t *__restrict__ count) {
  __shared__ double2 lds[kBlock / 64];
  __shared__ double2 beta;
  const int s = blockIdx.x;
  if (!active[s]) return;
  const double2 acc = block_dotu(v.r(s, j), v.RT + (int64_t)s * v.ld, v.ld, lds);
  if (threadIdx.x == 0) {
    sc.dot[s] = acc;
    bicg_s1_body(s, begin, commit, sc, fresh, count, acc);
    beta = sc.beta[s];
  }
  __syncthreads();
  const double2 mb = cneg(beta);
  for (int64_t e = threadIdx.x; e < v.ld; e += kBlock)
    for (int i = 0; i <= j; ++i) {
      double2 *pu = v.u(s, i) + e;
      double2 t = cmul(mb, *pu);
      *pu = cadd(t, v.r(s, i)[e]);
    }
}

__global__ __launch_bounds__(kBlock) void k_bicg_half2(int j, const int32_t *__restrict__ active,
                                                       Vecs v, int nsh, int l, Scal sc) {
  __shared__ double2 lds[kBlock / 64];
  __shared__ double2 alpha_s;
  const int s = blockIdx.x;
  if (!active[s]) return;
  const double2 acc = block_dotu(v.u(s, j + 1), v.RT + (int64_t)s * v.ld, v.ld, lds);
  if (threadIdx.x == 0) {
    sc.dot[s] = acc;
    alpha_s = cdiv(sc.rho[s], acc);  // the alpha bicg_s2_body stores as alpha_new
  }
  for (int k = threadIdx.x; k < max(nsh, 1); k += kBlock) bicg_s2_body(s, k, nsh, l, j, sc, acc);
  __syncthreads();
  const double2 alpha = alpha_s, malpha = cneg(alpha);
  for (int64_t e = threadIdx.x; e < v.ld; e += kBlock) {
    for (int i = 0; i <= j; ++i) {
      double2 t = v.r(s, i)[e];
      v.rs(s, snap_index(j, i))[e] = t;
      cmac(t, malpha, v.u(s, i + 1)[e]);
      v.r(s, i)[e] = t;
      if (j == v.l - 1 && i == j) v.rs(s, snap_final(v.l))[e] = t;
    }
    double2 *px = v.x(s) + e;
    double2 t = *px;
    cmac(t, alpha, v.u(s, 0)[e]);
    *px = t;
  }
}

// All updates of the shifted systems in one iteration, one pass over their
// vectors: the l BiCG steps (L14-L19 and L26 of Algorithm 3, with the seed
// residual snapshots of each step) and, for systems still running after the
// BiCG half (mr_active), the MR step (L28-L34 of Algorithm 4, with the
// Gram-Schmidt residuals before the delayed r_0 update).  u_0^sigma and
// x^sigma are read and written once per iteration; u_1..u_l^sigma live in
// registers only.
template <int L>
__global__ __launch_bounds__(kBlock) void k_shift_cycle(const int32_t *__restrict__ active,
                                                        const int32_t *__restrict__ mr_active,
                                                        Vecs v, Scal sc, int groups) {
  // block = (kBlock / groups elements) x (groups of shifted systems): the
  // groups of one element read the same snapshots, which reach HBM once
  const int s = blockIdx.y, ne = kBlock / groups;
  if (!active[s]) return;
  const int kg = threadIdx.x / ne;
  const int64_t e = (int64_t)blockIdx.x * ne + threadIdx.x % ne;
  if (e >= v.ld) return;
  constexpr int l = L;  // compile-time degree: every array below stays in registers
  constexpr int nsnap = L * (L + 1) / 2 + 1;
  const bool do_mr = mr_active[s] != 0;
  double2 rs[nsnap];
#pragma unroll
  for (int q = 0; q < nsnap; ++q) rs[q] = v.rs(s, q)[e];
  double2 rr[L + 1];
#pragma unroll
  for (int i = 0; i <= l; ++i) rr[i] = do_mr ? v.r(s, i)[e] : mk(0, 0);
  const double2 *gam = sc.gamma + s * l;
  for (int k = kg; k < v.nsh; k += groups) {
    const int idx = s * v.nsh + k;
    if (sc.frozen[idx]) continue;  // uniform over the wavefront (one group per wavefront)
    // u_1..u_l are set by L19/L26 of steps 0..l-1 before any use, so only
    // u_0 and x carry over from the previous iteration
    double2 u[L + 1];
    u[0] = v.su(s, k)[e];
#pragma unroll
    for (int i = 1; i <= l; ++i) u[i] = mk(0, 0);
    double2 x = v.sx(s, k)[e];
    const double2 msg = cneg(sc.sigma[idx]);
#pragma unroll
    for (int j = 0; j < l; ++j) {
      const int pj = idx * l + j;
      const double2 mbs = cneg(sc.beta_s[pj]), fa = sc.fa[pj];
#pragma unroll
      for (int i = 0; i <= j; ++i) {                                 // L15
        double2 t = cmul(mbs, u[i]);
        cmac(t, fa, rs[snap_index(j, i)]);
        u[i] = t;
      }
      cmac(x, sc.alpha_s[pj], u[0]);                                 // L17
      // L19 then L26: u_{j+1} = (fa r_j + fb r_j') / alpha^sigma - sigma u_j,
      // r_j' = r_j after the seed update of step j
      const double2 rjn = (j + 1 < l) ? rs[snap_index(j + 1, j)] : rs[snap_final(l)];
      double2 uj1 = cmul(fa, rs[snap_index(j, j)]);
      cmac(uj1, sc.fb[pj], rjn);
      uj1 = cmul(sc.ia[pj], uj1);
      cmac(uj1, msg, u[j]);
      u[j + 1] = uj1;
    }
    if (do_mr) {
      const double2 *cf = sc.coef + (int64_t)idx * (kMaxL + 2);
      cmac(x, cf[0], rr[0]);                                         // L28
      double2 u0 = u[0];
      cmac(u0, cneg(gam[l - 1]), u[l]);
#pragma unroll
      for (int jj = 1; jj <= l - 1; ++jj) {                          // L29-31
        cmac(u0, cneg(gam[jj - 1]), u[jj]);
        cmac(x, cf[jj], rr[jj]);
      }
      u[0] = cmul(cf[l], u0);                                        // L34
    }
    v.su(s, k)[e] = u[0];
    v.sx(s, k)[e] = x;
  }
}

// Algorithm 4 seed updates: x, u_0 (L14-17) and the delayed r_0 (L32),
// bicgstab.f90:833-846, 919-925.  Runs after k_shift_cycle, which needs r_0
// before the delayed update.
template <int L>
__global__ __launch_bounds__(kBlock) void k_mr_seed(const int32_t *__restrict__ active, Vecs v,
                                                    Scal sc) {
  const int s = blockIdx.y;
  if (!active[s]) return;
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= v.ld) return;
  constexpr int l = L;
  double2 rr[L + 1];
#pragma unroll
  for (int i = 0; i <= l; ++i) rr[i] = v.r(s, i)[e];
  const double2 *gam = sc.gamma + s * l, *gp = sc.gamma_p + s * l, *gpp = sc.gamma_pp + s * l;
  double2 x = v.x(s)[e];
  cmac(x, gam[0], rr[0]);                                            // L14
  double2 u0 = v.u(s, 0)[e];
  cmac(u0, cneg(gam[l - 1]), v.u(s, l)[e]);
#pragma unroll
  for (int jj = 1; jj <= l - 1; ++jj) {                              // L15-17
    cmac(u0, cneg(gam[jj - 1]), v.u(s, jj)[e]);
    cmac(x, gpp[jj - 1], rr[jj]);
  }
  v.x(s)[e] = x;
  v.u(s, 0)[e] = u0;
  double2 r0 = rr[0];
#pragma unroll
  for (int jj = 1; jj <= l; ++jj) cmac(r0, cneg(gp[jj - 1]), rr[jj]);  // L32
  v.r(s, 0)[e] = r0;
}

// Algorithm 4, MGS part and all scalars (one block per system).
// bicgstab.f90:780-829 (seed) and :850-884 (per-shift Horner and gammas).
template <int NT>
__global__ __launch_bounds__(NT) void k_mr_mgs(const int32_t *__restrict__ active, Vecs v, Scal sc,
                                               int ntau, int nmu) {
  const int s = blockIdx.x;
  if (!active[s]) return;
  __shared__ double2 lds[NT / 64];
  __shared__ double2 tau[(kMaxL + 1) * (kMaxL + 2) / 2];
  __shared__ double2 nu[kMaxL + 1];
  __shared__ double2 gam[kMaxL + 1], gp[kMaxL + 1], gpp[kMaxL + 1];
  const int l = v.l;
  const int64_t ld = v.ld;
  for (int jj = 1; jj <= l; ++jj) {
    double2 *rj = v.r(s, jj);
    int offset = jj * (jj + 1) / 2 + 1;
    for (int ii = 1; ii <= jj - 1; ++ii) {
      const double2 *ri = v.r(s, ii);
      double2 acc = mk(0, 0);
      for (int64_t e = threadIdx.x; e < ld; e += NT) cmac(acc, rj[e], ri[e]);
      acc = block_sum<NT>(acc, lds);
      double2 t = cdiv(acc, nu[ii]);
      if (threadIdx.x == 0) tau[offset + ii - 1] = t;
      const double2 mt = cneg(t);
      for (int64_t e = threadIdx.x; e < ld; e += NT) {
        double2 x = rj[e];
        cmac(x, mt, ri[e]);
        rj[e] = x;
      }
      __syncthreads();
    }
    double2 a1 = mk(0, 0), a2 = mk(0, 0);
    const double2 *r0 = v.r(s, 0);
    for (int64_t e = threadIdx.x; e < ld; e += NT) {
      double2 x = rj[e];
      cmac(a1, x, x);
      cmac(a2, r0[e], x);
    }
    a1 = block_sum<NT>(a1, lds);
    a2 = block_sum<NT>(a2, lds);
    if (threadIdx.x == 0) {
      nu[jj] = a1;
      gp[jj] = cdiv(a2, a1);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    gam[l] = gp[l];                                                // L7
    sc.omega[s] = gam[l];
    for (int jj = l - 1; jj >= 1; --jj) {                          // L8-9
      double2 g = gp[jj];
      for (int ii = jj + 1; ii <= l; ++ii) {
        int ij = ii * (ii + 1) / 2 + jj + 1;
        g = csub(g, cmul(tau[ij - 1], gam[ii]));
      }
      gam[jj] = g;
    }
    for (int jj = 1; jj <= l - 1; ++jj) {                          // L11-12
      double2 g = gam[jj + 1];
      for (int ii = jj + 1; ii <= l - 1; ++ii) {
        int ij = ii * (ii + 1) / 2 + jj + 1;
        g = cadd(g, cmul(tau[ij - 1], gam[ii + 1]));
      }
      gpp[jj] = g;
    }
    for (int jj = 1; jj <= l; ++jj) {
      sc.gamma[s * l + jj - 1] = gam[jj];
      sc.gamma_p[s * l + jj - 1] = gp[jj];
      sc.gamma_pp[s * l + jj - 1] = (jj < l) ? gpp[jj] : mk(0, 0);
    }
  }
  __syncthreads();
  // per shift (one thread each): Horner (Algorithm 5), theta, gamma', gamma''
  for (int k = threadIdx.x; k < v.nsh; k += NT) {
    const int idx = s * v.nsh + k;
    const double2 sg = sc.sigma[idx], one = mk(1, 0);
    double2 g[kMaxL + 1];
    g[l] = cneg(gam[l]);
    for (int jj = l - 1; jj >= 1; --jj) g[jj] = csub(cmul(cneg(sg), g[jj + 1]), gam[jj]);
    double2 psi = cadd(cmul(cneg(sg), g[1]), one);
    for (int ii = 1; ii <= l - 1; ++ii)
      for (int jj = l - 1; jj >= ii; --jj) g[jj] = cadd(cmul(cneg(sg), g[jj + 1]), g[jj]);
    for (int jj = 1; jj <= l; ++jj) g[jj] = cdiv(cneg(g[jj]), psi);
    double2 inv_xi = cmul(sc.inv_theta[idx], sc.inv_phi[idx]);   // L21
    sc.inv_theta[idx] = cdiv(sc.inv_theta[idx], psi);
    const double2 *mu = sc.mu + (int64_t)idx * nmu;
    double2 gps[kMaxL + 1];
    for (int jj = 1; jj <= l; ++jj) {                              // L22-23
      double2 a = mk(0, 0);
      for (int ii = jj; ii <= l; ++ii) {
        int ij = (ii - 1) * ii / 2 + jj;
        a = cadd(a, cmul(mu[ij - 1], g[ii]));
      }
      gps[jj] = a;
    }
    double2 *cf = sc.coef + (int64_t)idx * (kMaxL + 2);
    cf[0] = cmul(gps[1], inv_xi);                                  // L28 factor
    for (int jj = 1; jj <= l - 1; ++jj) {                          // L25-26, L31 factor
      double2 a = gps[jj + 1];
      for (int ii = jj + 1; ii <= l - 1; ++ii) {
        int ij = ii * (ii + 1) / 2 + jj + 1;
        a = cadd(a, cmul(tau[ij - 1], gps[ii + 1]));
      }
      cf[jj] = cmul(a, inv_xi);
    }
    cf[l] = cdiv(one, psi);                                        // L34 factor
  }
  (void)ntau;
}

// Convergence test (bicgstab.f90:237,245,296-299) and NaN guard on the
// systems in `in`: out[s] = in[s] and not converged.  After the BiCG half
// `out` is mr_active (the MR half runs only for those); after the MR half
// `in` is mr_active and `out` the active set of the next iteration.
// shift_test = 0 is the reference's test on the seed residual alone; with
// shift_test = 1 every shifted residual must pass too, through its estimate
// ||r^sigma|| = |1 / (theta^sigma phi^sigma)| ||r_seed|| (the collinearity of
// Frommer's recurrences, bicgstab.f90:637-643), which can exceed the seed's.
// `count` holds the iterations each system has begun; after the MR half
// (final = 1) a system that has run max_iter of them without converging stops
// with state 3, iterations max_iter + 1 (the reference's loop variable after
// DO iter = 1, max_iter, bicgstab.f90:229-253).
__global__ void k_check(int nsys, const int32_t *__restrict__ count, int max_iter, int final,
                        double thr, const int32_t *__restrict__ in,
                        int32_t *__restrict__ out, const double *__restrict__ norm,
                        int32_t *__restrict__ iters, int32_t *__restrict__ state, Scal sc,
                        int nsh, int shift_test, double relax_thr, int32_t *__restrict__ relax) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsys) return;
  if (!in[s]) {
    out[s] = 0;
    return;
  }
  double nr = norm[s];
  if (shift_test) {
    double zmax = 1.0;
    for (int k = 0; k < nsh; ++k) {
      if (sc.frozen[s * nsh + k]) continue;
      const double2 z = cmul(sc.inv_theta[s * nsh + k], sc.inv_phi[s * nsh + k]);
      const double az = sqrt(z.x * z.x + z.y * z.y);
      // at the end of an iteration x^sigma holds all of its updates: a
      // shifted system below half the threshold is done (the half leaves
      // room for the gap between this estimate and the true residual)
      if (final && az * nr < 0.5 * thr) sc.frozen[s * nsh + k] = 1;
      else zmax = fmax(zmax, az);
    }
    nr *= zmax;
  }
  // relaxed matrix-vector products (inexact Krylov): once the residual is
  // small the error of an FP32 local-potential FFT, proportional to the
  // vector it multiplies, stays below the final tolerance
  if (relax && nr < relax_thr) relax[s] = 1;
  int keep = 1;
  if (nr < thr) {
    keep = 0;
    iters[s] = count[s];
    state[s] = 1;  // converged
  } else if (isnan(nr)) {
    keep = 0;
    iters[s] = count[s];
    state[s] = 2;  // NaN: the reference would run to max_iter and report ierr 2
  } else if (final && count[s] >= max_iter) {
    keep = 0;
    iters[s] = max_iter + 1;
    state[s] = 3;  // out of iterations: ierr 1
  }
  out[s] = keep;
}

// init: mu_ij, shift scalars, seed scalars of system s from the shifts of
// row q of sigma_all (bicgstab.f90:345-348, 400-483); init_shift_one does the
// part of shifted system k alone
__device__ void init_scalars_one(int s, int q, int nsh, int l, const Scal &sc,
                                 const double2 *__restrict__ sigma_all, int nshift_all);
__device__ void init_shift_one(int s, int q, int k, int nsh, int l, const Scal &sc,
                               const double2 *__restrict__ sigma_all, int nshift_all);

__global__ void k_init_scalars(int nsys, int nsh, int l, Scal sc,
                               const double2 *__restrict__ sigma_all, int nshift_all) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsys) return;
  init_scalars_one(s, s, nsh, l, sc, sigma_all, nshift_all);
}

__device__ void init_scalars_one(int s, int q, int nsh, int l, const Scal &sc,
                                 const double2 *__restrict__ sigma_all, int nshift_all) {
  const double2 one = mk(1, 0);
  sc.rho[s] = mk(0, 0);
  sc.rho_old[s] = one;
  sc.alpha_old[s] = one;
  sc.alpha[s] = mk(0, 0);
  sc.omega[s] = one;
  sc.beta[s] = mk(0, 0);
  for (int k = 0; k < nsh; ++k) init_shift_one(s, q, k, nsh, l, sc, sigma_all, nshift_all);
}

__device__ void init_shift_one(int s, int q, int k, int nsh, int l, const Scal &sc,
                               const double2 *__restrict__ sigma_all, int nshift_all) {
  const double2 one = mk(1, 0);
  const double2 s1 = sigma_all[(int64_t)q * nshift_all];
  const int nmu = l * (l + 1) / 2;
  const int idx = s * nsh + k;
  sc.inv_phi_old[idx] = one;
  sc.inv_phi[idx] = one;
  sc.inv_theta[idx] = one;
  sc.frozen[idx] = 0;
  double2 sg = csub(sigma_all[(int64_t)q * nshift_all + k + 1], s1);
  sc.sigma[idx] = sg;
  double2 spow[kMaxL + 1];
  spow[0] = one;
  for (int i = 1; i <= l - 1; ++i) spow[i] = cmul(sg, spow[i - 1]);
  double bin[(kMaxL + 1) * (kMaxL + 2) / 2];
  for (int jj = 0; jj <= l - 1; ++jj) {
    int offset = jj * (jj + 1) / 2 + 1;
    for (int ii = 0; ii <= jj; ++ii) {
      int ij = offset + ii;
      bin[ij - 1] = (ii == 0) ? 1.0 : (bin[ij - 2] * (jj - ii + 1)) / ii;
      sc.mu[(int64_t)idx * nmu + ij - 1] = cmul(mk(bin[ij - 1], 0.0), spow[jj - ii]);
    }
  }
}

// Streaming refill: pairs[2p] = slot s, pairs[2p + 1] = queued system q.
// Vectors of slot s start as for a new solve: r_0 = r0~ = b_q, all else 0.
__global__ void k_refill_vectors(const int32_t *__restrict__ pairs, Vecs v,
                                 const double2 *__restrict__ rhs, int64_t rhs_ld) {
  const int s = pairs[2 * blockIdx.y], q = pairs[2 * blockIdx.y + 1];
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= v.ld) return;
  const double2 b = rhs[(int64_t)q * rhs_ld + e], z = mk(0, 0);
  for (int i = 0; i <= v.l; ++i) {
    v.r(s, i)[e] = i == 0 ? b : z;
    v.u(s, i)[e] = z;
  }
  v.x(s)[e] = z;
  v.RT[(int64_t)s * v.ld + e] = b;
  for (int k = 0; k < v.nsh; ++k) {
    v.su(s, k)[e] = z;
    v.sx(s, k)[e] = z;
  }
}

// Scalars and flags of the refilled slots: the seed shift of q, a fresh
// first iteration, active.
__global__ void k_refill_scalars(int npairs, const int32_t *__restrict__ pairs, int nsh, int l,
                                 Scal sc, const double2 *__restrict__ sigma_all, int nshift_all,
                                 double2 *__restrict__ sigma_seed, int32_t *__restrict__ active,
                                 int32_t *__restrict__ iters, int32_t *__restrict__ state,
                                 int32_t *__restrict__ fresh, int32_t *__restrict__ count,
                                 int32_t *__restrict__ relax) {
  // one thread per (pair, shifted system); the k = 0 thread also does the seed
  const int nshw = nsh > 0 ? nsh : 1;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npairs * nshw) return;
  const int p = t / nshw, k = t % nshw;
  const int s = pairs[2 * p], q = pairs[2 * p + 1];
  if (nsh > 0) init_shift_one(s, q, k, nsh, l, sc, sigma_all, nshift_all);
  if (k != 0) return;
  init_scalars_one(s, q, 0, l, sc, sigma_all, nshift_all);  // seed scalars only
  sigma_seed[s] = sigma_all[(int64_t)q * nshift_all];
  active[s] = 1;
  iters[s] = 0;
  state[s] = 0;
  fresh[s] = 1;
  count[s] = 0;
  if (relax) relax[s] = 0;
}

// Compaction of running systems into the leading slots: pairs[2p] = source
// slot, pairs[2p + 1] = destination slot (sources and destinations disjoint).
// Every vector a later iteration reads moves; the residual snapshots are
// rebuilt within each iteration and stay.
__global__ void k_move_vectors(const int32_t *__restrict__ pairs, Vecs v) {
  const int s = pairs[2 * blockIdx.y], d = pairs[2 * blockIdx.y + 1];
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= v.ld) return;
  for (int i = 0; i <= v.l; ++i) {
    v.r(d, i)[e] = v.r(s, i)[e];
    v.u(d, i)[e] = v.u(s, i)[e];
  }
  v.x(d)[e] = v.x(s)[e];
  v.RT[(int64_t)d * v.ld + e] = v.RT[(int64_t)s * v.ld + e];
  for (int k = 0; k < v.nsh; ++k) {
    v.su(d, k)[e] = v.su(s, k)[e];
    v.sx(d, k)[e] = v.sx(s, k)[e];
  }
}

// per-slot int arrays that move with a system (null entries are skipped)
struct SlotFlags {
  int32_t *f[8];
};

// Scalars and flags of the moved systems (pairs as above), one thread per
// (pair, shifted system); the k = 0 thread also moves the seed's.  S is the
// solver's slot count (the seed arrays' stride).
__global__ void k_move_scalars(int npairs, const int32_t *__restrict__ pairs, int S, int nsh,
                               int l, Scal sc, double2 *__restrict__ seed_all,
                               double2 *__restrict__ sigma_seed, SlotFlags fl) {
  const int nshw = nsh > 0 ? nsh : 1;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npairs * nshw) return;
  const int p = t / nshw, k = t % nshw;
  const int s = pairs[2 * p], d = pairs[2 * p + 1];
  if (nsh > 0) {
    const int is = s * nsh + k, id = d * nsh + k;
    sc.sigma[id] = sc.sigma[is];
    sc.inv_phi_old[id] = sc.inv_phi_old[is];
    sc.frozen[id] = sc.frozen[is];
    sc.inv_phi[id] = sc.inv_phi[is];
    sc.inv_phi_new[id] = sc.inv_phi_new[is];
    sc.inv_theta[id] = sc.inv_theta[is];
    for (int j = 0; j < l; ++j) {
      sc.alpha_s[id * l + j] = sc.alpha_s[is * l + j];
      sc.beta_s[id * l + j] = sc.beta_s[is * l + j];
      sc.fa[id * l + j] = sc.fa[is * l + j];
      sc.fb[id * l + j] = sc.fb[is * l + j];
      sc.ia[id * l + j] = sc.ia[is * l + j];
    }
    const int nmu = l * (l + 1) / 2;
    for (int j = 0; j < nmu; ++j) sc.mu[(int64_t)id * nmu + j] = sc.mu[(int64_t)is * nmu + j];
    for (int j = 0; j < kMaxL + 2; ++j)
      sc.coef[(int64_t)id * (kMaxL + 2) + j] = sc.coef[(int64_t)is * (kMaxL + 2) + j];
  }
  if (k != 0) return;
  for (int q = 0; q < 8; ++q) seed_all[(int64_t)q * S + d] = seed_all[(int64_t)q * S + s];
  for (int j = 0; j < l; ++j) {
    sc.gamma[d * l + j] = sc.gamma[s * l + j];
    sc.gamma_p[d * l + j] = sc.gamma_p[s * l + j];
    sc.gamma_pp[d * l + j] = sc.gamma_pp[s * l + j];
  }
  sigma_seed[d] = sigma_seed[s];
  for (int f = 0; f < 8; ++f)
    if (fl.f[f]) fl.f[f][d] = fl.f[f][s];
}

// Outcome of the systems of pairs (slot s, system q): iterations, state and
// the running flag by system.
__global__ void k_retire(int npairs, const int32_t *__restrict__ pairs,
                         const int32_t *__restrict__ iters, const int32_t *__restrict__ state,
                         const int32_t *__restrict__ active, int32_t *__restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  const int s = pairs[2 * p], q = pairs[2 * p + 1];
  out[3 * q] = iters[s];
  out[3 * q + 1] = state[s];
  out[3 * q + 2] = active[s];
}

// x_out[q][k][ldo] <- X / SX of slot s for every pair (s, q)
__global__ void k_gather_pairs(const int32_t *__restrict__ pairs, Vecs v, double2 *__restrict__ out,
                               int64_t ldo, int64_t n) {
  const int s = pairs[2 * blockIdx.y], q = pairs[2 * blockIdx.y + 1], k = blockIdx.z;
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= n) return;
  const double2 val = (k == 0) ? v.x(s)[e] : v.sx(s, k - 1)[e];
  out[((int64_t)q * (v.nsh + 1) + k) * ldo + e] = val;
}

__global__ void k_init_vectors(int nsys, Vecs v, const double2 *__restrict__ rhs, int64_t rhs_ld) {
  const int s = blockIdx.y;
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= v.ld) return;
  double2 b = rhs[(int64_t)s * rhs_ld + e];
  v.r(s, 0)[e] = b;
  v.RT[(int64_t)s * v.ld + e] = b;
}

// x_out[s][k][ldo] <- X (k = 0) or SX (k >= 1)
__global__ void k_gather_x(Vecs v, double2 *__restrict__ out, int64_t ldo, int64_t n) {
  const int s = blockIdx.y, k = blockIdx.z;
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= n) return;
  double2 val = (k == 0) ? v.x(s)[e] : v.sx(s, k - 1)[e];
  out[((int64_t)s * (v.nsh + 1) + k) * ldo + e] = val;
}

__global__ void k_nan_scan(int nsys, int nshift, const double2 *__restrict__ x, int64_t ldo,
                           int64_t n, int32_t *__restrict__ has_nan) {
  const int s = blockIdx.x;
  int found = 0;
  for (int64_t i = threadIdx.x; i < (int64_t)nshift * n; i += blockDim.x) {
    int k = (int)(i / n);
    int64_t e = i % n;
    double2 v = x[((int64_t)s * nshift + k) * ldo + e];
    if (isnan(v.x) || isnan(v.y)) found = 1;
  }
  if (found) has_nan[s] = 1;
}

inline dim3 grid1(int n, int bs) { return dim3((n + bs - 1) / bs); }

}  // namespace

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------
// launch a kernel templated on the BiCGStab degree l (1..kMaxL)
template <template <int> class K, typename... Args>
void launch_l(int l, dim3 grid, dim3 block, hipStream_t st, Args... args) {
  switch (l) {
#define SGW_L(n) \
  case n: K<n>::launch(grid, block, st, args...); break;
    SGW_L(1) SGW_L(2) SGW_L(3) SGW_L(4) SGW_L(5) SGW_L(6) SGW_L(7) SGW_L(8)
    SGW_L(9) SGW_L(10) SGW_L(11) SGW_L(12) SGW_L(13) SGW_L(14) SGW_L(15)
#undef SGW_L
    default: throw Error{SGW_E_ARG, "bicgstab: lmax must be in [1, 15]"};
  }
}
template <int L> struct ShiftCycle {
  template <typename... A> static void launch(dim3 g, dim3 b, hipStream_t st, A... a) {
    k_shift_cycle<L><<<g, b, 0, st>>>(a...);
  }
};
template <int L> struct MrSeed {
  template <typename... A> static void launch(dim3 g, dim3 b, hipStream_t st, A... a) {
    k_mr_seed<L><<<g, b, 0, st>>>(a...);
  }
};

struct BatchBicgstab::Impl {
  int nsys, nshift, nsh, l;
  int shift_groups = kShiftGroups;  // shifted-system groups per k_shift_cycle block
  int64_t ld;
  DevBuf<double2> R, U, X, RT, SU, SX, RS;
  DevBuf<double2> scal_seed, scal_shift, mu, coef, gam;
  DevBuf<double2> sigma_seed;  // per system seed shift sigma_1
  DevBuf<double> norms;
  DevBuf<int32_t> active, mr_active, iters, state, count, frozen;
  Scal sc{};
  Vecs vv{};
  std::vector<double2> h_sigma_seed;
  std::vector<int32_t> h_active;
  int32_t *pinned_active = nullptr;  // 2 x nsys flags for the lagged convergence test
  hipEvent_t flag_ev[2] = {nullptr, nullptr};
  // streaming solves: first-iteration flags, the fixed group of each slot,
  // (slot, system) lists on the device (4 x 2 nsys) and pinned host staging
  // (2 x 3 nsys flag copies, then 4 x 2 nsys lists)
  DevBuf<int32_t> fresh, slot_group, pairs_dev;
  int32_t *pinned_stream = nullptr;

  // Iterations as a captured graph: `chunk` iterations are captured once per
  // solve on the solver's own stream and replayed by one hipGraphLaunch each
  // (the caller's stream may be the legacy default stream, which cannot be
  // captured; the solve joins it through events).  Off while kernel timing
  // is on, whose events cannot sit inside a graph.
  int graph_mode = 0;  // SGW_SOLVER_GRAPH=1 replays captured iterations
  int graph_chunk = 1; // SGW_GRAPH_CHUNK: iterations per graph launch (lock-step solves)
  hipStream_t own = nullptr;
  hipEvent_t join_ev[2] = {nullptr, nullptr};
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int graph_n = 0;

  ~Impl() {
    drop_graph();
    if (pinned_active) (void)hipHostFree(pinned_active);
    if (pinned_stream) (void)hipHostFree(pinned_stream);
    for (auto &e : flag_ev)
      if (e) (void)hipEventDestroy(e);
    for (auto &e : join_ev)
      if (e) (void)hipEventDestroy(e);
    if (own) (void)hipStreamDestroy(own);
  }

  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
    graph_n = 0;
  }

  // the stream a solve runs on: the caller's, or, with graphs, the solver's
  // own stream after the caller's earlier work
  hipStream_t begin_on(hipStream_t st, bool graphs) {
    if (!graphs) return st;
    if (!own) {
      SGW_HIP(hipStreamCreateWithFlags(&own, hipStreamNonBlocking));
      SGW_HIP(hipEventCreateWithFlags(&join_ev[0], hipEventDisableTiming));
      SGW_HIP(hipEventCreateWithFlags(&join_ev[1], hipEventDisableTiming));
    }
    SGW_HIP(hipEventRecord(join_ev[0], st));
    SGW_HIP(hipStreamWaitEvent(own, join_ev[0], 0));
    return own;
  }
  // the caller's stream continues after the solve's work
  void end_on(hipStream_t st, hipStream_t ws) {
    drop_graph();
    if (ws == st) return;
    SGW_HIP(hipEventRecord(join_ev[1], ws));
    SGW_HIP(hipStreamWaitEvent(st, join_ev[1], 0));
  }
  bool graphs_wanted(const Operator &op) const {
    return graph_mode && !op.host_side() && ktimer_level() == 0;
  }

  // n iterations on ws: eagerly, or by replaying the graph of n iterations
  // (captured at the first call of the solve)
  void run(Operator &op, hipStream_t ws, bool graphs, int n, int max_iter, double thr) {
    if (!graphs) {
      for (int i = 0; i < n; ++i) iterate(op, ws, max_iter, thr);
      return;
    }
    if (gexec && graph_n != n) drop_graph();
    if (!gexec) {
      SGW_HIP(hipStreamBeginCapture(ws, hipStreamCaptureModeThreadLocal));
      try {
        for (int i = 0; i < n; ++i) iterate(op, ws, max_iter, thr);
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(ws, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      SGW_HIP(hipStreamEndCapture(ws, &graph));
      SGW_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
      graph_n = n;
    }
    SGW_HIP(hipGraphLaunch(gexec, ws));
  }

  Impl(int nsys_, int nshift_, int l_, int64_t ld_)
      : nsys(nsys_), nshift(nshift_), nsh(nshift_ - 1), l(l_), ld(ld_) {
    SGW_REQUIRE(l >= 1 && l <= kMaxL, SGW_E_ARG, "bicgstab: lmax must be in [1, 15]");
    if (const char *c = std::getenv("SGW_SHIFT_GROUPS")) {  // 1, 2, 4, 8 or 16
      const int g = std::atoi(c);
      if (g == 1 || g == 2 || g == 4 || g == 8 || g == 16) shift_groups = g;
    }
    if (const char *c = std::getenv("SGW_SOLVER_GRAPH")) graph_mode = std::atoi(c) != 0;
    if (const char *c = std::getenv("SGW_COMPACT")) compact_mode = std::atoi(c) != 0;
    ns = nsys_;
    if (const char *c = std::getenv("SGW_GRAPH_CHUNK"))
      graph_chunk = std::max(1, std::min(16, std::atoi(c)));
    SGW_REQUIRE(nshift >= 1, SGW_E_ARG, "bicgstab: need at least one shift");
    const size_t S = nsys;
    R.alloc(S * (l + 1) * ld);
    U.alloc(S * (l + 1) * ld);
    X.alloc(S * ld);
    RT.alloc(S * ld);
    if (nsh > 0) {
      SU.alloc(S * nsh * ld);
      SX.alloc(S * nsh * ld);
    }
    const int nsnap = snap_final(l) + 1;
    RS.alloc(S * nsnap * ld);
    scal_seed.alloc(S * 8);
    gam.alloc(S * 3 * l);
    const size_t nshs = std::max<size_t>(1, S * nsh);
    scal_shift.alloc(nshs * 5 + nshs * l * 5);  // state + per-step factors
    mu.alloc(nshs * (l * (l + 1) / 2));
    coef.alloc(nshs * (kMaxL + 2));
    sigma_seed.alloc(S);
    norms.alloc(S);
    active.alloc(S);
    mr_active.alloc(S);
    iters.alloc(S);
    state.alloc(S);
    count.alloc(S);
    fresh.alloc(S);
    slot_grp.alloc(S);
    outcome.alloc(3 * S);
    double2 *p = scal_seed.p;
    sc.rho = p; sc.rho_old = p + S; sc.alpha = p + 2 * S; sc.alpha_old = p + 3 * S;
    sc.beta = p + 4 * S; sc.omega = p + 5 * S; sc.dot = p + 6 * S; sc.alpha_new = p + 7 * S;
    sc.gamma = gam.p; sc.gamma_p = gam.p + S * l; sc.gamma_pp = gam.p + 2 * S * l;
    double2 *q = scal_shift.p;
    sc.sigma = q; sc.inv_phi_old = q + nshs; sc.inv_phi = q + 2 * nshs;
    sc.inv_phi_new = q + 3 * nshs; sc.inv_theta = q + 4 * nshs;
    double2 *qs = q + 5 * nshs;
    const size_t nstep = nshs * l;
    sc.alpha_s = qs; sc.beta_s = qs + nstep; sc.fa = qs + 2 * nstep; sc.fb = qs + 3 * nstep;
    sc.ia = qs + 4 * nstep;
    sc.mu = mu.p;
    sc.coef = coef.p;
    frozen.alloc(nshs);
    SGW_HIP(hipMemset(frozen.p, 0, sizeof(int32_t) * nshs));
    sc.frozen = frozen.p;
    vv.R = R.p; vv.U = U.p; vv.X = X.p; vv.RT = RT.p; vv.SU = SU.p; vv.SX = SX.p; vv.RS = RS.p;
    vv.ld = ld; vv.l = l; vv.nsh = nsh; vv.nsnap = nsnap;
    SGW_HIP(hipHostMalloc((void **)&pinned_active, sizeof(int32_t) * 2 * S, hipHostMallocDefault));
    SGW_HIP(hipEventCreateWithFlags(&flag_ev[0], hipEventDisableTiming));
    SGW_HIP(hipEventCreateWithFlags(&flag_ev[1], hipEventDisableTiming));
  }

  void sync_active(hipStream_t st) {
    SGW_HIP(hipMemcpyAsync(h_active.data(), active.p, sizeof(int32_t) * nsys,
                           hipMemcpyDeviceToHost, st));
    SGW_HIP(hipStreamSynchronize(st));
  }

  void op_apply(Operator &op, hipStream_t st, const double2 *in, double2 *out) {
    op.apply(st, ns, in, (l + 1) * ld, out, (l + 1) * ld, sigma_seed.p, h_sigma_seed.data(),
             active.p, h_active.data(), group_dev);
  }

  // Slots in use: the leading ns of the nsys slots.  Lock-step solves on
  // device operators compact their running systems into fewer leading slots
  // as systems finish (SGW_COMPACT=0 keeps all), so that late iterations
  // launch grids for the running systems only.
  int ns = 0;
  int compact_mode = 1;
  DevBuf<int32_t> slot_grp, outcome;  // group per slot; (iters, state, active) per system

  // pairs (a, b) through pinned staging buffer b of the stream buffers
  int32_t *stage_pairs(hipStream_t st, const std::vector<std::pair<int, int>> &pr, int buf) {
    int32_t *hp = pinned_stream + 6 * nsys + (size_t)buf * 2 * nsys;
    for (size_t i = 0; i < pr.size(); ++i) {
      hp[2 * i] = pr[i].first;
      hp[2 * i + 1] = pr[i].second;
    }
    int32_t *dp = pairs_dev.p + (size_t)buf * 2 * nsys;
    SGW_HIP(hipMemcpyAsync(dp, hp, sizeof(int32_t) * 2 * pr.size(), hipMemcpyHostToDevice, st));
    return dp;
  }
  void ensure_stream_buffers() {
    if (pairs_dev.n < (size_t)8 * nsys) pairs_dev.alloc((size_t)8 * nsys);
    if (!pinned_stream)
      SGW_HIP(hipHostMalloc((void **)&pinned_stream, sizeof(int32_t) * 14 * nsys,
                            hipHostMallocDefault));
  }
  // solutions and outcomes of the systems in slots `pr` (slot, system)
  void retire(hipStream_t st, const std::vector<std::pair<int, int>> &pr, int buf, double2 *d_x,
              int64_t x_ld, int64_t nout) {
    if (pr.empty()) return;
    int32_t *dp = stage_pairs(st, pr, buf);
    k_gather_pairs<<<dim3((unsigned)((nout + kBlock - 1) / kBlock), (unsigned)pr.size(), nshift),
                     kBlock, 0, st>>>(dp, vv, d_x, x_ld, nout);
    k_retire<<<grid1((int)pr.size(), 64), 64, 0, st>>>((int)pr.size(), dp, iters.p, state.p,
                                                        active.p, outcome.p);
  }
  // move the systems of `mv` (source slot, destination slot)
  void move(hipStream_t st, const std::vector<std::pair<int, int>> &mv, int buf) {
    if (mv.empty()) return;
    int32_t *dp = stage_pairs(st, mv, buf);
    k_move_vectors<<<dim3((unsigned)((ld + kBlock - 1) / kBlock), (unsigned)mv.size()), kBlock, 0,
                     st>>>(dp, vv);
    SlotFlags fl{{active.p, mr_active.p, iters.p, state.p, count.p, fresh.p,
                  relax_thr > 0 ? relax.p : nullptr, slot_grp.p}};
    k_move_scalars<<<grid1((int)mv.size() * std::max(nsh, 1), 64), 64, 0, st>>>(
        (int)mv.size(), dp, nsys, nsh, l, sc, scal_seed.p, sigma_seed.p, fl);
  }

  const int32_t *group_dev = nullptr;

  int shift_test = 0;
  double relax_thr = 0.0;
  DevBuf<int32_t> relax;

  // ||r_0|| test on the systems of `in`; survivors go to `out`
  void check(hipStream_t st, int max_iter, int final, double thr, int32_t *in, int32_t *out) {
    k_norm<<<ns, kBlock, 0, st>>>(ns, in, R.p, (l + 1) * ld, ld, norms.p);
    k_check<<<grid1(ns, 128), 128, 0, st>>>(ns, count.p, max_iter, final, thr, in, out,
                                               norms.p, iters.p, state.p, sc, nsh, shift_test,
                                               relax_thr, relax_thr > 0 ? relax.p : nullptr);
  }

  // One outer iteration (bicgstab.f90:229-247) on every active system.  Its
  // launches do not depend on the iteration number: the `fresh` flags mark
  // systems in their first iteration and `count` their iterations, so the
  // same sequence can be captured once and replayed.
  void iterate(Operator &op, hipStream_t st, int max_iter, double threshold) {
    const int S = ns;
    const dim3 gv((unsigned)((ld + kBlock - 1) / kBlock), S);
    const int nmu = l * (l + 1) / 2, ntau = (l + 1) * (l + 2) / 2;
    const int64_t sstride = (int64_t)(l + 1) * ld;
    // ---- bicg_part (Algorithm 3)
    for (int j = 0; j < l; ++j) {
      // many slots: each seed half in one launch, one block per system; few
      // (the lock-step tail): the dot product and the update over a grid of
      // tiles as two launches, which spreads a system over more CUs.  Both
      // give the same bits.
      const bool fused = S >= kFuseHalves;
      {
        KScope ks_("bicg_v1", st, 2);
        if (fused) {
          k_bicg_half1<<<S, kBlock, 0, st>>>(j, active.p, vv, j == 0 ? 1 : 0, 1, sc,
                                             j == 0 ? fresh.p : nullptr,
                                             j == 0 ? count.p : nullptr);
        } else {
          k_dotu_s1<<<S, kBlock, 0, st>>>(active.p, R.p + j * ld, sstride, RT.p, ld,
                                          j == 0 ? 1 : 0, 1, sc, j == 0 ? fresh.p : nullptr,
                                          j == 0 ? count.p : nullptr);
          k_bicg_v1<<<gv, kBlock, 0, st>>>(j, active.p, vv, sc);
        }
      }
      op_apply(op, st, U.p + j * ld, U.p + (j + 1) * ld);       // L10
      {
        KScope ks_("bicg_seed", st, 2);
        if (fused) {
          k_bicg_half2<<<S, kBlock, 0, st>>>(j, active.p, vv, nsh, l, sc);
        } else {
          k_dotu_s2<<<S, kBlock, 0, st>>>(active.p, U.p + (j + 1) * ld, sstride, RT.p, ld, nsh,
                                          l, j, sc);
          k_bicg_seed<<<gv, kBlock, 0, st>>>(j, active.p, vv, sc);
        }
      }
      op_apply(op, st, R.p + j * ld, R.p + (j + 1) * ld);       // L24
    }
    check(st, max_iter, 0, threshold, active.p, mr_active.p);
    // ---- mr_part (Algorithm 4) on the systems still running
    {
      KScope ks_("mr_mgs", st, 2);
      k_mr_mgs<1024><<<S, 1024, 0, st>>>(mr_active.p, vv, sc, ntau, nmu);
    }
    if (nsh > 0) {  // the shifted systems: BiCG half of this iteration and, if running, MR half
      KScope ks_("shift_cycle", st, 2);
      const int ne = kBlock / shift_groups;
      const dim3 gs((unsigned)((ld + ne - 1) / ne), S);
      launch_l<ShiftCycle>(l, gs, dim3(kBlock), st, (const int32_t *)active.p,
                           (const int32_t *)mr_active.p, vv, sc, shift_groups);
    }
    {
      KScope ks_("mr_seed", st, 2);
      launch_l<MrSeed>(l, gv, dim3(kBlock), st, (const int32_t *)mr_active.p, vv, sc);
    }
    check(st, max_iter, 1, threshold, mr_active.p, active.p);
  }
};

BatchBicgstab::BatchBicgstab(int nsys, int nshift, int lmax, int64_t ld)
    : impl_(new Impl(nsys, nshift, lmax, ld)) {}
BatchBicgstab::~BatchBicgstab() { delete impl_; }

void BatchBicgstab::solve(Operator &op, hipStream_t caller_st, const double2 *d_rhs,
                          int64_t rhs_ld, const double2 *d_sigma, const double2 *h_sigma,
                          const int32_t *d_group, double threshold, int max_iter,
                          double2 *d_x, int64_t x_ld, int32_t *h_ierr, int32_t *h_iters,
                          int check_every, int shift_test, double relax_thr) {
  Impl &m = *impl_;
  m.shift_test = shift_test;
  const int S = m.nsys, l = m.l;
  // relaxed FP32 local-potential FFTs per system (plane-wave operators only)
  auto *pw = dynamic_cast<PwOperator *>(&op);
  m.relax_thr = (pw != nullptr && relax_thr > 0) ? relax_thr : 0.0;
  struct RelaxGuard {
    PwOperator *pw;
    int mode;
    const int32_t *flags;
    ~RelaxGuard() {
      if (pw) pw->set_fft_precision(mode, flags);
    }
  } guard{nullptr, 0, nullptr};
  const bool graphs = m.graphs_wanted(op);
  hipStream_t st = m.begin_on(caller_st, graphs);
  struct Join {
    Impl &m;
    hipStream_t caller, ws;
    ~Join() {
      try {
        m.end_on(caller, ws);
      } catch (...) {
      }
    }
  } join{m, caller_st, st};
  if (m.relax_thr > 0) {
    if (m.relax.n < (size_t)S) m.relax.alloc(S);
    m.relax.zero(st);
    guard = RelaxGuard{pw, pw->fft_precision(), nullptr};
    pw->set_fft_precision(2, m.relax.p);
  }
  SGW_REQUIRE(d_group != nullptr || op.ngroups() == 1, SGW_E_ARG,
              "bicgstab: a group index per system is required with several groups");
  // every group index must address an operator group; slot s starts with system s
  std::vector<int32_t> slot_grp(S, 0), slot_sys(S);
  if (d_group) {
    SGW_HIP(hipMemcpyAsync(slot_grp.data(), d_group, sizeof(int32_t) * S, hipMemcpyDeviceToHost,
                           st));
    SGW_HIP(hipStreamSynchronize(st));
  }
  for (int s = 0; s < S; ++s) {
    SGW_REQUIRE(slot_grp[s] >= 0 && slot_grp[s] < op.ngroups(), SGW_E_ARG,
                "bicgstab: group index out of range");
    slot_sys[s] = s;
  }
  m.ns = S;
  op.prepare(slot_grp.data(), S);
  m.slot_grp.upload(slot_grp.data(), S, st);
  m.group_dev = m.slot_grp.p;
  m.ensure_stream_buffers();
  m.h_sigma_seed.resize(S);
  for (int s = 0; s < S; ++s) m.h_sigma_seed[s] = h_sigma[(int64_t)s * m.nshift];
  m.sigma_seed.upload(m.h_sigma_seed.data(), S, st);
  m.h_active.assign(S, 1);
  m.active.upload(m.h_active.data(), S, st);
  m.fresh.upload(m.h_active.data(), S, st);  // every system starts its first iteration
  SGW_HIP(hipMemsetAsync(m.iters.p, 0, sizeof(int32_t) * S, st));
  SGW_HIP(hipMemsetAsync(m.state.p, 0, sizeof(int32_t) * S, st));
  SGW_HIP(hipMemsetAsync(m.count.p, 0, sizeof(int32_t) * S, st));
  m.R.zero(st);
  m.U.zero(st);
  m.X.zero(st);
  m.SU.zero(st);
  m.SX.zero(st);
  const dim3 gv((unsigned)((m.ld + kBlock - 1) / kBlock), S);
  k_init_vectors<<<gv, kBlock, 0, st>>>(S, m.vv, d_rhs, rhs_ld);
  k_init_scalars<<<grid1(S, 64), 64, 0, st>>>(S, m.nsh, l, m.sc, d_sigma, m.nshift);
  const bool host_op = op.host_side();
  const bool compact = m.compact_mode && !host_op;
  const int64_t nout = std::min<int64_t>(x_ld, m.ld);
  // every system stops by itself after max_iter iterations (k_check), so the
  // loop ends once no flag is left; the bound is a guard
  const int chunk = graphs ? m.graph_chunk : 1;
  const int64_t cap = (int64_t)max_iter + chunk + 1;
  int64_t done_iters = 0;
  bool skip_check = false;
  for (int64_t it = 1; done_iters < cap; ++it) {
    m.run(op, st, graphs, chunk, max_iter, threshold);
    done_iters += chunk;
    if (host_op) {  // host operators read the active flags on every call
      m.sync_active(st);
      bool any = false;
      for (int s = 0; s < S; ++s) any |= (m.h_active[s] != 0);
      if (!any) break;
      continue;
    }
    // Lagged test: copy this launch's flags to pinned memory and wait only
    // for the previous launch's copy, so the next iterations are already
    // queued while the host looks.  Iterations enqueued after every system
    // has stopped are no-ops (all kernels skip inactive systems).
    const int slot = (int)(it & 1);
    const int ns = m.ns;
    SGW_HIP(hipMemcpyAsync(m.pinned_active + (size_t)slot * S, m.active.p, sizeof(int32_t) * ns,
                           hipMemcpyDeviceToHost, st));
    SGW_HIP(hipEventRecord(m.flag_ev[slot], st));
    const bool last = done_iters >= cap;
    if (it < 2 && !last) continue;
    const int prev = last ? slot : (slot ^ 1);
    SGW_HIP(hipEventSynchronize(m.flag_ev[prev]));
    if (skip_check) {  // flags of the launch before a compaction: old slot order
      skip_check = false;
      continue;
    }
    const int32_t *pa = m.pinned_active + (size_t)prev * S;
    int running = 0;
    for (int s = 0; s < ns; ++s) running += pa[s] != 0;
    if (!running) break;
    if (!compact || running > ns / 2 || ns < 8) continue;
    // Fewer than half the slots run: wait for the launch just queued (its
    // flags are exact), retire the finished systems, move the running ones
    // into the leading slots and continue on those.
    SGW_HIP(hipEventSynchronize(m.flag_ev[slot]));
    const int32_t *pc = m.pinned_active + (size_t)slot * S;
    std::vector<std::pair<int, int>> done, mv;
    int keep = 0;
    for (int s = 0; s < ns; ++s) keep += pc[s] != 0;
    if (keep == 0) break;
    for (int s = 0; s < ns; ++s)
      if (!pc[s]) done.push_back({s, slot_sys[s]});
    int dst = 0;
    for (int s = keep; s < ns; ++s) {
      if (!pc[s]) continue;
      while (pc[dst]) ++dst;  // a free slot below keep (one exists for every running one above)
      mv.push_back({s, dst});
      slot_sys[dst] = slot_sys[s];
      slot_grp[dst] = slot_grp[s];
      ++dst;
    }
    m.retire(st, done, 0, d_x, x_ld, nout);
    m.move(st, mv, 1);
    m.ns = keep;
    op.prepare(slot_grp.data(), keep);
    SGW_HIP(hipStreamSynchronize(st));  // the staging buffers are free again
    m.drop_graph();  // its launches were sized for the old slot count
    skip_check = true;
  }
  (void)check_every;
  // the systems still in slots: solutions and outcomes
  {
    std::vector<std::pair<int, int>> rest;
    if (host_op)
      for (int s = 0; s < S; ++s) rest.push_back({s, s});
    else
      for (int s = 0; s < m.ns; ++s) rest.push_back({s, slot_sys[s]});
    m.retire(st, rest, 0, d_x, x_ld, nout);
  }
  std::vector<int32_t> out_h((size_t)3 * S);
  SGW_HIP(hipMemcpyAsync(out_h.data(), m.outcome.p, sizeof(int32_t) * 3 * S,
                         hipMemcpyDeviceToHost, st));
  DevBuf<int32_t> nanflag(S);
  nanflag.zero(st);
  k_nan_scan<<<S, kBlock, 0, st>>>(S, m.nshift, d_x, x_ld, nout, nanflag.p);
  std::vector<int32_t> nan_h(S);
  SGW_HIP(hipMemcpyAsync(nan_h.data(), nanflag.p, sizeof(int32_t) * S, hipMemcpyDeviceToHost, st));
  SGW_HIP(hipStreamSynchronize(st));
  for (int q = 0; q < S; ++q) {
    const int its = out_h[3 * q], sts = out_h[3 * q + 1], act = out_h[3 * q + 2];
    int ierr = 0;
    int iters_q = its;
    // still running, NaN or out of iterations: ierr 1 (bicgstab.f90:249-253)
    if (act || sts == 2 || sts == 3) {
      ierr = 1;
      if (act) iters_q = max_iter + 1;
    }
    if (nan_h[q]) ierr = 2;  // bicgstab.f90:264-267
    if (h_ierr) h_ierr[q] = ierr;
    if (h_iters) h_iters[q] = iters_q;
  }
}

void BatchBicgstab::solve_stream(Operator &op, hipStream_t caller_st, int nq, const double2 *d_rhs,
                                 int64_t rhs_ld, const double2 *d_sigma, const double2 *h_sigma,
                                 const int32_t *h_group, double threshold, int max_iter,
                                 double2 *d_x, int64_t x_ld, int32_t *h_ierr, int32_t *h_iters,
                                 int shift_test, double relax_thr) {
  Impl &m = *impl_;
  const int S = m.nsys;
  SGW_REQUIRE(!op.host_side(), SGW_E_ARG, "bicgstab stream: device-resident operators only");
  const bool graphs = m.graphs_wanted(op);
  hipStream_t st = m.begin_on(caller_st, graphs);
  struct Join {
    Impl &m;
    hipStream_t caller, ws;
    ~Join() {
      try {
        m.end_on(caller, ws);
      } catch (...) {
      }
    }
  } join{m, caller_st, st};
  SGW_REQUIRE(nq >= 0 && h_group && rhs_ld >= m.ld, SGW_E_ARG,
              "bicgstab stream: need nq >= 0, host group indices and rhs_ld >= ld");
  m.shift_test = shift_test;
  // relaxed FP32 local-potential FFTs per slot (plane-wave operators only);
  // a refilled slot starts in FP64
  auto *pw = dynamic_cast<PwOperator *>(&op);
  m.relax_thr = (pw != nullptr && relax_thr > 0) ? relax_thr : 0.0;
  struct RelaxGuard {
    PwOperator *pw;
    int mode;
    ~RelaxGuard() {
      if (pw) pw->set_fft_precision(mode, nullptr);
    }
  } guard{nullptr, 0};
  if (m.relax_thr > 0) {
    if (m.relax.n < (size_t)S) m.relax.alloc(S);
    m.relax.zero(st);
    guard = RelaxGuard{pw, pw->fft_precision()};
    pw->set_fft_precision(2, m.relax.p);
  }
  // slots per group, in proportion to the queued systems of each group
  const int ng = op.ngroups();
  std::vector<std::vector<int>> queue(ng);
  for (int q = 0; q < nq; ++q) {
    SGW_REQUIRE(h_group[q] >= 0 && h_group[q] < ng, SGW_E_ARG,
                "bicgstab stream: group index out of range");
    queue[h_group[q]].push_back(q);
  }
  int present = 0;
  for (auto &qu : queue) present += qu.empty() ? 0 : 1;
  SGW_REQUIRE(present <= S, SGW_E_SHAPE, "bicgstab stream: more groups than slots");
  std::vector<int> share(ng, 0);
  int used = 0;
  for (int g = 0; g < ng; ++g)
    if (!queue[g].empty()) {
      share[g] = std::max(1, (int)((int64_t)S * (int64_t)queue[g].size() / std::max(nq, 1)));
      share[g] = std::min<int>(share[g], (int)queue[g].size());
      used += share[g];
    }
  while (used > S) {  // trim the largest shares
    int g = (int)(std::max_element(share.begin(), share.end()) - share.begin());
    --share[g];
    --used;
  }
  std::vector<int32_t> slot_group(S, 0);
  std::vector<int> slot_q(S, -1), slot_start(S, 0), head(ng, 0);
  {
    int s = 0;
    for (int g = 0; g < ng; ++g)
      for (int k = 0; k < share[g]; ++k) slot_group[s++] = g;
    for (int t = s; t < S; ++t) slot_group[t] = slot_group[0];  // idle slots
  }
  if (m.slot_group.n < (size_t)S) m.slot_group.alloc(S);
  m.ns = S;
  if (m.pairs_dev.n < (size_t)8 * S) m.pairs_dev.alloc((size_t)8 * S);
  if (!m.pinned_stream) {
    SGW_HIP(hipHostMalloc((void **)&m.pinned_stream, sizeof(int32_t) * 14 * S,
                          hipHostMallocDefault));
  }
  int32_t *h_flags = m.pinned_stream;          // 2 x [active | state | iters] (3 S each)
  int32_t *h_pairs = m.pinned_stream + 6 * S;  // 4 x 2 S lists
  // List buffer b (0..3) is written by the host and copied to the device on
  // the stream.  At host iteration it the gather list uses b = it & 1 and the
  // refill list b = 2 + (it & 1); their previous copies were enqueued at
  // it - 2, before iteration it - 1, whose flag event the host has waited
  // for before writing them again.
  auto stage = [&](const std::vector<std::pair<int, int>> &pr, int b) -> int32_t * {
    int32_t *hp = h_pairs + (size_t)b * 2 * S;
    for (size_t i = 0; i < pr.size(); ++i) {
      hp[2 * i] = pr[i].first;
      hp[2 * i + 1] = pr[i].second;
    }
    int32_t *dp = m.pairs_dev.p + (size_t)b * 2 * S;
    SGW_HIP(hipMemcpyAsync(dp, hp, sizeof(int32_t) * 2 * pr.size(), hipMemcpyHostToDevice, st));
    return dp;
  };
  m.slot_group.upload(slot_group.data(), S, st);
  op.prepare(slot_group.data(), S);  // fixed group per slot: one plan for the whole stream
  m.group_dev = m.slot_group.p;
  m.h_sigma_seed.assign(S, make_double2(0, 0));
  std::vector<int32_t> zero(S, 0);
  m.h_active.assign(S, 1);  // host mirror unused by device operators
  m.active.upload(zero.data(), S, st);
  m.mr_active.upload(zero.data(), S, st);
  m.fresh.upload(zero.data(), S, st);
  SGW_HIP(hipMemsetAsync(m.iters.p, 0, sizeof(int32_t) * S, st));
  SGW_HIP(hipMemsetAsync(m.state.p, 0, sizeof(int32_t) * S, st));
  for (int q = 0; q < nq; ++q) {
    if (h_ierr) h_ierr[q] = 1;
    if (h_iters) h_iters[q] = max_iter + 1;
  }
  const int64_t nout = std::min<int64_t>(x_ld, m.ld);
  const unsigned tiles = (unsigned)((m.ld + kBlock - 1) / kBlock);
  const unsigned otiles = (unsigned)((nout + kBlock - 1) / kBlock);
  // enqueue refills of the listed (slot, system) pairs through list buffer b
  auto refill = [&](const std::vector<std::pair<int, int>> &pr, int b) {
    if (pr.empty()) return;
    int32_t *dp = stage(pr, b);
    k_refill_vectors<<<dim3(tiles, (unsigned)pr.size()), kBlock, 0, st>>>(dp, m.vv, d_rhs,
                                                                           rhs_ld);
    k_refill_scalars<<<grid1((int)pr.size() * std::max(m.nsh, 1), 64), 64, 0, st>>>(
        (int)pr.size(), dp, m.nsh, m.l, m.sc, d_sigma, m.nshift, m.sigma_seed.p, m.active.p,
        m.iters.p, m.state.p, m.fresh.p, m.count.p, m.relax_thr > 0 ? m.relax.p : nullptr);
  };
  // first fill
  std::vector<std::pair<int, int>> pr;
  for (int s = 0; s < used; ++s) {
    const int g = slot_group[s];
    if (head[g] < (int)queue[g].size()) {
      slot_q[s] = queue[g][head[g]++];
      slot_start[s] = 1;
      pr.push_back({s, slot_q[s]});
    }
  }
  refill(pr, 3);  // list buffer 3 is next written at host iteration 3
  int running = (int)pr.size();
  for (int it = 1; running > 0; ++it) {
    m.run(op, st, graphs, 1, max_iter, threshold);
    const int slot = it & 1;
    int32_t *hf = h_flags + (size_t)slot * 3 * S;
    SGW_HIP(hipMemcpyAsync(hf, m.active.p, sizeof(int32_t) * S, hipMemcpyDeviceToHost, st));
    SGW_HIP(hipMemcpyAsync(hf + S, m.state.p, sizeof(int32_t) * S, hipMemcpyDeviceToHost, st));
    SGW_HIP(hipMemcpyAsync(hf + 2 * S, m.iters.p, sizeof(int32_t) * S, hipMemcpyDeviceToHost,
                           st));
    SGW_HIP(hipEventRecord(m.flag_ev[slot], st));
    // look at the previous iteration's flags while this one runs (at it = 1
    // there is none; a final pass below drains the last iteration)
    const int p = it - 1;
    if (p < 1) continue;
    SGW_HIP(hipEventSynchronize(m.flag_ev[slot ^ 1]));
    const int32_t *fa = h_flags + (size_t)(slot ^ 1) * 3 * S;
    std::vector<std::pair<int, int>> done, fill;
    for (int s = 0; s < S; ++s) {
      if (slot_q[s] < 0 || slot_start[s] > p) continue;
      const int q = slot_q[s];
      if (!fa[s]) {  // converged, NaN or out of iterations (ierr 1) at iteration <= p
        if (h_iters) h_iters[q] = fa[2 * S + s];
        if (h_ierr) h_ierr[q] = fa[S + s] == 1 ? 0 : 1;
        done.push_back({s, q});
      }
    }
    if (done.empty()) continue;
    {  // solutions of the finished systems
      int32_t *dp = stage(done, slot);
      k_gather_pairs<<<dim3(otiles, (unsigned)done.size(), m.nshift), kBlock, 0, st>>>(
          dp, m.vv, d_x, x_ld, nout);
    }
    for (auto &d : done) {
      const int s = d.first, g = slot_group[s];
      if (head[g] < (int)queue[g].size()) {
        slot_q[s] = queue[g][head[g]++];
        slot_start[s] = it + 1;
        fill.push_back({s, slot_q[s]});
      } else {
        slot_q[s] = -1;
        --running;
      }
    }
    refill(fill, 2 + slot);
  }
  // NaN scan of every solution (bicgstab.f90:264-267)
  if (nq > 0) {
    DevBuf<int32_t> nanflag(nq);
    nanflag.zero(st);
    k_nan_scan<<<nq, kBlock, 0, st>>>(nq, m.nshift, d_x, x_ld, nout, nanflag.p);
    std::vector<int32_t> nan_h(nq);
    SGW_HIP(hipMemcpyAsync(nan_h.data(), nanflag.p, sizeof(int32_t) * nq, hipMemcpyDeviceToHost,
                           st));
    SGW_HIP(hipStreamSynchronize(st));
    for (int q = 0; q < nq; ++q)
      if (nan_h[q] && h_ierr) h_ierr[q] = 2;
  } else {
    SGW_HIP(hipStreamSynchronize(st));
  }
}

}  // namespace sgw
