// Linear operators applied by the solvers: (A + sigma) x for batches of vectors.
//
//  * DenseOperator    A given as a dense column-major matrix (the reference's
//                     unit-test operator, algo/linear_solver/test/linear_solver.pf:100-113)
//  * CallbackOperator the reference's procedure argument AA(sigma, xx, Ax)
//                     (algo/linear_solver/src/select_solver.f90:77-87); runs
//                     on the host, vectors cross PCIe per call
//  * PwOperator       the plane-wave Hamiltonian of linear_op
//                     (algo/linear_solver/src/linear_op.f90:46-206):
//                     (|k+q+G|^2 + V_loc + sum_p |p> D_p <p| + sigma) x,
//                     where the projector list holds the nonlocal beta's and,
//                     with D = alpha_pv, the valence states of P_v.
#include "operators.h"
#include "ktimer.h"
#include "regfft.h"
#include "splitfft.h"
#include "tune.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>

// The fixed work area of PwOperator::apply_noalloc: one copy per device,
// held until the process ends.  Host code never uses this name as an address
// (on the host it is only a stand-in for the symbol); it takes the device
// address from hipGetSymbolAddress once per device (noalloc_scratch below).
static __device__ __align__(256) unsigned char scratch[SGW_NOALLOC_SCRATCH_BYTES];

namespace sgw {

namespace {

constexpr int kBlock = 256;

// ---------------------------------------------------------------------------
// dense
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_dense(int64_t n, const double2 *__restrict__ a,
                                                  const double2 *__restrict__ in, int64_t sin,
                                                  double2 *__restrict__ out, int64_t sout,
                                                  int64_t ld, const double2 *__restrict__ sigma,
                                                  const int32_t *__restrict__ active) {
  const int v = blockIdx.y;
  if (active && !active[v]) return;
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= ld) return;
  const double2 *x = in + v * sin;
  double2 acc = mk(0, 0);
  if (i < n) {
    for (int64_t j = 0; j < n; ++j) cmac(acc, a[i + j * n], x[j]);
    cmac(acc, sigma[v], x[i]);
  }
  out[v * sout + i] = acc;
}

// ---------------------------------------------------------------------------
// plane waves
// ---------------------------------------------------------------------------
struct PwDev {
  int n1, n2, n3;
  int M1, M2;           // column grid half widths; nc1 = 2M1+1, nc2 = 2M2+1
  int nc1, nc2, nslot;
  int64_t ld;
  int nproj;
  const double *vloc;          // [n3][n2][n1]
  const double *vloc_zxy;      // [n3][n1][n2]: lanes of the x stage (one per y) read rows
  const int32_t *npw;          // [ngroup]
  const int32_t *col_start;    // [ngroup][nslot]
  const int32_t *col_cnt;      // [ngroup][nslot]
  const int32_t *sorted_idx;   // [ngroup][ld]
  const int32_t *sorted_m3;    // [ngroup][ld] (reduced mod n3)
  const int32_t *slot_of;      // [ngroup][ld]
  const int32_t *m3_of;        // [ngroup][ld] (reduced mod n3)
  const int32_t *col_lo;       // [ngroup][nslot] smallest m3 of the column
  // column layout of the split z / plane transforms' T and S: ncs columns per
  // plane; packed column j of group g is grid slot col_info[g][j].x (-1:
  // none), with that column's count, start and lowest m3 in .y, .z, .w;
  // row a of the grid holds columns m2 = row_lo .. row_lo + row_cnt - 1 at
  // j = row_start[g][a] + m2 - row_lo (row_start[g][nc1]: the group's count)
  int ncs;
  const int4 *col_info;        // [ngroup][ncs]: (slot, col_cnt, col_start, col_lo)
  const int32_t *row_start;    // [ngroup][nc1 + 1]
  const int32_t *row_lo;       // [ngroup][nc1]
  const int32_t *row_cnt;      // [ngroup][nc1]
  const double *kin;           // [ngroup][ld]
  const double2 *proj;         // [ngroup][nproj][ld]
  const double *dcoef;         // [ngroup][nproj]
  const double2 *tw1, *tw2, *tw3;  // e^{+2 pi i k/n}
  double inv_nbox;
};

// One block per (plane z, vector v): pencil DFTs of one z-plane.
// Output S[v][slot][z]: the plane's contribution to sum_z (...) before the
// final z-transform.
__global__ __launch_bounds__(kBlock) void k_pw_plane(PwDev d, const double2 *__restrict__ in,
                                                     int64_t sin, const int32_t *__restrict__ group,
                                                     const int32_t *__restrict__ active,
                                                     double2 *__restrict__ S) {
  extern __shared__ double2 smem[];
  const int z = blockIdx.x, v = blockIdx.y;
  if (active && !active[v]) return;
  const int g = group ? group[v] : 0;
  const int n1 = d.n1, n2 = d.n2, n3 = d.n3, nc1 = d.nc1, nc2 = d.nc2, nslot = d.nslot;
  double2 *w1 = smem;                 // n1
  double2 *w2 = w1 + n1;              // n2
  double2 *pin = w2 + n2;             // nslot   (m1, m2)
  double2 *A = pin + nslot;           // nc1 * n2 (m1, y)
  double2 *B = A + nc1 * n2;          // n2 * n1  (y, x)
  for (int i = threadIdx.x; i < n1; i += kBlock) w1[i] = d.tw1[i];
  for (int i = threadIdx.x; i < n2; i += kBlock) w2[i] = d.tw2[i];
  const double2 *x = in + v * sin;
  const int32_t *cs = d.col_start + (int64_t)g * nslot, *cc = d.col_cnt + (int64_t)g * nslot;
  const int32_t *sidx = d.sorted_idx + (int64_t)g * d.ld, *sm3 = d.sorted_m3 + (int64_t)g * d.ld;
  // z inverse transform of each column at this z
  for (int c = threadIdx.x; c < nslot; c += kBlock) {
    double2 acc = mk(0, 0);
    const int st = cs[c], cnt = cc[c];
    for (int j = st; j < st + cnt; ++j) {
      int e = (sm3[j] * z) % n3;
      cmac(acc, x[sidx[j]], d.tw3[e]);
    }
    pin[c] = acc;
  }
  __syncthreads();
  // y inverse: A[m1][y] = sum_m2 pin[m1][m2] e^{+i 2pi m2 y/n2}
  for (int idx = threadIdx.x; idx < nc1 * n2; idx += kBlock) {
    const int a = idx / n2, y = idx % n2;
    double2 acc = mk(0, 0);
    for (int b = 0; b < nc2; ++b) {
      int m2 = b - d.M2;
      int e = ((m2 * y) % n2 + n2) % n2;
      cmac(acc, pin[a * nc2 + b], w2[e]);
    }
    A[idx] = acc;
  }
  __syncthreads();
  // x inverse and local potential: B[y][x] = V(x,y,z) sum_m1 A[m1][y] e^{+i 2pi m1 x/n1}
  const double *vz = d.vloc + (int64_t)z * n1 * n2;
  for (int idx = threadIdx.x; idx < n2 * n1; idx += kBlock) {
    const int y = idx / n1, xx = idx % n1;
    double2 acc = mk(0, 0);
    for (int a = 0; a < nc1; ++a) {
      int m1 = a - d.M1;
      int e = ((m1 * xx) % n1 + n1) % n1;
      cmac(acc, A[a * n2 + y], w1[e]);
    }
    const double vv = vz[idx];
    B[idx] = mk(acc.x * vv, acc.y * vv);
  }
  __syncthreads();
  // x forward: A[m1][y] = sum_x B[y][x] e^{-i 2pi m1 x/n1}
  for (int idx = threadIdx.x; idx < nc1 * n2; idx += kBlock) {
    const int a = idx / n2, y = idx % n2;
    const int m1 = a - d.M1;
    double2 acc = mk(0, 0);
    for (int xx = 0; xx < n1; ++xx) {
      int e = ((m1 * xx) % n1 + n1) % n1;
      cmac(acc, B[y * n1 + xx], cconj(w1[e]));
    }
    A[idx] = acc;
  }
  __syncthreads();
  // y forward into the sphere columns
  double2 *Sv = S + (int64_t)v * nslot * n3;
  for (int c = threadIdx.x; c < nslot; c += kBlock) {
    if (cc[c] == 0) continue;
    const int a = c / nc2, b = c % nc2;
    const int m2 = b - d.M2;
    double2 acc = mk(0, 0);
    for (int y = 0; y < n2; ++y) {
      int e = ((m2 * y) % n2 + n2) % n2;
      cmac(acc, A[a * n2 + y], cconj(w2[e]));
    }
    Sv[(int64_t)c * n3 + z] = acc;
  }
}

// The local-potential FFT pipeline runs in the element type C: double2, or
// float2 for vectors whose solve has been relaxed to FP32 (input and output
// stay FP64).  With a flag array, a kernel instance handles the vectors whose
// flag relax[v] equals `want`; without one, every vector.  The FP32 scratch
// of vector v sits in the first half of its FP64 slot.
template <typename C> __device__ inline C to_c(double2 v);
template <> __device__ inline double2 to_c<double2>(double2 v) { return v; }
template <> __device__ inline float2 to_c<float2>(double2 v) {
  return make_float2((float)v.x, (float)v.y);
}
template <typename C> __device__ inline C czero() { return CplxT<C>::make(0, 0); }
template <typename C> constexpr int kSlotScale = (int)(sizeof(double2) / sizeof(C));

__device__ inline bool precision_skip(const int32_t *__restrict__ relax, int want, int v) {
  return relax != nullptr && relax[v] != want;
}
// RL false: a kernel instance that never reads relax flags, so it compiles to
// the same code (and rounds the same) as one without them
template <bool RL>
__device__ inline bool precision_skip(const int32_t *__restrict__ relax, int want, int v) {
  return RL && precision_skip(relax, want, v);
}

// dst[c] = src[c], c < n, by one wavefront: 16 loads per lane are issued
// before the first LDS store, so a plane of up to 1024 column values costs
// one memory latency instead of one per 64 values.
template <typename C>
__device__ inline void stage_plane(const C *__restrict__ src, C *dst, int n) {
  constexpr int kB = 16;
  const int lane = threadIdx.x & 63;
  for (int c0 = 0; c0 < n; c0 += 64 * kB) {
    C t[kB];
#pragma unroll
    for (int b = 0; b < kB; ++b) {
      const int c = c0 + b * 64 + lane;
      t[b] = c < n ? src[c] : CplxT<C>::make(0, 0);
    }
#pragma unroll
    for (int b = 0; b < kB; ++b) {
      const int c = c0 + b * 64 + lane;
      if (c < n) dst[c] = t[b];
    }
  }
}

// z inverse transform of the sphere columns: T[v][z][c] = sum_m3 x(c, m3)
// e^{+2 pi i m3 z / N}.  One lane per column (m1, m2); the column's plane
// waves are consecutive in sorted order with m3 = lo .. lo + cnt - 1 (the
// sphere is convex).
template <int N, typename C>
__device__ inline void zinv_body(const PwDev &d, const double2 *__restrict__ in, int64_t sin,
                                 const int32_t *__restrict__ group,
                                 const int32_t *__restrict__ active, C *__restrict__ T,
                                 const int32_t *__restrict__ relax, int want, int v, int c);

template <int N, typename C = double2>
__global__ __launch_bounds__(64) void k_pw_zinv(PwDev d, const double2 *__restrict__ in,
                                                int64_t sin, const int32_t *__restrict__ group,
                                                const int32_t *__restrict__ active,
                                                C *__restrict__ T,
                                                const int32_t *__restrict__ relax = nullptr,
                                                int want = 0) {
  zinv_body<N, C>(d, in, sin, group, active, T, relax, want, blockIdx.y,
                  blockIdx.x * 64 + threadIdx.x);
}

template <int N, typename C>
__device__ inline void zinv_body(const PwDev &d, const double2 *__restrict__ in, int64_t sin,
                                 const int32_t *__restrict__ group,
                                 const int32_t *__restrict__ active, C *__restrict__ T,
                                 const int32_t *__restrict__ relax, int want, int v, int c) {
  if (active && !active[v]) return;
  if (precision_skip(relax, want, v)) return;
  if (c >= d.nslot) return;
  const int g = group ? group[v] : 0;
  const int cnt = d.col_cnt[(int64_t)g * d.nslot + c];
  C *Tv = T + (int64_t)v * N * d.nslot * kSlotScale<C> + c;
  if (cnt == 0) {
#pragma unroll
    for (int z = 0; z < N; ++z) Tv[(int64_t)z * d.nslot] = czero<C>();
    return;
  }
  const int st = d.col_start[(int64_t)g * d.nslot + c], lo = d.col_lo[(int64_t)g * d.nslot + c];
  const int32_t *sidx = d.sorted_idx + (int64_t)g * d.ld + st;
  const double2 *x = in + v * sin;
  C buf[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const int m = i < N / 2 ? i : i - N;
    const int k = m - lo;
    buf[i] = (k >= 0 && k < cnt) ? to_c<C>(x[sidx[k]]) : czero<C>();
  }
  reg_fft<N, +1>(buf);
#pragma unroll
  for (int z = 0; z < N; ++z) Tv[(int64_t)z * d.nslot] = buf[z];
}

// z forward transform of the sphere columns back to the plane waves:
// loc[v][i] = (1/Nbox) sum_z S[v][z][c] e^{-2 pi i m3 z / N} for the plane
// waves i of column c.
template <int N, typename C = double2>
__global__ __launch_bounds__(64) void k_pw_zfwd(PwDev d, const C *__restrict__ S,
                                                const int32_t *__restrict__ group,
                                                const int32_t *__restrict__ active,
                                                double2 *__restrict__ loc,
                                                const int32_t *__restrict__ relax = nullptr,
                                                int want = 0) {
  const int v = blockIdx.y, c = blockIdx.x * 64 + threadIdx.x;
  if (active && !active[v]) return;
  if (precision_skip(relax, want, v)) return;
  if (c >= d.nslot) return;
  const int g = group ? group[v] : 0;
  const int cnt = d.col_cnt[(int64_t)g * d.nslot + c];
  if (cnt == 0) return;
  const int st = d.col_start[(int64_t)g * d.nslot + c], lo = d.col_lo[(int64_t)g * d.nslot + c];
  const int32_t *sidx = d.sorted_idx + (int64_t)g * d.ld + st;
  const C *Sv = S + (int64_t)v * N * d.nslot * kSlotScale<C> + c;
  C buf[N];
#pragma unroll
  for (int z = 0; z < N; ++z) buf[z] = Sv[(int64_t)z * d.nslot];
  reg_fft<N, -1>(buf);
  double2 *lv = loc + (int64_t)v * d.ld;
  const double sc = d.inv_nbox;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const int m = i < N / 2 ? i : i - N;
    const int k = m - lo;
    if (k >= 0 && k < cnt) lv[sidx[k]] = mk((double)buf[i].x * sc, (double)buf[i].y * sc);
  }
}

// z transforms with each column split over a lane pair (splitfft.h): lane
// 2j + odd of the pair of column j holds N/2 of its values, so these need
// about half the registers of k_pw_zinv / k_pw_zfwd.  Columns are counted in
// the layout of d.col_info (T[v][z][j], d.ncs columns per plane).  Both lanes
// of a pair take the same early exits.  MB >= 0: the sphere's z half-width
// M3 as a compile-time constant, so values outside |m3| <= MB are literal
// zeros (inverse) or never stored (forward); -1: no such bound.
template <int N, int MB, typename C = double2, bool RL = false>
__device__ inline void zinv_split_body(const PwDev &d, const double2 *__restrict__ in,
                                       int64_t sin, const int32_t *__restrict__ group,
                                       const int32_t *__restrict__ active, C *__restrict__ T,
                                       int v, int jc, bool odd,
                                       const int32_t *__restrict__ relax = nullptr,
                                       int want = 0) {
  constexpr int H = N / 2, Q = N / 4;
  if (active && !active[v]) return;
  if (precision_skip<RL>(relax, want, v)) return;
  const int ncs = d.ncs;
  if (jc >= ncs) return;
  const int g = group ? group[v] : 0;
  const int4 ci = d.col_info[(int64_t)g * ncs + jc];  // one load for the column's tables
  if (ci.x < 0) return;  // past this group's columns
  const int cnt = ci.y;
  C *Tv = T + (int64_t)v * N * ncs * kSlotScale<C> + jc;
  const int zo = odd ? Q : 0;  // this lane's z: zo + j and H + zo + j
  if (cnt == 0) {
#pragma unroll
    for (int j = 0; j < Q; ++j) {
      Tv[(int64_t)(zo + j) * ncs] = czero<C>();
      Tv[(int64_t)(H + zo + j) * ncs] = czero<C>();
    }
    return;
  }
  const int st = ci.z, lo = ci.w;
  const int32_t *sidx = d.sorted_idx + (int64_t)g * d.ld + st;
  const double2 *x = in + v * sin;
  C b[H];
#pragma unroll
  for (int n = 0; n < H; ++n) {
    const int i = 2 * n + (odd ? 1 : 0), m = i < N / 2 ? i : i - N, k = m - lo;
    const bool band = MB < 0 || (m >= -MB && m <= MB);
    b[n] = (band && k >= 0 && k < cnt) ? to_c<C>(x[sidx[k]]) : czero<C>();
  }
  split_dit<N, +1>(b, odd);
#pragma unroll
  for (int j = 0; j < Q; ++j) {
    Tv[(int64_t)(zo + j) * ncs] = b[j];
    Tv[(int64_t)(H + zo + j) * ncs] = b[Q + j];
  }
}

template <int N, int MB, typename C = double2, bool RL = false>
__global__ __launch_bounds__(64) void k_pw_zfwd_split(PwDev d, const C *__restrict__ S,
                                                      const int32_t *__restrict__ group,
                                                      const int32_t *__restrict__ active,
                                                      double2 *__restrict__ loc,
                                                      const int32_t *__restrict__ relax = nullptr,
                                                      int want = 0) {
  constexpr int H = N / 2, Q = N / 4;
  const int v = blockIdx.y, jc = blockIdx.x * 32 + (int)(threadIdx.x >> 1);
  const bool odd = (threadIdx.x & 1) != 0;
  if (active && !active[v]) return;
  if (precision_skip<RL>(relax, want, v)) return;
  const int ncs = d.ncs;
  if (jc >= ncs) return;
  const int g = group ? group[v] : 0;
  const int4 ci = d.col_info[(int64_t)g * ncs + jc];
  if (ci.x < 0) return;  // past this group's columns
  const int cnt = ci.y;
  if (cnt == 0) return;
  const int st = ci.z, lo = ci.w;
  const int32_t *sidx = d.sorted_idx + (int64_t)g * d.ld + st;
  const C *Sv = S + (int64_t)v * N * ncs * kSlotScale<C> + jc;
  const int zo = odd ? Q : 0;
  C b[H];
#pragma unroll
  for (int j = 0; j < Q; ++j) {
    b[j] = Sv[(int64_t)(zo + j) * ncs];
    b[Q + j] = Sv[(int64_t)(H + zo + j) * ncs];
  }
  split_dif<N, -1>(b, odd);
  double2 *lv = loc + (int64_t)v * d.ld;
  const double sc = d.inv_nbox;
#pragma unroll
  for (int k = 0; k < H; ++k) {
    const int i = 2 * k + (odd ? 1 : 0), m = i < N / 2 ? i : i - N, k3 = m - lo;
    const bool band = MB < 0 || (m >= -MB && m <= MB);
    if (band && k3 >= 0 && k3 < cnt) lv[sidx[k3]] = mk((double)b[k].x * sc, (double)b[k].y * sc);
  }
}

// One wavefront per (plane z, vector v), cubic box of compile-time side N.
// The plane's pencil transforms run as register FFTs (regfft.h):
//   Z  each lane sums its sphere columns at this z (<= 2M3+1 terms),
//   Y  lane a < nc1 transforms row m1 = a - M1 over m2 -> y,
//   X  lane y < N transforms over m1 -> x, multiplies by V(x, y, z) and
//      transforms back, keeping the nc1 frequencies m1 of the sphere,
//   Y' lane a < nc1 transforms back over y -> m2 and writes the sphere
//      columns' values S[v][c][z].
// LDS holds the column values T[nc1][nc2] and one row buffer P[nc1][N] that
// the X stage overwrites column by column (each lane owns its column y).
template <int N, typename C>
__device__ inline void plane_body(const PwDev &d, const C *__restrict__ Tin,
                                  const int32_t *__restrict__ group,
                                  const int32_t *__restrict__ active, C *__restrict__ S,
                                  const int32_t *__restrict__ relax, int want, int z, int v,
                                  C *smem);

template <int N, typename C = double2>
__global__ __launch_bounds__(64) void k_pw_plane_fft(PwDev d, const C *__restrict__ Tin,
                                                     const int32_t *__restrict__ group,
                                                     const int32_t *__restrict__ active,
                                                     C *__restrict__ S,
                                                     const int32_t *__restrict__ relax = nullptr,
                                                     int want = 0) {
  extern __shared__ double2 smem_d[];
  plane_body<N, C>(d, Tin, group, active, S, relax, want, blockIdx.x, blockIdx.y,
                   reinterpret_cast<C *>(smem_d));
}

template <int N, typename C>
__device__ inline void plane_body(const PwDev &d, const C *__restrict__ Tin,
                                  const int32_t *__restrict__ group,
                                  const int32_t *__restrict__ active, C *__restrict__ S,
                                  const int32_t *__restrict__ relax, int want, int z, int v,
                                  C *smem) {
  using R = typename CplxT<C>::S;
  if (active && !active[v]) return;
  if (precision_skip(relax, want, v)) return;
  const int g = group ? group[v] : 0;
  const int nc1 = d.nc1, nc2 = d.nc2, nslot = d.nslot, M1 = d.M1, M2 = d.M2;
  const int lane = threadIdx.x;
  C *T = smem;          // nslot
  C *P = T + nslot;     // nc1 * N
  const int32_t *cc = d.col_cnt + (int64_t)g * nslot;
  const C *Tz = Tin + ((int64_t)v * N * kSlotScale<C> + z) * nslot;
  stage_plane(Tz, T, nslot);
  __syncthreads();
  C buf[N];
  if (lane < nc1) {  // Y inverse of row a
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const int m = i < N / 2 ? i : i - N;
      buf[i] = (m >= -M2 && m <= M2) ? T[lane * nc2 + m + M2] : czero<C>();
    }
    reg_fft<N, +1>(buf);
#pragma unroll
    for (int i = 0; i < N; ++i) P[lane * N + i] = buf[i];
  }
  __syncthreads();
  if (lane < N) {  // X inverse, V(r), X forward of column y = lane
    const int y = lane;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const int m = i < N / 2 ? i : i - N;
      buf[i] = (m >= -M1 && m <= M1) ? P[(m + M1) * N + y] : czero<C>();
    }
    reg_fft<N, +1>(buf);
    const double *vcol = d.vloc_zxy + (int64_t)z * N * N + y;  // V(x, y, z) at vcol[x * N]
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const R vv = (R)vcol[i * N];
      buf[i] = CplxT<C>::make(buf[i].x * vv, buf[i].y * vv);
    }
    reg_fft<N, -1>(buf);
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const int m = i < N / 2 ? i : i - N;
      if (m >= -M1 && m <= M1) P[(m + M1) * N + y] = buf[i];
    }
  }
  __syncthreads();
  if (lane < nc1) {  // Y forward of row a into the sphere columns
#pragma unroll
    for (int i = 0; i < N; ++i) buf[i] = P[lane * N + i];
    reg_fft<N, -1>(buf);
    C *Sz = S + ((int64_t)v * N * kSlotScale<C> + z) * nslot;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const int m = i < N / 2 ? i : i - N;
      if (m >= -M2 && m <= M2) {
        const int c = lane * nc2 + m + M2;
        if (cc[c] > 0) Sz[c] = buf[i];
      }
    }
  }
}

// FP64 plane transforms of two vectors per wavefront, planes (z, v0) and
// (z, v0 + 1): the Y stages of both planes run side by side (plane p on lanes
// [p nc1, (p + 1) nc1), which needs 2 nc1 <= 64), the X stage plane by plane on
// lanes [0, N).  The Y stages use only nc1 lanes each, so pairing the planes
// halves their share of the wavefront's time, and the loads of both planes are
// in flight together.  LDS: T[2][nslot], then P[2][nc1][N].
template <int N>
__device__ inline void plane2_body(const PwDev &d, const double2 *__restrict__ Tin,
                                   const int32_t *__restrict__ group,
                                   const int32_t *__restrict__ active, double2 *__restrict__ S,
                                   int z, int v0, int nvec, double2 *smem) {
  const int nc1 = d.nc1, nc2 = d.nc2, nslot = d.nslot, M2 = d.M2;
  const int lane = threadIdx.x;
  bool live[2];
#pragma unroll
  for (int p = 0; p < 2; ++p) live[p] = v0 + p < nvec && (!active || active[v0 + p]);
  if (!live[0] && !live[1]) return;
  double2 *T = smem, *P = smem + 2 * nslot;
#pragma unroll
  for (int p = 0; p < 2; ++p)
    if (live[p]) stage_plane(Tin + ((int64_t)(v0 + p) * N + z) * nslot, T + p * nslot, nslot);
  __syncthreads();
  // Y-stage lane mapping: plane yp, row a
  const int yp = lane < nc1 ? 0 : 1, a = lane - yp * nc1;
  const bool yrow = a < nc1 && live[yp];
  double2 buf[N];
  if (yrow) {  // Y inverse of row a
    const double2 *Tp = T + yp * nslot;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const int m = i < N / 2 ? i : i - N;
      buf[i] = (m >= -M2 && m <= M2) ? Tp[a * nc2 + m + M2] : mk(0, 0);
    }
    reg_fft<N, +1>(buf);
    double2 *Pp = P + yp * nc1 * N;
#pragma unroll
    for (int i = 0; i < N; ++i) Pp[a * N + i] = buf[i];
  }
  __syncthreads();
  if (lane < N) {  // X inverse, V(r), X forward of column y = lane, plane by plane
#pragma unroll 1
    for (int p = 0; p < 2; ++p) {
      if (!live[p]) continue;
      // Per-plane copies the compiler cannot see through: V loads, LDS
      // addresses and band masks are then formed inside the loop.  Hoisted
      // out of it they would stay live across both planes (2N loaded values,
      // N addresses, 2N mask bits) and spill.
      int y = lane, M1 = d.M1;
      __asm__ volatile("" : "+v"(y), "+s"(M1)::"memory");
      const double *vcol = d.vloc_zxy + (int64_t)z * N * N + y;  // V(x, y, z) at vcol[x * N]
      double2 *Pp = P + p * nc1 * N;
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const int m = i < N / 2 ? i : i - N;
        buf[i] = (m >= -M1 && m <= M1) ? Pp[(m + M1) * N + y] : mk(0, 0);
      }
      reg_fft<N, +1>(buf);
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const double vv = vcol[i * N];
        buf[i] = mk(buf[i].x * vv, buf[i].y * vv);
      }
      reg_fft<N, -1>(buf);
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const int m = i < N / 2 ? i : i - N;
        if (m >= -M1 && m <= M1) Pp[(m + M1) * N + y] = buf[i];
      }
    }
  }
  __syncthreads();
  if (yrow) {  // Y forward of row a into the sphere columns
    const int v = v0 + yp;
    const double2 *Pp = P + yp * nc1 * N;
#pragma unroll
    for (int i = 0; i < N; ++i) buf[i] = Pp[a * N + i];
    reg_fft<N, -1>(buf);
    const int32_t *cc = d.col_cnt + (int64_t)(group ? group[v] : 0) * nslot;
    double2 *Sz = S + ((int64_t)v * N + z) * nslot;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const int m = i < N / 2 ? i : i - N;
      if (m >= -M2 && m <= M2) {
        const int c = a * nc2 + m + M2;
        if (cc[c] > 0) Sz[c] = buf[i];
      }
    }
  }
}

// Plane transforms of PL vectors, planes (z, v0 + p) for p < PL, by a
// block of 64 W lanes whose transforms are split over lane pairs
// (splitfft.h): every lane holds N/2 values, so the block needs about half
// the registers per lane of plane2_body and each SIMD holds two wavefronts.
// T and S are in the column layout of d.row_* ([v][z][j], d.ncs columns per
// plane; row a of the grid holds its columns m2 = lo .. hi - 1).  The split
// transforms are in rotation form (fewer per-lane selects than the select
// form, 6.5 % less time here; the two forms round differently).
//   Y  lane pair r < PL nc1: row a of plane yp (r = yp nc1 + a), m2 -> y;
//   X  the PL N (plane, column) tasks in rounds of 32 W lane pairs: column y
//      of plane p, m1 -> x, times V(x, y, z), x -> m1 (nc1 frequencies kept);
//   Y' as Y, y -> m2, into the row's columns of S.
// With PL N a multiple of 32 W (4 planes of 48 on 3 wavefronts) every lane
// pair works in every X round.  LDS holds P[PL][nc1][N + 1] (rows padded by
// one value, so that the lanes of the Y stages, one row each, fall on
// different banks); the column values T[PL][ncs] are staged into its start
// and read into registers before the Y stage overwrites them.
// MB >= 0: the sphere's half-widths M1 = M2 = MB are compile-time constants.
// Every transform input outside |m| <= MB is then a literal zero and every
// output outside it is never stored, so the compiler drops the butterflies
// that only touch those (the file is built without signed zeros and
// non-finite values, which lets x + 0 and x * 0 fold).  MB = -1: run-time
// half-widths.
template <int N, int MB, int PL, int W, typename C = double2, bool RL = false>
__device__ inline void plane_split_body(const PwDev &d, const C *__restrict__ Tin,
                                        const int32_t *__restrict__ group,
                                        const int32_t *__restrict__ active,
                                        C *__restrict__ S, int z, int v0, int nvec, C *smem,
                                        const int32_t *__restrict__ relax = nullptr,
                                        int want = 0,
                                        const int32_t *__restrict__ vlist = nullptr) {
  using R = typename CplxT<C>::S;
  constexpr int H = N / 2, Q = N / 4, NP = 32 * W;  // NP lane pairs
  constexpr int PN = N + 1;  // row stride of P: rows of N values would share LDS banks
  const int nc1 = MB >= 0 ? 2 * MB + 1 : d.nc1;
  const int ncs = d.ncs, M2 = MB >= 0 ? MB : d.M2;
  const int t = threadIdx.x, r = t >> 1;
  const bool odd = (t & 1) != 0;
  // plane p of the block belongs to vector vlist[v0 + p] (the running
  // vectors in order, then -1) or, without a list, to vector v0 + p
  int vv[PL];
  bool live[PL], any = false;
#pragma unroll
  for (int p = 0; p < PL; ++p) {
    const int i = v0 + p;
    vv[p] = i < nvec ? (vlist ? vlist[i] : i) : -1;
    live[p] = vv[p] >= 0 && (!active || active[vv[p]]) &&
              !precision_skip<RL>(relax, want, vv[p]);
    any |= live[p];
  }
  if (!any) return;
  C *P = smem, *T = smem;  // T[PL][ncs] aliases the start of P[PL][nc1][PN]
#pragma unroll
  for (int p = 0; p < PL; ++p)  // plane p staged by wavefront p mod W
    if ((t >> 6) == p % W && live[p]) {
      const int gp = group ? group[vv[p]] : 0;
      stage_plane(Tin + ((int64_t)vv[p] * N * kSlotScale<C> + z) * ncs, T + p * ncs,
                  d.row_start[(int64_t)gp * (nc1 + 1) + nc1]);
    }
  __syncthreads();
  const int yp = r / nc1, a = r - yp * nc1;
  const bool yrow = yp < PL && live[yp];
  const int vy = yrow ? (vlist ? vlist[v0 + yp] : v0 + yp) : 0;  // this lane pair's vector
  C b[H];
  if (yrow) {  // row a: columns m2 = lo .. hi - 1 of the group's layout
    const int64_t g = group ? group[vy] : 0;
    const int lo = d.row_lo[g * nc1 + a], hi = lo + d.row_cnt[g * nc1 + a];
    const C *Tp = T + yp * ncs + d.row_start[g * (nc1 + 1) + a] - lo;
#pragma unroll
    for (int n = 0; n < H; ++n) {
      const int i = 2 * n + (odd ? 1 : 0), m = i < N / 2 ? i : i - N;
      b[n] = (m >= -M2 && m <= M2 && m >= lo && m < hi) ? Tp[m] : czero<C>();
    }
  }
  __syncthreads();  // every column value is in registers before P overwrites T
  if (yrow) {  // Y inverse of row a
    split_dit<N, +1>(b, odd);
    C *Pr = P + (yp * nc1 + a) * PN + (odd ? Q : 0);
#pragma unroll
    for (int j = 0; j < Q; ++j) {
      Pr[j] = b[j];
      Pr[H + j] = b[Q + j];
    }
  }
  __syncthreads();
  // X inverse, V(r), X forward: task = (plane p, column y), round by round
#pragma unroll 1
  for (int task = r; task < PL * N; task += NP) {
    const int p = task / N;
    if (!live[p]) continue;
    // per-round copies the compiler cannot see through, so that addresses
    // and masks are formed inside the loop rather than kept live across it
    int y = task - p * N, M1 = MB >= 0 ? MB : d.M1;
    if constexpr (MB >= 0)
      __asm__ volatile("" : "+v"(y)::"memory");
    else
      __asm__ volatile("" : "+v"(y), "+s"(M1)::"memory");
    C *Pp = P + p * nc1 * PN + y;
#pragma unroll
    for (int n = 0; n < H; ++n) {
      const int i = 2 * n + (odd ? 1 : 0), m = i < N / 2 ? i : i - N;
      b[n] = (m >= -M1 && m <= M1) ? Pp[(m + M1) * PN] : czero<C>();
    }
    split_dit<N, +1>(b, odd);
    const double *vcol = d.vloc_zxy + ((int64_t)z * N + (odd ? Q : 0)) * N + y;
#pragma unroll
    for (int j = 0; j < Q; ++j) {  // x = j (+Q) and H + j (+Q)
      const R v1 = (R)vcol[j * N], v2 = (R)vcol[(H + j) * N];
      b[j] = CplxT<C>::make(b[j].x * v1, b[j].y * v1);
      b[Q + j] = CplxT<C>::make(b[Q + j].x * v2, b[Q + j].y * v2);
    }
    split_dif<N, -1>(b, odd);
#pragma unroll
    for (int k = 0; k < H; ++k) {
      const int i = 2 * k + (odd ? 1 : 0), m = i < N / 2 ? i : i - N;
      if (m >= -M1 && m <= M1) Pp[(m + M1) * PN] = b[k];
    }
  }
  __syncthreads();
  if (yrow) {  // Y forward of row a into the sphere columns
    const int v = vy;
    const C *Pr = P + (yp * nc1 + a) * PN + (odd ? Q : 0);
#pragma unroll
    for (int j = 0; j < Q; ++j) {
      b[j] = Pr[j];
      b[Q + j] = Pr[H + j];
    }
    split_dif<N, -1>(b, odd);
    const int64_t g = group ? group[v] : 0;
    const int lo = d.row_lo[g * nc1 + a], hi = lo + d.row_cnt[g * nc1 + a];
    C *Sz = S + ((int64_t)v * N * kSlotScale<C> + z) * ncs + d.row_start[g * (nc1 + 1) + a] - lo;
#pragma unroll
    for (int k = 0; k < H; ++k) {
      const int i = 2 * k + (odd ? 1 : 0), m = i < N / 2 ? i : i - N;
      if (m >= -M2 && m <= M2 && m >= lo && m < hi) Sz[m] = b[k];
    }
  }
}

// Y then X inverse transforms of one z-plane of the sphere columns Tz (in
// LDS as T) into the registers of lane y: buf[x] = value at (x, y, z).
// P (nc1 x N, LDS) is the row buffer; lanes >= N return with buf untouched.
template <int N>
__device__ inline void plane_to_real(const PwDev &d, const double2 *__restrict__ Tz, double2 *T,
                                     double2 *P, double2 (&buf)[N]) {
  const int nc1 = d.nc1, nc2 = d.nc2, nslot = d.nslot, M1 = d.M1, M2 = d.M2;
  const int lane = threadIdx.x;
  __syncthreads();
  stage_plane(Tz, T, nslot);
  __syncthreads();
  if (lane < nc1) {
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const int m = i < N / 2 ? i : i - N;
      buf[i] = (m >= -M2 && m <= M2) ? T[lane * nc2 + m + M2] : mk(0, 0);
    }
    reg_fft<N, +1>(buf);
#pragma unroll
    for (int i = 0; i < N; ++i) P[lane * N + i] = buf[i];
  }
  __syncthreads();
  if (lane < N) {
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const int m = i < N / 2 ? i : i - N;
      buf[i] = (m >= -M1 && m <= M1) ? P[(m + M1) * N + lane] : mk(0, 0);
    }
    reg_fft<N, +1>(buf);
  }
}

// Density response of one z-plane for one shift k (incdrhoscf):
//   D[k][z][x][y] (+)= wgt * sum_v conj(psi_v(r)) dpsi_{v,k}(r),
// dpsi from the z-transformed columns T (vector v * nshift + k of the block)
// and psi_v(r) given as Pr[v][z][x][y].  Block (z, k) owns its plane of D.
template <int N>
__global__ __launch_bounds__(64) void k_pw_plane_drho(PwDev d, const double2 *__restrict__ T,
                                                      int nbnd, int nshift,
                                                      const double2 *__restrict__ Pr, double wgt,
                                                      double2 *__restrict__ D, int accumulate) {
  extern __shared__ double2 smem[];
  const int z = blockIdx.x, k = blockIdx.y;
  const int lane = threadIdx.x;
  double2 *Ts = smem;              // nslot
  double2 *P = Ts + d.nslot;       // nc1 * N
  double2 *Dz = D + ((int64_t)k * N + z) * N * N;
  if (!accumulate && lane < N)
    for (int x = 0; x < N; ++x) Dz[(int64_t)x * N + lane] = mk(0, 0);
  double2 buf[N];
  for (int v = 0; v < nbnd; ++v) {
    const int64_t vec = (int64_t)v * nshift + k;
    plane_to_real<N>(d, T + (vec * N + z) * d.nslot, Ts, P, buf);
    if (lane < N) {
      const double2 *pz = Pr + ((int64_t)v * N + z) * N * N + lane;
#pragma unroll
      for (int x = 0; x < N; ++x) {
        const double2 p = pz[(int64_t)x * N];
        double2 *o = Dz + (int64_t)x * N + lane;
        double2 acc = *o;
        // conj(p) * dpsi
        acc.x = fma(wgt, p.x * buf[x].x + p.y * buf[x].y, acc.x);
        acc.y = fma(wgt, p.x * buf[x].y - p.y * buf[x].x, acc.y);
        *o = acc;
      }
    }
  }
}

// The same density response with the transforms split over lane pairs
// (splitfft.h), one block of 128 lanes per (z, k): lane pair r < N owns
// column y = r of the plane for all bands, so its 2 x N/2 values of D stay in
// registers instead of a read-modify-write of D per band.  Per band: the
// column values are staged in LDS, the Y stage (pairs r < nc1) writes the
// rows P[nc1][N] over them, the X stage (pairs r < N) transforms column y
// and accumulates conj(psi_v) dpsi.  MB as in plane_split_body.
template <int N, int MB>
__global__ __launch_bounds__(128, 2) void k_pw_plane_drho_split(PwDev d,
                                                             const double2 *__restrict__ T,
                                                             int nbnd, int nshift,
                                                             const double2 *__restrict__ Pr,
                                                             double wgt, double2 *__restrict__ D,
                                                             int accumulate) {
  constexpr int H = N / 2, Q = N / 4;
  extern __shared__ double2 smem[];  // P[nc1][N]; the column values staged at its start
  const int z = blockIdx.x, k = blockIdx.y;
  const int t = threadIdx.x, r = t >> 1;
  const bool odd = (t & 1) != 0;
  const int nc1 = MB >= 0 ? 2 * MB + 1 : d.nc1, nc2 = MB >= 0 ? 2 * MB + 1 : d.nc2;
  const int nslot = d.nslot, M1 = MB >= 0 ? MB : d.M1, M2 = MB >= 0 ? MB : d.M2;
  double2 *P = smem, *Ts = smem;
  const int xo = odd ? Q : 0;  // this lane's x: xo + j and H + xo + j
  double2 *Dz = D + ((int64_t)k * N + z) * N * N + r;
  double2 acc[H];
#pragma unroll
  for (int j = 0; j < H; ++j) acc[j] = mk(0, 0);
  if (accumulate && r < N) {
#pragma unroll
    for (int j = 0; j < Q; ++j) {
      acc[j] = Dz[(int64_t)(xo + j) * N];
      acc[Q + j] = Dz[(int64_t)(H + xo + j) * N];
    }
  }
  for (int v = 0; v < nbnd; ++v) {
    const int64_t vec = (int64_t)v * nshift + k;
    __syncthreads();  // the previous band's X stage is done with P
    if (t < 64) stage_plane(T + (vec * N + z) * nslot, Ts, nslot);
    __syncthreads();
    double2 b[H];
    if (r < nc1) {
      const double2 *Tr = Ts + r * nc2 + M2;
#pragma unroll
      for (int n = 0; n < H; ++n) {
        const int i = 2 * n + (odd ? 1 : 0), m = i < N / 2 ? i : i - N;
        b[n] = (m >= -M2 && m <= M2) ? Tr[m] : mk(0, 0);
      }
    }
    __syncthreads();  // every column value is in registers before P overwrites them
    if (r < nc1) {  // Y inverse of row r
      split_dit_sel<N, +1>(b, odd);
      double2 *Prow = P + r * N + xo;
#pragma unroll
      for (int j = 0; j < Q; ++j) {
        Prow[j] = b[j];
        Prow[H + j] = b[Q + j];
      }
    }
    __syncthreads();
    if (r < N) {  // X inverse of column y = r, times conj(psi_v(r)), into D
      const double2 *Pc = P + r;
#pragma unroll
      for (int n = 0; n < H; ++n) {
        const int i = 2 * n + (odd ? 1 : 0), m = i < N / 2 ? i : i - N;
        b[n] = (m >= -M1 && m <= M1) ? Pc[(m + M1) * N] : mk(0, 0);
      }
      split_dit_sel<N, +1>(b, odd);
      const double2 *pz = Pr + ((int64_t)v * N + z) * N * N + r;
#pragma unroll
      for (int j = 0; j < H; ++j) {
        const int x = (j < Q ? j : H + j - Q) + xo;
        const double2 p = pz[(int64_t)x * N];
        acc[j].x = fma(wgt, p.x * b[j].x + p.y * b[j].y, acc[j].x);
        acc[j].y = fma(wgt, p.x * b[j].y - p.y * b[j].x, acc[j].y);
      }
    }
  }
  if (r < N) {
#pragma unroll
    for (int j = 0; j < Q; ++j) {
      Dz[(int64_t)(xo + j) * N] = acc[j];
      Dz[(int64_t)(H + xo + j) * N] = acc[Q + j];
    }
  }
}

// Adjoint gradient of one z-plane over a strided set of vector pairs:
//   G[b][z][x][y] += sum_{p = b, b + nb, ...} w_p Re(conj(y_p(r)) x_p(r)),
// from the z-transformed columns Tx (x_p) and Ty (y_p).  Block (z, b) owns
// its plane of G exclusively; the x plane waits in LDS (Xs[x][y]) for y.
template <int N>
__global__ __launch_bounds__(64) void k_pw_plane_grad(PwDev d, const double2 *__restrict__ Tx,
                                                      const double2 *__restrict__ Ty, int np,
                                                      const double *__restrict__ w,
                                                      double *__restrict__ G, int accumulate) {
  extern __shared__ double2 smem[];
  const int z = blockIdx.x, b = blockIdx.y, nb = gridDim.y;
  const int lane = threadIdx.x;
  double2 *T = smem;               // nslot
  double2 *P = T + d.nslot;        // nc1 * N
  double2 *Xs = P + d.nc1 * N;     // N * N
  double *Gz = G + ((int64_t)b * N + z) * N * N;
  if (!accumulate && lane < N)
    for (int x = 0; x < N; ++x) Gz[(int64_t)x * N + lane] = 0.0;
  double2 buf[N];
  for (int p = b; p < np; p += nb) {
    plane_to_real<N>(d, Tx + ((int64_t)p * N + z) * d.nslot, T, P, buf);
    if (lane < N)
#pragma unroll
      for (int x = 0; x < N; ++x) Xs[x * N + lane] = buf[x];
    plane_to_real<N>(d, Ty + ((int64_t)p * N + z) * d.nslot, T, P, buf);
    if (lane < N) {
      const double wp = w[p];
#pragma unroll
      for (int x = 0; x < N; ++x) {
        const double2 a = Xs[x * N + lane];
        double *g = Gz + (int64_t)x * N + lane;
        *g = fma(wp, buf[x].x * a.x + buf[x].y * a.y, *g);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Projector contraction per k+q group: the vectors of one group share the
// projector block proj[g] (nproj x npw), so the two complex GEMMs
//   becp = proj^H X   (nproj x nv, K = npw)   and   HX += proj (D becp)
// run per (group, chunk of kChunk plane waves, tile of kVT vectors): each
// block reads its projector chunk once for kVT vectors.
// ---------------------------------------------------------------------------
constexpr int kChunk = 64;
constexpr int kVT = 16;
constexpr int kAsmChunk = 64, kAsmThreads = 256;  // k_pw_assemble2 default block shape
constexpr int kPT = 16;  // projectors staged in LDS at a time

// does any vector of the tile still run?  (uniform across the block)
__device__ inline bool tile_active(const int32_t *__restrict__ active,
                                   const int32_t *__restrict__ vecs, int nv) {
  int any = 0;
  for (int k = 0; k < nv; ++k) any |= active[vecs[k]];
  return any != 0;
}

// becp[g][vt][p][vl] = D_p sum over the chunk partials (once per tile); one
// lane per output (p, vl), 64 outputs per block
__device__ inline void becp_reduce_body(const PwDev &d, const int32_t *__restrict__ active,
                                        const int32_t *__restrict__ gvec,
                                        const int32_t *__restrict__ gcnt, int maxper, int tiles,
                                        int nchunk, const double2 *__restrict__ part,
                                        double2 *__restrict__ becp, int bx, int g, int vt, int t);

__global__ __launch_bounds__(64) void k_pw_becp_reduce(PwDev d,
                                                       const int32_t *__restrict__ active,
                                                       const int32_t *__restrict__ gvec,
                                                       const int32_t *__restrict__ gcnt,
                                                       int maxper, int tiles, int nchunk,
                                                       const double2 *__restrict__ part,
                                                       double2 *__restrict__ becp) {
  becp_reduce_body(d, active, gvec, gcnt, maxper, tiles, nchunk, part, becp, blockIdx.x,
                   blockIdx.y, blockIdx.z, threadIdx.x);
}

__device__ inline void becp_reduce_body(const PwDev &d, const int32_t *__restrict__ active,
                                        const int32_t *__restrict__ gvec,
                                        const int32_t *__restrict__ gcnt, int maxper, int tiles,
                                        int nchunk, const double2 *__restrict__ part,
                                        double2 *__restrict__ becp, int bx, int g, int vt, int t) {
  const int cnt = gcnt[g], v0 = vt * kVT;
  if (v0 >= cnt) return;
  const int nv = min(kVT, cnt - v0), nproj = d.nproj;
  const int o = bx * 64 + t;  // t: thread of the 64 handling outputs [64 bx, 64 bx + 64)
  if (o >= nproj * kVT) return;
  if (active && !tile_active(active, gvec + g * maxper + v0, nv)) return;
  const int nck = (d.npw[g] + kChunk - 1) / kChunk;
  const int64_t stride = (int64_t)nproj * kVT;
  const double2 *pp = part + ((int64_t)g * tiles + vt) * nchunk * stride + o;
  double2 a0 = mk(0, 0), a1 = mk(0, 0), a2 = mk(0, 0), a3 = mk(0, 0);
  int c = 0;
  for (; c + 4 <= nck; c += 4) {  // four independent loads in flight
    const double2 t0 = pp[(int64_t)c * stride], t1 = pp[(int64_t)(c + 1) * stride];
    const double2 t2 = pp[(int64_t)(c + 2) * stride], t3 = pp[(int64_t)(c + 3) * stride];
    a0.x += t0.x; a0.y += t0.y;
    a1.x += t1.x; a1.y += t1.y;
    a2.x += t2.x; a2.y += t2.y;
    a3.x += t3.x; a3.y += t3.y;
  }
  for (; c < nck; ++c) {
    const double2 t = pp[(int64_t)c * stride];
    a0.x += t.x; a0.y += t.y;
  }
  const double sx = (a0.x + a1.x) + (a2.x + a3.x), sy = (a0.y + a1.y) + (a2.y + a3.y);
  const double dc = d.dcoef[(int64_t)g * nproj + o / kVT];
  becp[((int64_t)g * tiles + vt) * stride + o] = mk(sx * dc, sy * dc);
}

// part[((g*tiles + vt)*nchunk + chunk)][p][vl] = sum_{i in chunk} conj(proj_p(i)) x_v(i)
__device__ inline void becp_body(const PwDev &d, const double2 *__restrict__ in, int64_t sin,
                                 const int32_t *__restrict__ active,
                                 const int32_t *__restrict__ gvec,
                                 const int32_t *__restrict__ gcnt, int maxper, int tiles,
                                 int nchunk, double2 *__restrict__ part, int chunk, int g, int vt);

__global__ __launch_bounds__(256) void k_pw_becp(PwDev d, const double2 *__restrict__ in,
                                                 int64_t sin, const int32_t *__restrict__ active,
                                                 const int32_t *__restrict__ gvec,
                                                 const int32_t *__restrict__ gcnt, int maxper,
                                                 int tiles, int nchunk,
                                                 double2 *__restrict__ part) {
  becp_body(d, in, sin, active, gvec, gcnt, maxper, tiles, nchunk, part, blockIdx.x, blockIdx.y,
            blockIdx.z);
}

__device__ inline void becp_body(const PwDev &d, const double2 *__restrict__ in, int64_t sin,
                                 const int32_t *__restrict__ active,
                                 const int32_t *__restrict__ gvec,
                                 const int32_t *__restrict__ gcnt, int maxper, int tiles,
                                 int nchunk, double2 *__restrict__ part, int chunk, int g, int vt) {
  const int cnt = gcnt[g], v0 = vt * kVT;
  const int npw = d.npw[g], i0 = chunk * kChunk;
  if (v0 >= cnt || i0 >= npw) return;
  const int nv = min(kVT, cnt - v0), lim = min(kChunk, npw - i0), nproj = d.nproj;
  if (active && !tile_active(active, gvec + g * maxper + v0, nv)) return;
  __shared__ double2 xs[kVT][kChunk + 1];
  __shared__ double2 ps[kPT][kChunk + 1];
  for (int t = threadIdx.x; t < kVT * kChunk; t += 256) {
    const int vl = t / kChunk, i = t % kChunk;
    double2 val = mk(0, 0);
    if (vl < nv && i < lim) {
      const int v = gvec[g * maxper + v0 + vl];
      if (!active || active[v]) val = in[v * sin + i0 + i];
    }
    xs[vl][i] = val;
  }
  const double2 *pg = d.proj + (int64_t)g * nproj * d.ld + i0;
  double2 *pout = part + (((int64_t)g * tiles + vt) * nchunk + chunk) * nproj * kVT;
  for (int p0 = 0; p0 < nproj; p0 += kPT) {
    __syncthreads();
    for (int t = threadIdx.x; t < kPT * kChunk; t += 256) {
      const int pl = t / kChunk, i = t % kChunk;
      ps[pl][i] = (p0 + pl < nproj && i < lim) ? pg[(int64_t)(p0 + pl) * d.ld + i] : mk(0, 0);
    }
    __syncthreads();
    const int o = threadIdx.x;  // kPT * kVT == 256 outputs
    const int pl = o / kVT, vl = o % kVT;
    double2 acc = mk(0, 0);
    for (int i = 0; i < lim; ++i) cmacc(acc, ps[pl][i], xs[vl][i]);
    if (p0 + pl < nproj) pout[(p0 + pl) * kVT + vl] = acc;
  }
}

// out_v = kin x_v + fwfft(V invfft x_v) + proj (D becp_v) + sigma_v x_v for the
// plane waves of one chunk and the vectors of one tile; becp_v is the sum of
// the chunk partials, the local term the z forward transform of S.
// CH plane waves per block of NT lanes: thread = (plane wave i < CH, vector
// group vq < NT / CH), each thread kVT / (NT / CH) vectors.
template <int CH, int NT>
__global__ __launch_bounds__(NT) void k_pw_assemble2(
    PwDev d, const double2 *__restrict__ in, int64_t sin, double2 *__restrict__ out, int64_t sout,
    const double2 *__restrict__ sigma, const int32_t *__restrict__ active,
    const int32_t *__restrict__ gvec, const int32_t *__restrict__ gcnt, int maxper, int tiles,
    int nchunk, const double2 *__restrict__ part, const double2 *__restrict__ S,
    const double2 *__restrict__ locv) {
  extern __shared__ double2 smem[];
  const int chunk = blockIdx.x, g = blockIdx.y, vt = blockIdx.z;
  const int cnt = gcnt[g], v0 = vt * kVT;
  if (v0 >= cnt) return;
  const int nv = min(kVT, cnt - v0), nproj = d.nproj, n3 = d.n3;
  const int npw = d.npw[g], i0 = chunk * CH;
  if (active && !tile_active(active, gvec + g * maxper + v0, nv)) return;
  double2 *becp = smem;            // [nproj][kVT]
  double2 *w3 = becp + nproj * kVT;  // n3
  const double2 *bsrc = part + ((int64_t)g * tiles + vt) * nproj * kVT;  // reduced D becp
  for (int o = threadIdx.x; o < nproj * kVT; o += NT) becp[o] = bsrc[o];
  for (int i = threadIdx.x; i < n3; i += NT) w3[i] = d.tw3[i];
  __syncthreads();
  // thread = (plane wave i of the chunk, vector quarter vq): vectors
  // vl = vq + 4k, k < 4, accumulate together so each projector value is
  // loaded once per thread
  constexpr int kVQ = NT / CH;
  constexpr int kPer = kVT / kVQ;
  static_assert(kVT % kVQ == 0, "vector groups must divide the tile");
  const int i = threadIdx.x % CH, vq = threadIdx.x / CH;
  const int64_t e = (int64_t)i0 + i;
  if (e >= d.ld) return;
  const double2 *pg = d.proj + (int64_t)g * nproj * d.ld;
  const int64_t ge = (int64_t)g * d.ld + e;
  int vv[kPer];
  bool on[kPer];
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int vl = vq + kVQ * k;
    on[k] = vl < nv;
    vv[k] = on[k] ? gvec[g * maxper + v0 + vl] : 0;
    if (on[k] && active && !active[vv[k]]) on[k] = false;
  }
  bool any = false;
#pragma unroll
  for (int k = 0; k < kPer; ++k) any |= on[k];
  if (!any) return;  // no running vector in this thread's group
  if (e >= npw) {  // padding stays zero
#pragma unroll
    for (int k = 0; k < kPer; ++k)
      if (on[k]) out[vv[k] * sout + e] = mk(0, 0);
    return;
  }
  const double kin = d.kin[ge];
  double2 xi[kPer], acc[kPer];
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    xi[k] = on[k] ? in[vv[k] * sin + e] : mk(0, 0);
    acc[k] = mk(kin * xi[k].x, kin * xi[k].y);
    if (!on[k]) continue;
    if (locv) {  // z transform done by k_pw_zfwd (already scaled)
      const double2 l = locv[vv[k] * d.ld + e];
      acc[k].x += l.x;
      acc[k].y += l.y;
    } else {
      const double2 *col = S + ((int64_t)vv[k] * d.nslot + d.slot_of[ge]) * n3;
      const int mz = d.m3_of[ge];
      double2 loc = mk(0, 0);
      for (int z = 0; z < n3; ++z) cmac(loc, col[z], cconj(w3[(mz * z) % n3]));
      acc[k].x = fma(loc.x, d.inv_nbox, acc[k].x);
      acc[k].y = fma(loc.y, d.inv_nbox, acc[k].y);
    }
  }
  for (int p = 0; p < nproj; ++p) {
    const double2 pv = pg[(int64_t)p * d.ld + e];
#pragma unroll
    for (int k = 0; k < kPer; ++k) cmac(acc[k], pv, becp[p * kVT + vq + kVQ * k]);
  }
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    if (!on[k]) continue;
    cmac(acc[k], sigma[vv[k]], xi[k]);
    out[vv[k] * sout + e] = acc[k];
  }
}

// f(std::integral_constant<int, M>) for the M of the list equal to m; false if none
template <int... Ms, typename F> inline bool dispatch_band(int m, F &&f) {
  bool hit = false;
  ((m == Ms ? (f(std::integral_constant<int, Ms>{}), hit = true) : false), ...);
  return hit;
}

// f(std::integral_constant<int, M>) with M = m when m is one of the sphere
// half-widths specialised for box N (the usual ones: the box is the smallest
// FFT size of at least 4M + 1, or one size up), else with M = -1
template <int N, typename F> inline void with_band(bool enable, int m, F &&f) {
  bool hit = false;
  if (enable) {
    if constexpr (N == 48) hit = dispatch_band<8, 9, 10, 11>(m, f);
    if constexpr (N == 36) hit = dispatch_band<7, 8>(m, f);
    if constexpr (N == 32) hit = dispatch_band<6, 7>(m, f);
    if constexpr (N == 24) hit = dispatch_band<4, 5>(m, f);
  }
  if (!hit) f(std::integral_constant<int, -1>{});
}

// Two independent stages of one FP64 application in one launch each, which
// saves a launch boundary (and the tail of the smaller stage) per pair:
//   k_pw_becp_zinv:    blocks [0, nbecp) are projector partial sums (256
//                      lanes); the rest hold four 64-lane z inverse groups;
//   k_pw_reduce_plane: blocks [0, nred) reduce the partials (64 lanes); the
//                      rest are the plane transforms (z fastest).
// The block index orders match the separate kernels' grids, and every branch
// on it is uniform within a block, so the barriers of either part are safe.
// SPLIT: the z inverse transforms split over lane pairs (32 columns per
// 64-lane group), else one lane per column.
// vlist[0 .. k) = the vectors v < nvec with active[v] != 0 (all of them
// without flags) in increasing order, vlist[k .. nvec) = -1; one block of
// 256 lanes
__device__ inline void vlist_body(int nvec, const int32_t *__restrict__ active,
                                  int32_t *__restrict__ vlist) {
  __shared__ int wsum[4];
  __shared__ int base;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  if (t == 0) base = 0;
  __syncthreads();
  for (int v0 = 0; v0 < nvec; v0 += 256) {
    const int v = v0 + t;
    const bool on = v < nvec && (!active || active[v] != 0);
    const unsigned long long m = __ballot(on);
    if (lane == 0) wsum[w] = __popcll(m);
    __syncthreads();
    int off = base;
    for (int i = 0; i < w; ++i) off += wsum[i];
    if (on) vlist[off + __popcll(m & ((1ull << lane) - 1))] = v;
    __syncthreads();  // every lane has read base and wsum
    if (t == 0) base += wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
  }
  for (int i = base + t; i < nvec; i += 256) vlist[i] = -1;
}

template <int N, bool SPLIT, int MB, typename C = double2, bool RL = false>
__global__ __launch_bounds__(256) void k_pw_becp_zinv(
    PwDev d, const double2 *__restrict__ in, int64_t sin, const int32_t *__restrict__ group,
    const int32_t *__restrict__ active, const int32_t *__restrict__ gvec,
    const int32_t *__restrict__ gcnt, int maxper, int tiles, int nchunk, int ngroup,
    double2 *__restrict__ part, C *__restrict__ T, int nvec,
    const int32_t *__restrict__ relax = nullptr, int want = 0,
    int32_t *__restrict__ vlist = nullptr) {
  // with a list to build, block 0 builds it (first, so that its serial
  // scan does not trail the launch) and the others shift down by one
  const int nbecp = nchunk * ngroup * tiles, b = (int)blockIdx.x - (vlist ? 1 : 0);
  if (b < 0) {
    vlist_body(nvec, active, vlist);
    return;
  }
  if (b < nbecp) {
    becp_body(d, in, sin, active, gvec, gcnt, maxper, tiles, nchunk, part, b % nchunk,
              (b / nchunk) % ngroup, b / (nchunk * ngroup));
    return;
  }
  constexpr int kCols = SPLIT ? 32 : 64;  // columns per 64-lane group
  const int gx = ((SPLIT ? d.ncs : d.nslot) + kCols - 1) / kCols;
  const int zb = (b - nbecp) * 4 + (int)(threadIdx.x >> 6);
  if (zb >= gx * nvec) return;
  const int lane = (int)(threadIdx.x & 63);
  if constexpr (SPLIT)
    zinv_split_body<N, MB, C, RL>(d, in, sin, group, active, T, zb / gx,
                                  (zb % gx) * 32 + (lane >> 1), (lane & 1) != 0, relax, want);
  else
    zinv_body<N, C>(d, in, sin, group, active, T, RL ? relax : nullptr, want, zb / gx,
                    (zb % gx) * 64 + lane);
}

// PAIR: two vectors per plane block (plane2_body), else one (plane_body)
template <int N, bool PAIR>
__global__ __launch_bounds__(64) void k_pw_reduce_plane(
    PwDev d, const int32_t *__restrict__ group, const int32_t *__restrict__ active,
    const int32_t *__restrict__ gvec, const int32_t *__restrict__ gcnt, int maxper, int tiles,
    int nchunk, int ngroup, int nred_x, const double2 *__restrict__ part,
    double2 *__restrict__ becp, const double2 *__restrict__ T, double2 *__restrict__ S,
    int nvec) {
  extern __shared__ double2 smem_d[];
  const int nred = nred_x * ngroup * tiles, b = blockIdx.x;
  if (b < nred) {
    becp_reduce_body(d, active, gvec, gcnt, maxper, tiles, nchunk, part, becp, b % nred_x,
                     (b / nred_x) % ngroup, b / (nred_x * ngroup), threadIdx.x);
    return;
  }
  const int pb = b - nred;
  if constexpr (PAIR)
    plane2_body<N>(d, T, group, active, S, pb % N, 2 * (pb / N), nvec, smem_d);
  else
    plane_body<N, double2>(d, T, group, active, S, nullptr, 0, pb % N, pb / N, smem_d);
}

// As k_pw_reduce_plane with the split plane transforms (blocks of W
// wavefronts, PL vectors per plane block): a reduction block covers W of the
// 64-output ranges.  MB as in plane_split_body.
template <int N, int MB, int PL, int W, typename C = double2, bool RL = false>
__global__ __launch_bounds__(64 * W) void k_pw_reduce_split(
    PwDev d, const int32_t *__restrict__ group, const int32_t *__restrict__ active,
    const int32_t *__restrict__ gvec, const int32_t *__restrict__ gcnt, int maxper, int tiles,
    int nchunk, int ngroup, int nred_x, const double2 *__restrict__ part,
    double2 *__restrict__ becp, const C *__restrict__ T, C *__restrict__ S, int nvec,
    const int32_t *__restrict__ relax = nullptr, int want = 0,
    const int32_t *__restrict__ vlist = nullptr) {
  extern __shared__ double2 smem_d[];
  const int nred_w = (nred_x + W - 1) / W, nred = nred_w * ngroup * tiles, b = blockIdx.x;
  if (b < nred) {
    const int bx = W * (b % nred_w) + (int)(threadIdx.x >> 6);
    if (bx < nred_x)
      becp_reduce_body(d, active, gvec, gcnt, maxper, tiles, nchunk, part, becp, bx,
                       (b / nred_w) % ngroup, b / (nred_w * ngroup), threadIdx.x & 63);
    return;
  }
  const int pb = b - nred;
  plane_split_body<N, MB, PL, W, C, RL>(d, T, group, active, S, pb % N, PL * (pb / N), nvec,
                                        reinterpret_cast<C *>(smem_d), relax, want, vlist);
}

// planes per block and wavefronts per block of the split plane kernel where
// no other shape is tried (mixed-precision batches): one plane on one
// wavefront (on the 48 box the X stage in 1.5 rounds of 32 lane pairs, 13 KB
// of LDS).  FP64 batches take the autotuned shapes at the launch (three
// planes on two wavefronts first, then this).
template <int N> struct SplitShape {
  static constexpr int PL = 1;
  static constexpr int W = 1;
};

}  // namespace

// ---------------------------------------------------------------------------
// DenseOperator
// ---------------------------------------------------------------------------
DenseOperator::DenseOperator(const double2 *h_a, int64_t n_) : a_(n_ * n_) {
  n = n_;
  ld = n_;
  a_.upload(h_a, n_ * n_);
  SGW_HIP(hipDeviceSynchronize());
}

void DenseOperator::apply(hipStream_t st, int nvec, const double2 *in, int64_t in_stride,
                          double2 *out, int64_t out_stride, const double2 *d_sigma,
                          const double2 *, const int32_t *d_active, const int32_t *,
                          const int32_t *) {
  dim3 grid((unsigned)((ld + kBlock - 1) / kBlock), nvec);
  {
    KScope ks("happly", st);
    k_dense<<<grid, kBlock, 0, st>>>(n, a_.p, in, in_stride, out, out_stride, ld, d_sigma,
                                     d_active);
  }
  SGW_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// CallbackOperator
// ---------------------------------------------------------------------------
void CallbackOperator::apply(hipStream_t st, int nvec, const double2 *in, int64_t in_stride,
                             double2 *out, int64_t out_stride, const double2 *,
                             const double2 *h_sigma, const int32_t *, const int32_t *h_active,
                             const int32_t *) {
  hx_.resize(n);
  hax_.resize(n);
  SGW_HIP(hipStreamSynchronize(st));
  for (int v = 0; v < nvec; ++v) {
    if (h_active && !h_active[v]) continue;
    SGW_HIP(hipMemcpy(hx_.data(), in + v * in_stride, sizeof(double2) * n, hipMemcpyDeviceToHost));
    fn_(&h_sigma[v], hx_.data(), hax_.data(), n, ctx_);
    SGW_HIP(hipMemcpy(out + v * out_stride, hax_.data(), sizeof(double2) * n,
                      hipMemcpyHostToDevice));
  }
}

// ---------------------------------------------------------------------------
// PwOperator
// ---------------------------------------------------------------------------
static std::vector<double2> twiddles(int n) {
  std::vector<double2> w(n);
  for (int k = 0; k < n; ++k) {
    double ang = 2.0 * M_PI * (double)k / n;
    w[k] = mk(std::cos(ang), std::sin(ang));
  }
  return w;
}

static inline int modn(int a, int n) { return ((a % n) + n) % n; }

PwOperator::PwOperator(const PwDesc &desc) {
  n1_ = desc.n1; n2_ = desc.n2; n3_ = desc.n3;
  if (const char *f = std::getenv("SGW_PW_FUSE")) fuse_stages_ = std::strcmp(f, "0") != 0;
  if (const char *f = std::getenv("SGW_PW_PAIR")) pair_planes_ = std::strcmp(f, "0") != 0;
  if (const char *f = std::getenv("SGW_PW_SPLIT")) split_planes_ = std::strcmp(f, "0") != 0;
  if (const char *f = std::getenv("SGW_PW_BAND")) band_planes_ = std::strcmp(f, "0") != 0;
  SGW_REQUIRE(n1_ > 0 && n2_ > 0 && n3_ > 0 && n1_ <= 512 && n2_ <= 512 && n3_ <= 512, SGW_E_ARG,
              "pw operator: FFT box dimensions must be in [1, 512]");
  ngroup_ = desc.ngroup;
  SGW_REQUIRE(ngroup_ >= 1, SGW_E_ARG, "pw operator: need at least one group");
  ld = desc.ld;
  nproj_ = desc.nproj;
  SGW_REQUIRE(nproj_ >= 0 && nproj_ <= 4096, SGW_E_ARG, "pw operator: nproj out of range");
  npw_.assign(desc.npw, desc.npw + ngroup_);
  int M1 = 0, M2 = 0, M3 = 0;
  for (int g = 0; g < ngroup_; ++g) {
    SGW_REQUIRE(npw_[g] >= 1 && npw_[g] <= ld, SGW_E_SHAPE, "pw operator: npw must be in [1, ld]");
    for (int i = 0; i < npw_[g]; ++i) {
      const int32_t *m = desc.mill + ((int64_t)g * ld + i) * 3;
      M1 = std::max(M1, std::abs(m[0]));
      M2 = std::max(M2, std::abs(m[1]));
      M3 = std::max(M3, std::abs(m[2]));
    }
  }
  SGW_REQUIRE(2 * M1 + 1 <= n1_ && 2 * M2 + 1 <= n2_ && 2 * M3 + 1 <= n3_, SGW_E_SHAPE,
              "pw operator: the plane-wave sphere does not fit in the FFT box");
  n = npw_[0];
  M1_ = M1; M2_ = M2; M3_ = M3;
  nc1_ = 2 * M1 + 1; nc2_ = 2 * M2 + 1;
  nslot_ = nc1_ * nc2_;
  // column tables
  std::vector<int32_t> col_start((size_t)ngroup_ * nslot_, 0), col_cnt((size_t)ngroup_ * nslot_, 0);
  std::vector<int32_t> sidx((size_t)ngroup_ * ld, 0), sm3((size_t)ngroup_ * ld, 0);
  std::vector<int32_t> slot_of((size_t)ngroup_ * ld, 0), m3_of((size_t)ngroup_ * ld, 0);
  std::vector<int32_t> col_lo((size_t)ngroup_ * nslot_, 0);
  for (int g = 0; g < ngroup_; ++g) {
    std::vector<std::pair<int64_t, int>> key(npw_[g]);
    std::map<int64_t, int> seen;
    for (int i = 0; i < npw_[g]; ++i) {
      const int32_t *m = desc.mill + ((int64_t)g * ld + i) * 3;
      int slot = (m[0] + M1) * nc2_ + (m[1] + M2);
      key[i] = {(int64_t)slot * 4096 + (m[2] + 2048), i};
      SGW_REQUIRE(seen.emplace(key[i].first, i).second, SGW_E_ARG,
                  "pw operator: duplicate Miller index in a sphere");
      slot_of[(size_t)g * ld + i] = slot;
      m3_of[(size_t)g * ld + i] = modn(m[2], n3_);
    }
    std::sort(key.begin(), key.end());
    for (int j = 0; j < npw_[g]; ++j) {
      int i = key[j].second;
      int slot = slot_of[(size_t)g * ld + i];
      const int m3raw = desc.mill[((int64_t)g * ld + i) * 3 + 2];
      if (col_cnt[(size_t)g * nslot_ + slot] == 0) {
        col_start[(size_t)g * nslot_ + slot] = j;
        col_lo[(size_t)g * nslot_ + slot] = m3raw;
      } else if (m3raw != col_lo[(size_t)g * nslot_ + slot] + col_cnt[(size_t)g * nslot_ + slot]) {
        columns_contiguous_ = false;  // a column with a gap in m3: no z-FFT kernels
      }
      col_cnt[(size_t)g * nslot_ + slot]++;
      sidx[(size_t)g * ld + j] = i;
      sm3[(size_t)g * ld + j] = m3_of[(size_t)g * ld + i];
    }
  }
  {
    // packed column layout: per group, row a of the grid from its first to
    // its last column with plane waves (a row of a sphere's columns is one
    // run, so the packed planes hold exactly the sphere's columns; a gap
    // would be kept, as zeros)
    const int nr = nc1_ + 1;
    std::vector<int32_t> rs((size_t)ngroup_ * nr, 0), rl((size_t)ngroup_ * nc1_, 0),
        rc((size_t)ngroup_ * nc1_, 0);
    int ncs = 1;
    for (int g = 0; g < ngroup_; ++g) {
      int j = 0;
      for (int a = 0; a < nc1_; ++a) {
        int lo = nc2_, hi = -1;
        for (int b = 0; b < nc2_; ++b)
          if (col_cnt[(size_t)g * nslot_ + a * nc2_ + b] > 0) {
            lo = std::min(lo, b);
            hi = std::max(hi, b);
          }
        rs[(size_t)g * nr + a] = j;
        if (hi >= lo) {
          rl[(size_t)g * nc1_ + a] = lo - M2_;
          rc[(size_t)g * nc1_ + a] = hi - lo + 1;
          j += hi - lo + 1;
        }
      }
      rs[(size_t)g * nr + nc1_] = j;
      ncs = std::max(ncs, j);
    }
    ncs_ = ncs;
    // (slot, count, start, lowest m3) of every packed column, (-1, 0, 0, 0)
    // past a group's columns
    auto info = [&](int g, int c) {
      const size_t i = (size_t)g * nslot_ + c;
      return make_int4(c, col_cnt[i], col_start[i], col_lo[i]);
    };
    std::vector<int4> cs((size_t)ngroup_ * ncs_, make_int4(-1, 0, 0, 0));
    for (int g = 0; g < ngroup_; ++g)
      for (int a = 0; a < nc1_; ++a)
        for (int k = 0; k < rc[(size_t)g * nc1_ + a]; ++k)
          cs[(size_t)g * ncs_ + rs[(size_t)g * nr + a] + k] =
              info(g, a * nc2_ + rl[(size_t)g * nc1_ + a] + M2_ + k);
    pk_info_.alloc(cs.size()); pk_info_.upload(cs.data(), cs.size());
    pk_row_start_.alloc(rs.size()); pk_row_start_.upload(rs.data(), rs.size());
    pk_row_lo_.alloc(rl.size()); pk_row_lo_.upload(rl.data(), rl.size());
    pk_row_cnt_.alloc(rc.size()); pk_row_cnt_.upload(rc.data(), rc.size());
    // the full grid in the same terms
    std::vector<int4> gs((size_t)ngroup_ * nslot_);
    std::vector<int32_t> grs((size_t)ngroup_ * nr), grl((size_t)ngroup_ * nc1_, -M2_),
        grc((size_t)ngroup_ * nc1_, nc2_);
    for (int g = 0; g < ngroup_; ++g) {
      for (int c = 0; c < nslot_; ++c) gs[(size_t)g * nslot_ + c] = info(g, c);
      for (int a = 0; a <= nc1_; ++a) grs[(size_t)g * nr + a] = a * nc2_;
    }
    gr_info_.alloc(gs.size()); gr_info_.upload(gs.data(), gs.size());
    gr_row_start_.alloc(grs.size()); gr_row_start_.upload(grs.data(), grs.size());
    gr_row_lo_.alloc(grl.size()); gr_row_lo_.upload(grl.data(), grl.size());
    gr_row_cnt_.alloc(grc.size()); gr_row_cnt_.upload(grc.data(), grc.size());
  }
  const int64_t nbox = (int64_t)n1_ * n2_ * n3_;
  vloc_.alloc(nbox); vloc_.upload(desc.vloc, nbox);
  {  // [z][x][y] copy for the x stage of the plane kernels
    std::vector<double> t((size_t)nbox);
    for (int z = 0; z < n3_; ++z)
      for (int y = 0; y < n2_; ++y)
        for (int x = 0; x < n1_; ++x)
          t[((size_t)z * n1_ + x) * n2_ + y] = desc.vloc[((size_t)z * n2_ + y) * n1_ + x];
    vloc_zxy_.alloc(nbox);
    vloc_zxy_.upload(t.data(), nbox);
  }
  d_npw_.alloc(ngroup_); d_npw_.upload(npw_.data(), ngroup_);
  col_start_.alloc(col_start.size()); col_start_.upload(col_start.data(), col_start.size());
  col_cnt_.alloc(col_cnt.size()); col_cnt_.upload(col_cnt.data(), col_cnt.size());
  sidx_.alloc(sidx.size()); sidx_.upload(sidx.data(), sidx.size());
  sm3_.alloc(sm3.size()); sm3_.upload(sm3.data(), sm3.size());
  slot_of_.alloc(slot_of.size()); slot_of_.upload(slot_of.data(), slot_of.size());
  m3_of_.alloc(m3_of.size()); m3_of_.upload(m3_of.data(), m3_of.size());
  col_lo_.alloc(col_lo.size()); col_lo_.upload(col_lo.data(), col_lo.size());
  kin_.alloc((size_t)ngroup_ * ld); kin_.upload(desc.kin, (size_t)ngroup_ * ld);
  if (nproj_ > 0) {
    proj_.alloc((size_t)ngroup_ * nproj_ * ld);
    proj_.upload(desc.proj, (size_t)ngroup_ * nproj_ * ld);
    dcoef_.alloc((size_t)ngroup_ * nproj_);
    dcoef_.upload(desc.dcoef, (size_t)ngroup_ * nproj_);
  }
  auto t1 = twiddles(n1_), t2 = twiddles(n2_), t3 = twiddles(n3_);
  tw_.alloc(n1_ + n2_ + n3_);
  tw_.upload(t1.data(), n1_);
  SGW_HIP(hipMemcpy(tw_.p + n1_, t2.data(), sizeof(double2) * n2_, hipMemcpyHostToDevice));
  SGW_HIP(hipMemcpy(tw_.p + n1_ + n2_, t3.data(), sizeof(double2) * n3_, hipMemcpyHostToDevice));
  SGW_HIP(hipDeviceSynchronize());
}

int64_t PwOperator::group_n(int g) const { return npw_.at(g); }


void PwOperator::ensure_scratch(int nvec) {
  if (vlist_.n < (size_t)nvec) vlist_.alloc(nvec);
  size_t need = (size_t)nvec * nslot_ * n3_;
  if (S_.n < need) S_.alloc(need);
  if (T_.n < need) T_.alloc(need);
  if (loc_.n < (size_t)nvec * ld) loc_.alloc((size_t)nvec * ld);
}

void PwOperator::apply(hipStream_t st, int nvec, const double2 *in, int64_t in_stride,
                       double2 *out, int64_t out_stride, const double2 *d_sigma, const double2 *,
                       const int32_t *d_active, const int32_t *, const int32_t *d_group) {
  if (nvec <= 0) return;
  SGW_REQUIRE(d_group != nullptr || ngroup_ == 1, SGW_E_ARG,
              "pw operator: a group index per vector is required with several groups");
  SGW_REQUIRE(plan_nvec_ == nvec, SGW_E_STATE,
              "pw operator: prepare() was not called for this batch of vectors");
  ensure_scratch(nvec);
  PwWork w{T_.p,       S_.p,      loc_.p,    part_.p,      becp_.p,
           plan_vec_.p, plan_cnt_.p, plan_max_, plan_tiles_, vlist_.p};
  apply_work(st, nvec, in, in_stride, out, out_stride, d_sigma, d_active, d_group, w);
}

void PwOperator::apply_work(hipStream_t st, int nvec, const double2 *in, int64_t in_stride,
                            double2 *out, int64_t out_stride, const double2 *d_sigma,
                            const int32_t *d_active, const int32_t *d_group, const PwWork &wk) {
  KScope whole("happly", st);
  PwDev d;
  d.n1 = n1_; d.n2 = n2_; d.n3 = n3_;
  d.M1 = M1_; d.M2 = M2_; d.nc1 = nc1_; d.nc2 = nc2_; d.nslot = nslot_;
  d.ld = ld; d.nproj = nproj_;
  d.vloc = vloc_.p; d.vloc_zxy = vloc_zxy_.p; d.npw = d_npw_.p;
  d.col_start = col_start_.p; d.col_cnt = col_cnt_.p;
  d.sorted_idx = sidx_.p; d.sorted_m3 = sm3_.p; d.slot_of = slot_of_.p; d.m3_of = m3_of_.p;
  d.col_lo = col_lo_.p;
  d.ncs = nslot_; d.col_info = gr_info_.p;
  d.row_start = gr_row_start_.p; d.row_lo = gr_row_lo_.p; d.row_cnt = gr_row_cnt_.p;
  d.kin = kin_.p; d.proj = proj_.p; d.dcoef = dcoef_.p;
  d.tw1 = tw_.p; d.tw2 = tw_.p + n1_; d.tw3 = tw_.p + n1_ + n2_;
  d.inv_nbox = 1.0 / ((double)n1_ * n2_ * n3_);
  const bool cubic = n1_ == n2_ && n2_ == n3_ && columns_contiguous_;
  const size_t lds_fft = sizeof(double2) * (nslot_ + (size_t)nc1_ * n1_);
  const dim3 gcol((unsigned)((nslot_ + 63) / 64), nvec);
  const double2 *locv = nullptr;
  const int nchunk = (int)((ld + kChunk - 1) / kChunk);
  const int nred_x = (nproj_ * kVT + 63) / 64;
  // FP32 (fft_mode_ 1, or 2 per vector) takes the fused split kernels where
  // the split plane kernel applies, the separate ones otherwise
  const bool split_plane_ok = split_planes_ && n1_ <= 48 && nc1_ <= 32 &&
                              sizeof(double2) * (size_t)nc1_ * (n1_ + 1) <= 64 * 1024;
  const bool fused = fuse_stages_ && cubic &&
                     (n1_ == 24 || n1_ == 32 || n1_ == 36 || n1_ == 48 || n1_ == 64) &&
                     (fft_mode_ == 0 || split_plane_ok);
  if (nproj_ > 0 && !fused) {
    // projector contraction becp = D proj^H x (a concurrent stream for it was
    // measured: no gain, the plane kernel already fills the device)
    KScope ks("pw_becp", st, 2);
    k_pw_becp<<<dim3(nchunk, ngroup_, wk.plan_tiles), 256, 0, st>>>(
        d, in, in_stride, d_active, wk.plan_vec, wk.plan_cnt, wk.plan_max, wk.plan_tiles, nchunk,
        wk.part);
    k_pw_becp_reduce<<<dim3((unsigned)nred_x, ngroup_, wk.plan_tiles), 64, 0, st>>>(
        d, d_active, wk.plan_vec, wk.plan_cnt, wk.plan_max, wk.plan_tiles, nchunk, wk.part, wk.becp);
    SGW_HIP(hipGetLastError());
  }
  auto launch_fft = [&](auto tag) {
    constexpr int NB = decltype(tag)::value;
    if (fused) {
      // no projectors: the part counts are zero and only the transforms run
      const int tiles = nproj_ > 0 ? wk.plan_tiles : 0;
      const int64_t nbecp = (int64_t)nchunk * ngroup_ * tiles;
      // fft_mode_: 0 FP64, 1 FP32, 2 per vector from the relax flags.  With
      // both precisions the FP64 instances carry the projector sums and the
      // FP32 ones (tiles 0) the transforms of the relaxed vectors only.  The
      // FP64 instances read the flags only in mode 2 (template RL), so mode 0
      // runs the code it runs without an FP32 path.
      const int32_t *rl = fft_mode_ == 2 ? relax_ : nullptr;
      auto with_rl = [&](auto &&f) {
        if (rl) f(std::true_type{});
        else f(std::false_type{});
      };
      const bool run64 = fft_mode_ != 1, run32 = fft_mode_ != 0;
      const int tiles32 = run64 ? 0 : tiles;
      const int64_t nbecp32 = (int64_t)nchunk * ngroup_ * tiles32;
      float2 *T32 = reinterpret_cast<float2 *>(wk.T), *S32 = reinterpret_cast<float2 *>(wk.S);
      // z transforms split over lane pairs: 32 columns per 64 lanes
      const bool zsplit = split_planes_;
      // T and S in the packed column layout where the split plane kernel
      // runs (it and the split z transforms are their only readers and
      // writers here); the full grid for the other plane kernels
      const bool packed = zsplit && split_plane_ok;
      // the split plane kernel's blocks take the running vectors from a list
      // (built by the z inverse launch), so that a batch with few running
      // vectors launches no blocks with dead planes among live ones
      int32_t *vl = (packed && d_active && wk.vlist) ? wk.vlist : nullptr;
      if (packed) {
        d.ncs = ncs_; d.col_info = pk_info_.p;
        d.row_start = pk_row_start_.p; d.row_lo = pk_row_lo_.p; d.row_cnt = pk_row_cnt_.p;
      }
      const int64_t gxz = zsplit ? (d.ncs + 31) / 32 : (int64_t)gcol.x;
      const int64_t nzinv = (gxz * nvec + 3) / 4;
      const int64_t nred = (int64_t)nred_x * ngroup_ * tiles;
      SGW_REQUIRE(nbecp + nzinv < INT32_MAX && nred + (int64_t)NB * nvec < INT32_MAX, SGW_E_ARG,
                  "pw operator: too many vectors for one launch");
      {
        KScope ks("pw_becp_zinv", st, 2);
        if (zsplit)
          with_band<NB>(band_planes_, M3_, [&](auto mb) {
            // the first launch's extra block (block 0) lists the running vectors
            int32_t *vl64 = run64 ? vl : nullptr, *vl32 = run64 ? nullptr : vl;
            if (run64)
              with_rl([&](auto rlt) {
                k_pw_becp_zinv<NB, true, decltype(mb)::value, double2, decltype(rlt)::value>
                    <<<(unsigned)(nbecp + nzinv + (vl64 ? 1 : 0)), 256, 0, st>>>(
                        d, in, in_stride, d_group, d_active, wk.plan_vec, wk.plan_cnt,
                        wk.plan_max, tiles, nchunk, ngroup_, wk.part, wk.T, nvec, rl, 0, vl64);
              });
            if (run32)
              k_pw_becp_zinv<NB, true, decltype(mb)::value, float2, true>
                  <<<(unsigned)(nbecp32 + nzinv + (vl32 ? 1 : 0)), 256, 0, st>>>(
                      d, in, in_stride, d_group, d_active, wk.plan_vec, wk.plan_cnt, wk.plan_max,
                      tiles32, nchunk, ngroup_, wk.part, T32, nvec, rl, 1, vl32);
          });
        else
          k_pw_becp_zinv<NB, false, -1><<<(unsigned)(nbecp + nzinv), 256, 0, st>>>(
              d, in, in_stride, d_group, d_active, wk.plan_vec, wk.plan_cnt, wk.plan_max, tiles,
              nchunk, ngroup_, wk.part, wk.T, nvec);
      }
      {
        KScope ks("pw_reduce_plane", st, 2);
        // the transforms split over lane pairs (N <= 48); else, without the
        // split, two vectors per plane block while both Y stages fit in one
        // wavefront and four such blocks still fit in a CU's LDS (N = 64
        // would spill)
        bool paired = false;
        if constexpr (NB <= 48) {
          // split transforms over lane pairs (LDS: P of the block's planes only)
          constexpr int PL = SplitShape<NB>::PL, W = SplitShape<NB>::W;
          const size_t lds_split = sizeof(double2) * PL * (size_t)nc1_ * (NB + 1);
          auto launch_split = [&](auto shape, auto mb, auto rlt) {
            constexpr int pl = decltype(shape)::value / 8, w = decltype(shape)::value % 8;
            const int64_t nred_s = (int64_t)((nred_x + w - 1) / w) * ngroup_ * tiles;
            k_pw_reduce_split<NB, decltype(mb)::value, pl, w, double2, decltype(rlt)::value>
                <<<(unsigned)(nred_s + (int64_t)NB * ((nvec + pl - 1) / pl)), 64 * w,
                   sizeof(double2) * pl * (size_t)nc1_ * (NB + 1), st>>>(
                    d, d_group, d_active, wk.plan_vec, wk.plan_cnt, wk.plan_max, tiles, nchunk,
                    ngroup_, nred_x, wk.part, wk.becp, wk.T, wk.S, nvec, rl, 0, vl);
          };
          // the relaxed vectors' FP32 plane transforms, one plane per wavefront
          auto launch_split32 = [&](auto mb) {
            const int64_t nred_s = (int64_t)nred_x * ngroup_ * tiles32;
            k_pw_reduce_split<NB, decltype(mb)::value, 1, 1, float2, true>
                <<<(unsigned)(nred_s + (int64_t)NB * nvec), 64,
                   sizeof(float2) * (size_t)nc1_ * (NB + 1), st>>>(
                    d, d_group, d_active, wk.plan_vec, wk.plan_cnt, wk.plan_max, tiles32, nchunk,
                    ngroup_, nred_x, wk.part, wk.becp, T32, S32, nvec, rl, 1, vl);
          };
          const bool band = band_planes_ && M1_ == M2_;
          static_assert(PL == 1 && W == 1,
                        "split_plane_ok is the default shape's fit test for one plane per wave");
          if (packed) {  // split_planes_, nc1_ <= 32 and lds_split <= 64 KB
            SGW_REQUIRE(lds_split <= 64 * 1024 && PL * nc1_ <= 32 * W, SGW_E_STATE,
                        "pw operator: split plane kernel does not fit");
            // compile-time sphere half-widths for the usual ones of this box;
            // the autotuner picks among three planes on two wavefronts (on
            // the 48 box: Y stage on 51 of 64 lane pairs, X stage in 2.25
            // rounds), one plane on one, one plane on two (the X stage in one
            // round) and two planes on two for this box and size class of the
            // batch (mode 2: one plane on one)
            with_band<NB>(band, M1_, [&](auto mb) {
              using I11 = std::integral_constant<int, PL * 8 + W>;
              constexpr std::false_type no_rl{};
              if (!run64) return;
              if (rl) {
                launch_split(I11{}, mb, std::true_type{});
                return;
              }
              {
                const size_t row = sizeof(double2) * (size_t)nc1_ * (NB + 1);
                const bool fit2 = nc1_ <= 32 && 2 * row <= 64 * 1024;
                const bool fit3 = fit2 && 3 * nc1_ <= 64 && 3 * row <= 64 * 1024;
                const int ncand = fit2 ? (fit3 ? 4 : 3) : 1;
                const std::string key = "plane" + std::to_string(NB) + "/M" +
                                        std::to_string(M1_) + "/b" +
                                        std::to_string((int)decltype(mb)::value) + "/n" +
                                        std::to_string(size_class(nvec));
                // shapes (planes, wavefronts): (3, 2) first where it fits (the
                // streamed 48^3 solve: 365.8k solves/s against 359.2k with
                // (1, 1)), then (1, 1), (1, 2), (2, 2)
                autotune(key, ncand, st, [&](int c) {
                  const int k = fit3 ? c : c + 1;
                  if (k == 0) launch_split(std::integral_constant<int, 3 * 8 + 2>{}, mb, no_rl);
                  else if (k == 1) launch_split(I11{}, mb, no_rl);
                  else if (k == 2)
                    launch_split(std::integral_constant<int, 1 * 8 + 2>{}, mb, no_rl);
                  else launch_split(std::integral_constant<int, 2 * 8 + 2>{}, mb, no_rl);
                });
              }
            });
            if (run32) with_band<NB>(band, M1_, [&](auto mb) { launch_split32(mb); });
            paired = true;
          } else if (pair_planes_ && 2 * nc1_ <= 64 && 2 * lds_fft <= 40 * 1024) {
            k_pw_reduce_plane<NB, true>
                <<<(unsigned)(nred + (int64_t)NB * ((nvec + 1) / 2)), 64, 2 * lds_fft, st>>>(
                    d, d_group, d_active, wk.plan_vec, wk.plan_cnt, wk.plan_max, tiles, nchunk,
                    ngroup_, nred_x, wk.part, wk.becp, wk.T, wk.S, nvec);
            paired = true;
          }
        }
        if (!paired) {
          k_pw_reduce_plane<NB, false><<<(unsigned)(nred + (int64_t)NB * nvec), 64, lds_fft, st>>>(
              d, d_group, d_active, wk.plan_vec, wk.plan_cnt, wk.plan_max, tiles, nchunk, ngroup_,
              nred_x, wk.part, wk.becp, wk.T, wk.S, nvec);
        }
      }
      {
        KScope ks("pw_zfwd", st, 2);
        if (zsplit)
          with_band<NB>(band_planes_, M3_, [&](auto mb) {
            if (run64)
              with_rl([&](auto rlt) {
                k_pw_zfwd_split<NB, decltype(mb)::value, double2, decltype(rlt)::value>
                    <<<dim3((unsigned)gxz, nvec), 64, 0, st>>>(d, wk.S, d_group, d_active,
                                                               wk.loc, rl, 0);
              });
            if (run32)
              k_pw_zfwd_split<NB, decltype(mb)::value, float2, true>
                  <<<dim3((unsigned)gxz, nvec), 64, 0, st>>>(d, S32, d_group, d_active, wk.loc,
                                                             rl, 1);
          });
        else
          k_pw_zfwd<NB><<<gcol, 64, 0, st>>>(d, wk.S, d_group, d_active, wk.loc, nullptr, 0);
      }
      locv = wk.loc;
      return;
    }
    // fft_mode_: 0 FP64, 1 FP32, 2 per vector from the relax flags
    const int32_t *rl = fft_mode_ == 2 ? relax_ : nullptr;
    float2 *T32 = reinterpret_cast<float2 *>(wk.T), *S32 = reinterpret_cast<float2 *>(wk.S);
    const bool run64 = fft_mode_ != 1, run32 = fft_mode_ != 0;
    {
      KScope ks("pw_zinv", st, 2);
      if (run64) k_pw_zinv<NB><<<gcol, 64, 0, st>>>(d, in, in_stride, d_group, d_active, wk.T, rl, 0);
      if (run32)
        k_pw_zinv<NB, float2><<<gcol, 64, 0, st>>>(d, in, in_stride, d_group, d_active, T32, rl, 1);
    }
    {
      KScope ks("pw_plane", st, 2);
      if (run64)
        k_pw_plane_fft<NB><<<dim3(NB, nvec), 64, lds_fft, st>>>(d, wk.T, d_group, d_active,
                                                                 wk.S, rl, 0);
      if (run32)
        k_pw_plane_fft<NB, float2><<<dim3(NB, nvec), 64, lds_fft / 2, st>>>(
            d, T32, d_group, d_active, S32, rl, 1);
    }
    {
      KScope ks("pw_zfwd", st, 2);
      if (run64) k_pw_zfwd<NB><<<gcol, 64, 0, st>>>(d, wk.S, d_group, d_active, wk.loc, rl, 0);
      if (run32)
        k_pw_zfwd<NB, float2><<<gcol, 64, 0, st>>>(d, S32, d_group, d_active, wk.loc, rl, 1);
    }
    locv = wk.loc;
  };
  if (cubic && n1_ == 24) {
    launch_fft(std::integral_constant<int, 24>{});
  } else if (cubic && n1_ == 32) {
    launch_fft(std::integral_constant<int, 32>{});
  } else if (cubic && n1_ == 36) {
    launch_fft(std::integral_constant<int, 36>{});
  } else if (cubic && n1_ == 48) {
    launch_fft(std::integral_constant<int, 48>{});
  } else if (cubic && n1_ == 64) {
    launch_fft(std::integral_constant<int, 64>{});
  } else {
    size_t lds_plane =
        sizeof(double2) * (n1_ + n2_ + nslot_ + (size_t)nc1_ * n2_ + (size_t)n2_ * n1_);
    SGW_REQUIRE(lds_plane <= 64 * 1024, SGW_E_ARG, "pw operator: FFT plane does not fit in LDS");
    KScope ks("pw_plane", st, 2);
    k_pw_plane<<<dim3(n3_, nvec), kBlock, lds_plane, st>>>(d, in, in_stride, d_group, d_active,
                                                          wk.S);
  }
  SGW_HIP(hipGetLastError());
  {
    const size_t lds_asm = sizeof(double2) * ((size_t)nproj_ * kVT + n3_);
    KScope ks("pw_assemble", st, 2);
    // 64 plane waves x 4 vector groups per 256-lane block by default: on the
    // 48^3 batch 21.9 us per 512 vectors against 26.4 with 32 plane waves x 8
    // groups and 33.2 with 16 x 16, and the one the autotuner keeps there;
    // the autotuner tries 32 and 16 as well for each size class of the batch
    // (not when the output overwrites the input, which a second launch would
    // read)
    constexpr int NT = kAsmThreads;
    auto launch_asm = [&](auto chunk) {
      constexpr int CH = decltype(chunk)::value;
      k_pw_assemble2<CH, NT><<<dim3((unsigned)((ld + CH - 1) / CH), ngroup_, wk.plan_tiles), NT,
                               lds_asm, st>>>(d, in, in_stride, out, out_stride, d_sigma, d_active,
                                              wk.plan_vec, wk.plan_cnt, wk.plan_max,
                                              wk.plan_tiles, nchunk, wk.becp, wk.S, locv);
    };
    const bool disjoint = out + (int64_t)(nvec - 1) * out_stride + ld <= in ||
                          in + (int64_t)(nvec - 1) * in_stride + ld <= out;
    const std::string key = "assemble/n" + std::to_string(size_class(nvec)) + "/g" +
                            std::to_string(ngroup_) + "/p" + std::to_string(nproj_);
    autotune(key, disjoint ? 3 : 1, st, [&](int c) {
      if (c == 0) launch_asm(std::integral_constant<int, kAsmChunk>{});
      else if (c == 1) launch_asm(std::integral_constant<int, 32>{});
      else launch_asm(std::integral_constant<int, 16>{});
    });
  }
  SGW_HIP(hipGetLastError());
}

void PwOperator::drho_accumulate(hipStream_t st, int group, int nbnd, int nshift,
                                 const double2 *dpsi, int64_t stride, const double2 *Pr,
                                 double wgt, double2 *D, bool accumulate) {
  const int nv = nbnd * nshift;
  if (nv <= 0) return;
  SGW_REQUIRE(group >= 0 && group < ngroup_, SGW_E_ARG, "pw operator: group index out of range");
  const bool cubic = n1_ == n2_ && n2_ == n3_ && columns_contiguous_;
  SGW_REQUIRE(cubic && (n1_ == 24 || n1_ == 32 || n1_ == 36 || n1_ == 48), SGW_E_ARG,
              "pw operator: the density response needs a cubic 24/32/36/48 box");
  const size_t lds = sizeof(double2) * (nslot_ + (size_t)nc1_ * n1_);
  SGW_REQUIRE(nshift < 65536, SGW_E_ARG, "pw operator: too many shifts");
  ensure_scratch(nv);
  std::vector<int32_t> hg(nv, group);
  DevBuf<int32_t> dg(nv);
  dg.upload(hg.data(), nv, st);
  PwDev d;
  d.n1 = n1_; d.n2 = n2_; d.n3 = n3_;
  d.M1 = M1_; d.M2 = M2_; d.nc1 = nc1_; d.nc2 = nc2_; d.nslot = nslot_;
  d.ld = ld; d.nproj = nproj_;
  d.vloc = vloc_.p; d.vloc_zxy = vloc_zxy_.p; d.npw = d_npw_.p;
  d.col_start = col_start_.p; d.col_cnt = col_cnt_.p;
  d.sorted_idx = sidx_.p; d.sorted_m3 = sm3_.p; d.slot_of = slot_of_.p; d.m3_of = m3_of_.p;
  d.col_lo = col_lo_.p;
  d.ncs = nslot_; d.col_info = gr_info_.p;
  d.row_start = gr_row_start_.p; d.row_lo = gr_row_lo_.p; d.row_cnt = gr_row_cnt_.p;
  d.kin = kin_.p; d.proj = proj_.p; d.dcoef = dcoef_.p;
  d.tw1 = tw_.p; d.tw2 = tw_.p + n1_; d.tw3 = tw_.p + n1_ + n2_;
  d.inv_nbox = 1.0 / ((double)n1_ * n2_ * n3_);
  const dim3 gcol((unsigned)((nslot_ + 63) / 64), nv);
  auto run = [&](auto tag) {
    constexpr int NB = decltype(tag)::value;
    k_pw_zinv<NB><<<gcol, 64, 0, st>>>(d, dpsi, stride, dg.p, nullptr, T_.p);
    if (split_planes_ && nc1_ <= 64) {
      // transforms split over lane pairs, D of a column in registers
      with_band<NB>(band_planes_ && M1_ == M2_, M1_, [&](auto mb) {
        k_pw_plane_drho_split<NB, decltype(mb)::value>
            <<<dim3(NB, nshift), 128, sizeof(double2) * (size_t)nc1_ * NB, st>>>(
                d, T_.p, nbnd, nshift, Pr, wgt, D, accumulate ? 1 : 0);
      });
    } else {
      k_pw_plane_drho<NB><<<dim3(NB, nshift), 64, lds, st>>>(d, T_.p, nbnd, nshift, Pr, wgt, D,
                                                             accumulate ? 1 : 0);
    }
  };
  switch (n1_) {
    case 24: run(std::integral_constant<int, 24>{}); break;
    case 32: run(std::integral_constant<int, 32>{}); break;
    case 36: run(std::integral_constant<int, 36>{}); break;
    default: run(std::integral_constant<int, 48>{}); break;
  }
  SGW_HIP(hipGetLastError());
  SGW_HIP(hipStreamSynchronize(st));  // dg is released on return
}

void PwOperator::grad_accumulate(hipStream_t st, int np, const double2 *x, const double2 *y,
                                 int64_t stride, const int32_t *d_group, const double *d_w,
                                 int nb, double *G, bool accumulate) {
  if (np <= 0) return;
  SGW_REQUIRE(d_group != nullptr || ngroup_ == 1, SGW_E_ARG,
              "pw operator: a group index per vector is required with several groups");
  const bool cubic = n1_ == n2_ && n2_ == n3_ && columns_contiguous_;
  SGW_REQUIRE(cubic && (n1_ == 24 || n1_ == 32 || n1_ == 36 || n1_ == 48), SGW_E_ARG,
              "pw operator: the V_loc gradient needs a cubic 24/32/36/48 box");
  const size_t lds = sizeof(double2) * (nslot_ + (size_t)nc1_ * n1_ + (size_t)n1_ * n1_);
  SGW_REQUIRE(lds <= 64 * 1024, SGW_E_ARG, "pw operator: gradient plane does not fit in LDS");
  SGW_REQUIRE(nb >= 1 && nb < 65536, SGW_E_ARG, "pw operator: bad gradient partial count");
  ensure_scratch(np);
  PwDev d;
  d.n1 = n1_; d.n2 = n2_; d.n3 = n3_;
  d.M1 = M1_; d.M2 = M2_; d.nc1 = nc1_; d.nc2 = nc2_; d.nslot = nslot_;
  d.ld = ld; d.nproj = nproj_;
  d.vloc = vloc_.p; d.vloc_zxy = vloc_zxy_.p; d.npw = d_npw_.p;
  d.col_start = col_start_.p; d.col_cnt = col_cnt_.p;
  d.sorted_idx = sidx_.p; d.sorted_m3 = sm3_.p; d.slot_of = slot_of_.p; d.m3_of = m3_of_.p;
  d.col_lo = col_lo_.p;
  d.ncs = nslot_; d.col_info = gr_info_.p;
  d.row_start = gr_row_start_.p; d.row_lo = gr_row_lo_.p; d.row_cnt = gr_row_cnt_.p;
  d.kin = kin_.p; d.proj = proj_.p; d.dcoef = dcoef_.p;
  d.tw1 = tw_.p; d.tw2 = tw_.p + n1_; d.tw3 = tw_.p + n1_ + n2_;
  d.inv_nbox = 1.0 / ((double)n1_ * n2_ * n3_);
  const dim3 gcol((unsigned)((nslot_ + 63) / 64), np);
  auto run = [&](auto tag) {
    constexpr int NB = decltype(tag)::value;
    k_pw_zinv<NB><<<gcol, 64, 0, st>>>(d, x, stride, d_group, nullptr, T_.p);
    k_pw_zinv<NB><<<gcol, 64, 0, st>>>(d, y, stride, d_group, nullptr, S_.p);
    k_pw_plane_grad<NB><<<dim3(NB, nb), 64, lds, st>>>(d, T_.p, S_.p, np, d_w, G,
                                                       accumulate ? 1 : 0);
  };
  switch (n1_) {
    case 24: run(std::integral_constant<int, 24>{}); break;
    case 32: run(std::integral_constant<int, 32>{}); break;
    case 36: run(std::integral_constant<int, 36>{}); break;
    default: run(std::integral_constant<int, 48>{}); break;
  }
  SGW_HIP(hipGetLastError());
}

void PwOperator::prepare(const int32_t *h_group, int nvec) {
  std::vector<int32_t> key(h_group ? h_group : nullptr, h_group ? h_group + nvec : nullptr);
  if (!h_group) key.assign(nvec, 0);
  if (plan_nvec_ == nvec && key == plan_key_) return;
  std::vector<std::vector<int32_t>> lists(ngroup_);
  for (int v = 0; v < nvec; ++v) {
    SGW_REQUIRE(key[v] >= 0 && key[v] < ngroup_, SGW_E_ARG, "pw operator: group index out of range");
    lists[key[v]].push_back(v);
  }
  int maxper = 1;
  for (auto &l : lists) maxper = std::max<int>(maxper, (int)l.size());
  std::vector<int32_t> vec((size_t)ngroup_ * maxper, 0), cnt(ngroup_);
  for (int g = 0; g < ngroup_; ++g) {
    cnt[g] = (int32_t)lists[g].size();
    std::copy(lists[g].begin(), lists[g].end(), vec.begin() + (size_t)g * maxper);
  }
  plan_vec_.alloc(vec.size());
  plan_vec_.upload(vec.data(), vec.size());
  plan_cnt_.alloc(cnt.size());
  plan_cnt_.upload(cnt.data(), cnt.size());
  plan_max_ = maxper;
  plan_tiles_ = (maxper + kVT - 1) / kVT;
  const int nchunk = (int)((ld + kChunk - 1) / kChunk);
  const size_t need = (size_t)ngroup_ * plan_tiles_ * nchunk * std::max(nproj_, 1) * kVT;
  if (part_.n < need) part_.alloc(need);
  const size_t needb = (size_t)ngroup_ * plan_tiles_ * std::max(nproj_, 1) * kVT;
  if (becp_.n < needb) becp_.alloc(needb);
  if (nproj_ == 0) becp_.zero();
  ensure_scratch(nvec);
  SGW_HIP(hipDeviceSynchronize());
  plan_key_ = std::move(key);
  plan_nvec_ = nvec;
}


// ---------------------------------------------------------------------------
// application without device allocation
// ---------------------------------------------------------------------------
namespace {

constexpr int kNoAllocMaxDevices = 64;
constexpr int kNoAllocMaxVec = 1024;
constexpr int kNoAllocMaxGroups = 256;
constexpr int kNoAllocMaxPlan = 16384;  // ngroup x (largest group's vector count)

std::mutex g_noalloc_mu;
void *g_noalloc_ptr[kNoAllocMaxDevices] = {};

// device address of `scratch` on the current device (looked up once per device)
unsigned char *noalloc_scratch() {
  int dev = 0;
  SGW_HIP(hipGetDevice(&dev));
  SGW_REQUIRE(dev >= 0 && dev < kNoAllocMaxDevices, SGW_E_ARG,
              "pw operator no-alloc: device index beyond the supported range");
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void *p = nullptr;
    SGW_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    SGW_REQUIRE(p != nullptr, SGW_E_STATE, "pw operator no-alloc: scratch symbol has no address");
    g_noalloc_ptr[dev] = p;
  }
  return static_cast<unsigned char *>(g_noalloc_ptr[dev]);
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

void PwOperator::apply_noalloc(hipStream_t st, int nvec, const int32_t *d_group,
                               const double2 *d_sigma, const double2 *in, int64_t ldx,
                               double2 *out) {
  SGW_REQUIRE(nvec >= 1 && nvec <= kNoAllocMaxVec, SGW_E_ARG,
              "pw operator no-alloc: nvec must be in [1, 1024]");
  SGW_REQUIRE(in && out && d_sigma, SGW_E_ARG, "pw operator no-alloc: null argument");
  SGW_REQUIRE(ldx >= ld, SGW_E_SHAPE, "pw operator no-alloc: ldx < operator ld");
  SGW_REQUIRE(d_group != nullptr || ngroup_ == 1, SGW_E_ARG,
              "pw operator no-alloc: a group index per vector is required with several groups");
  SGW_REQUIRE(ngroup_ <= kNoAllocMaxGroups, SGW_E_ARG,
              "pw operator no-alloc: more than 256 k+q groups");
  unsigned char *base = noalloc_scratch();
  // group of every vector; the wait also orders this call after earlier
  // work on the stream, which may still read the scratch array
  int32_t grp[kNoAllocMaxVec];
  if (d_group)
    SGW_HIP(hipMemcpyAsync(grp, d_group, sizeof(int32_t) * nvec, hipMemcpyDeviceToHost, st));
  SGW_HIP(hipStreamSynchronize(st));
  if (!d_group) std::fill(grp, grp + nvec, 0);
  // vectors of each group, in increasing order: plan[g * maxper + j]
  int32_t cnt[kNoAllocMaxGroups] = {};
  for (int v = 0; v < nvec; ++v) {
    SGW_REQUIRE(grp[v] >= 0 && grp[v] < ngroup_, SGW_E_ARG,
                "pw operator no-alloc: group index out of range");
    ++cnt[grp[v]];
  }
  int maxper = 1;
  for (int g = 0; g < ngroup_; ++g) maxper = std::max<int>(maxper, cnt[g]);
  SGW_REQUIRE((int64_t)ngroup_ * maxper + ngroup_ <= kNoAllocMaxPlan, SGW_E_SHAPE,
              "pw operator no-alloc: vector lists exceed 16384 entries (groups x largest group)");
  int32_t plan[kNoAllocMaxPlan];
  std::fill(plan, plan + (size_t)ngroup_ * maxper, 0);
  int32_t fill[kNoAllocMaxGroups] = {};
  for (int v = 0; v < nvec; ++v) plan[(size_t)grp[v] * maxper + fill[grp[v]]++] = v;
  std::copy(cnt, cnt + ngroup_, plan + (size_t)ngroup_ * maxper);
  // carve the work buffers
  const int tiles = (maxper + kVT - 1) / kVT;
  const int nchunk = (int)((ld + kChunk - 1) / kChunk);
  const size_t np = std::max(nproj_, 1);
  const size_t b_plan = align256(sizeof(int32_t) * ((size_t)ngroup_ * maxper + ngroup_));
  const size_t b_part = align256(sizeof(double2) * (size_t)ngroup_ * tiles * nchunk * np * kVT);
  const size_t b_becp = align256(sizeof(double2) * (size_t)ngroup_ * tiles * np * kVT);
  const size_t b_box = align256(sizeof(double2) * (size_t)nvec * nslot_ * n3_);
  const size_t b_loc = align256(sizeof(double2) * (size_t)nvec * ld);
  const size_t b_vlist = align256(sizeof(int32_t) * (size_t)nvec);
  const size_t total = b_plan + b_part + b_becp + 2 * b_box + b_loc + b_vlist;
  SGW_REQUIRE(total <= (size_t)SGW_NOALLOC_SCRATCH_BYTES, SGW_E_SHAPE,
              "pw operator no-alloc: this batch needs " + std::to_string(total) +
                  " bytes of work space, more than the fixed " +
                  std::to_string((size_t)SGW_NOALLOC_SCRATCH_BYTES) +
                  "-byte scratch array; use fewer vectors per call");
  unsigned char *p = base;
  int32_t *d_plan = reinterpret_cast<int32_t *>(p);
  p += b_plan;
  PwWork w{};
  w.part = reinterpret_cast<double2 *>(p);
  p += b_part;
  w.becp = reinterpret_cast<double2 *>(p);
  p += b_becp;
  w.T = reinterpret_cast<double2 *>(p);
  p += b_box;
  w.S = reinterpret_cast<double2 *>(p);
  p += b_box;
  w.loc = reinterpret_cast<double2 *>(p);
  p += b_loc;
  w.vlist = reinterpret_cast<int32_t *>(p);
  w.plan_vec = d_plan;
  w.plan_cnt = d_plan + (size_t)ngroup_ * maxper;
  w.plan_max = maxper;
  w.plan_tiles = tiles;
  SGW_HIP(hipMemcpyAsync(d_plan, plan, sizeof(int32_t) * ((size_t)ngroup_ * maxper + ngroup_),
                         hipMemcpyHostToDevice, st));
  if (nproj_ == 0) SGW_HIP(hipMemsetAsync(w.becp, 0, b_becp, st));
  // the host lists above live on this stack frame: the copy must be done
  SGW_HIP(hipStreamSynchronize(st));
  apply_work(st, nvec, in, ldx, out, ldx, d_sigma, nullptr, d_group, w);
}

}  // namespace sgw
