// Length-N FFTs split over a lane pair (lanes 2r and 2r+1 of a wavefront),
// each lane holding N/2 complex values (double2 or float2): one radix-2 step
// across the pair, exchanged with a DPP lane swap, and one register FFT of
// length N/2 per lane.  Halving the values per lane halves the registers a
// transform needs, so a kernel built on these runs more wavefronts per SIMD
// than one on reg_fft<N>.
//
// Layouts (Q = N/4, H = N/2, `odd` = lane index & 1):
//   split_dit: in  b[n] = x[2n + odd]                      (decimated)
//              out b[j] = X[j + Q odd], b[Q + j] = X[H + j + Q odd]
//   split_dif: in  b[j] = X[j + Q odd], b[Q + j] = X[H + j + Q odd]
//              out b[k] = F[2k + odd]                      (decimated)
// with X (F) the transform of x (X) in the direction SIGN as in reg_fft:
// sum_n x[n] e^{SIGN 2 pi i n k / N}, unscaled.  Both lanes of a pair must
// execute these together (the swap reads the partner's registers).
#pragma once

#include <hip/hip_runtime.h>

#include <utility>

#include "regfft.h"

namespace sgw {

// the partner lane's value (lane ^ 1): DPP quad permutation [1, 0, 3, 2]
__device__ inline double lane_swap(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, 0xB1, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0xB1, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ inline double2 lane_swap(double2 v) {
  return make_double2(lane_swap(v.x), lane_swap(v.y));
}
__device__ inline float lane_swap(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));
}
__device__ inline float2 lane_swap(float2 v) {
  return make_float2(lane_swap(v.x), lane_swap(v.y));
}
__device__ inline double2 neg(double2 v) { return make_double2(-v.x, -v.y); }
__device__ inline float2 neg(float2 v) { return make_float2(-v.x, -v.y); }

namespace detail {
// v[k] *= e^{SIGN 2 pi i ((k + Q) mod H) / N}, k in the sequence
template <int N, int SIGN, typename C, int... K>
__device__ inline void twiddle_rot(C *v, std::integer_sequence<int, K...>) {
  constexpr int H = N / 2, Q = N / 4;
  ((v[K] = twiddle<(K + Q) % H, N, N, SIGN>(v[K])), ...);
}
}  // namespace detail

// The odd lane works on its half rotated by Q = H/2 positions, which makes
// the register it sends and the one it keeps the same as the even lane's: no
// per-lane selects around the swap.  A rotation by H/2 of a length-H
// transform's input or output is a factor (-1)^n on the other side.
template <int N, int SIGN, typename C> __device__ inline void split_dit(C (&b)[N / 2], bool odd) {
  static_assert(N % 4 == 0, "split FFT needs N divisible by 4");
  constexpr int H = N / 2, Q = N / 4;
  if (odd) {  // (-1)^n: the odd lane's transform comes out rotated by Q
#pragma unroll
    for (int n = 1; n < H; n += 2) b[n] = neg(b[n]);
  }
  C y[H];
  RegFft<H, N, SIGN, C>::run(b, y);  // even: E[k]; odd: O[(k + Q) mod H]
  if (odd) detail::twiddle_rot<N, SIGN>(y, std::make_integer_sequence<int, H>{});
  // even lane: y[j] = E[j], receives O'[j]          -> X[j], X[H + j]
  // odd lane:  y[j] = O'[Q + j], receives E[Q + j]  -> X[Q + j], -X[H + Q + j]
#pragma unroll
  for (int j = 0; j < Q; ++j) {
    const C recv = lane_swap(y[Q + j]);
    b[j] = rf_add(y[j], recv);
    b[Q + j] = rf_sub(y[j], recv);
  }
  if (odd) {
#pragma unroll
    for (int j = 0; j < Q; ++j) b[Q + j] = neg(b[Q + j]);
  }
}

template <int N, int SIGN, typename C> __device__ inline void split_dif(C (&b)[N / 2], bool odd) {
  static_assert(N % 4 == 0, "split FFT needs N divisible by 4");
  constexpr int H = N / 2, Q = N / 4;
  if (odd) {  // the odd lane forms (difference, sum) where the even lane forms (sum, difference)
#pragma unroll
    for (int j = 0; j < Q; ++j) b[Q + j] = neg(b[Q + j]);
  }
  // even lane: x = the sums over all x (its own, then the partner's);
  // odd lane: the differences rotated by Q (its own at x = Q + j, then the
  // partner's at x = j), twiddled by W^x
  C x[H];
#pragma unroll
  for (int j = 0; j < Q; ++j) {
    x[j] = rf_add(b[j], b[Q + j]);
    x[Q + j] = lane_swap(rf_sub(b[j], b[Q + j]));
  }
  if (odd) detail::twiddle_rot<N, SIGN>(x, std::make_integer_sequence<int, H>{});
  RegFft<H, N, SIGN, C>::run(x, b);
  if (odd) {  // undo the rotation: (-1)^k
#pragma unroll
    for (int k = 1; k < H; k += 2) b[k] = neg(b[k]);
  }
}

__device__ inline double2 pick(bool c, double2 a, double2 b) {
  return make_double2(c ? a.x : b.x, c ? a.y : b.y);
}
__device__ inline float2 pick(bool c, float2 a, float2 b) {
  return make_float2(c ? a.x : b.x, c ? a.y : b.y);
}

namespace detail {
// v[k] *= e^{SIGN 2 pi i k / N}, k in the sequence
template <int N, int SIGN, typename C, int... K>
__device__ inline void twiddle_seq(C *v, std::integer_sequence<int, K...>) {
  ((v[K] = twiddle<K, N, N, SIGN>(v[K])), ...);
}
}  // namespace detail

// Select-based forms: each lane picks what it sends and keeps (per-lane
// selects around the swap, no rotation).  Same layouts and results.
template <int N, int SIGN, typename C>
__device__ inline void split_dit_sel(C (&b)[N / 2], bool odd) {
  static_assert(N % 4 == 0, "split FFT needs N divisible by 4");
  constexpr int H = N / 2, Q = N / 4;
  C y[H];
  RegFft<H, N, SIGN, C>::run(b, y);  // E (even lane) or O (odd lane)
  if (odd) detail::twiddle_seq<N, SIGN>(y, std::make_integer_sequence<int, H>{});
  // even lane: e = E[j], o = O'[j] from the partner -> X[j], X[H + j]
  // odd lane:  e = E[Q + j] from the partner, o = O'[Q + j] -> X[Q + j], X[H + Q + j]
#pragma unroll
  for (int j = 0; j < Q; ++j) {
    const C recv = lane_swap(pick(odd, y[j], y[Q + j]));
    const C e = pick(odd, recv, y[j]), o = pick(odd, y[Q + j], recv);
    b[j] = rf_add(e, o);
    b[Q + j] = rf_sub(e, o);
  }
}

template <int N, int SIGN, typename C>
__device__ inline void split_dif_sel(C (&b)[N / 2], bool odd) {
  static_assert(N % 4 == 0, "split FFT needs N divisible by 4");
  constexpr int H = N / 2, Q = N / 4;
  C a[Q], d[Q];
#pragma unroll
  for (int j = 0; j < Q; ++j) {
    a[j] = rf_add(b[j], b[Q + j]);
    d[j] = rf_sub(b[j], b[Q + j]);
  }
  // d *= W^(j + Q odd), W^Q = SIGN i
  detail::twiddle_seq<N, SIGN>(d, std::make_integer_sequence<int, Q>{});
  if (odd) {
#pragma unroll
    for (int j = 0; j < Q; ++j) d[j] = rf_muli<SIGN>(d[j]);
  }
  // even lane keeps a over all x (its half, then the partner's), odd lane d
  C x[H];
#pragma unroll
  for (int j = 0; j < Q; ++j) {
    const C recv = lane_swap(pick(odd, a[j], d[j]));
    x[j] = pick(odd, recv, a[j]);
    x[Q + j] = pick(odd, d[j], recv);
  }
  RegFft<H, N, SIGN, C>::run(x, b);
}

}  // namespace sgw
