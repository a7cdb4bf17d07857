// Register-resident complex FFTs of compile-time length N (mixed radix 4/2/3,
// decimation in time).  Every index is a compile-time constant after
// unrolling, so the data stay in VGPRs and the twiddles fold into literal
// constants from twiddles.h.  SIGN = +1 computes sum_n x[n] e^{+2 pi i nk/N}
// (QE invfft direction), SIGN = -1 the conjugate transform (fwfft, unscaled).
// The element type C is double2 (FP64) or float2 (FP32; twiddles rounded once
// from the FP64 table).
#pragma once

#include <hip/hip_runtime.h>

#include "twiddles.h"

namespace sgw {

template <typename C> struct CplxT;
template <> struct CplxT<double2> {
  using S = double;
  __host__ __device__ static double2 make(double a, double b) { return make_double2(a, b); }
};
template <> struct CplxT<float2> {
  using S = float;
  __host__ __device__ static float2 make(float a, float b) { return make_float2(a, b); }
};

template <int NTOP, int SIGN> struct Root {
  // e^{SIGN 2 pi i e / n} for n dividing NTOP
  template <int E, int n> __host__ __device__ static constexpr double re() {
    return tw::Table<NTOP>::re[((E % n) * (NTOP / n)) % NTOP];
  }
  template <int E, int n> __host__ __device__ static constexpr double im() {
    return SIGN * tw::Table<NTOP>::im[((E % n) * (NTOP / n)) % NTOP];
  }
};

template <typename C> __host__ __device__ inline C rf_add(C a, C b) {
  return CplxT<C>::make(a.x + b.x, a.y + b.y);
}
template <typename C> __host__ __device__ inline C rf_sub(C a, C b) {
  return CplxT<C>::make(a.x - b.x, a.y - b.y);
}
// a * (SIGN i)
template <int SIGN, typename C> __host__ __device__ inline C rf_muli(C a) {
  return SIGN > 0 ? CplxT<C>::make(-a.y, a.x) : CplxT<C>::make(a.y, -a.x);
}
template <typename C, typename S>
__host__ __device__ inline C rf_mul(C a, S c, S s) {
  return CplxT<C>::make(a.x * c - a.y * s, a.x * s + a.y * c);
}

template <int N, int NTOP, int SIGN, typename C> struct RegFft;

template <int NTOP, int SIGN, typename C> struct RegFft<1, NTOP, SIGN, C> {
  __host__ __device__ static void run(const C (&x)[1], C (&y)[1]) { y[0] = x[0]; }
};

namespace detail {
template <int N> struct Radix {
  static constexpr int R = (N % 4 == 0) ? 4 : (N % 2 == 0) ? 2 : (N % 3 == 0) ? 3 : 0;
  static_assert(R != 0, "register FFT supports lengths 2^a 3^b");
};

// twiddle multiply v * W_N^(RK) with compile-time exponent
template <int RK, int N, int NTOP, int SIGN, typename C>
__host__ __device__ inline C twiddle(C v) {
  if constexpr (RK % N == 0) {
    return v;
  } else {
    using S = typename CplxT<C>::S;
    constexpr S c = (S)Root<NTOP, SIGN>::template re<RK, N>();
    constexpr S s = (S)Root<NTOP, SIGN>::template im<RK, N>();
    return rf_mul(v, c, s);
  }
}

template <int R, int SIGN> struct Butterfly;
template <int SIGN> struct Butterfly<2, SIGN> {
  template <typename C> __host__ __device__ static void run(const C (&v)[2], C (&y)[2]) {
    y[0] = rf_add(v[0], v[1]);
    y[1] = rf_sub(v[0], v[1]);
  }
};
template <int SIGN> struct Butterfly<4, SIGN> {
  template <typename C> __host__ __device__ static void run(const C (&v)[4], C (&y)[4]) {
    C t0 = rf_add(v[0], v[2]), t1 = rf_sub(v[0], v[2]);
    C t2 = rf_add(v[1], v[3]), t3 = rf_muli<SIGN>(rf_sub(v[1], v[3]));
    y[0] = rf_add(t0, t2);
    y[2] = rf_sub(t0, t2);
    y[1] = rf_add(t1, t3);
    y[3] = rf_sub(t1, t3);
  }
};
template <int SIGN> struct Butterfly<3, SIGN> {
  template <typename C> __host__ __device__ static void run(const C (&v)[3], C (&y)[3]) {
    using S = typename CplxT<C>::S;
    constexpr S h = (S)0.8660254037844386;  // sqrt(3)/2
    constexpr S half = (S)0.5;
    C t = rf_add(v[1], v[2]), d = rf_sub(v[1], v[2]);
    C m = CplxT<C>::make(v[0].x - half * t.x, v[0].y - half * t.y);
    C r = rf_muli<SIGN>(CplxT<C>::make(h * d.x, h * d.y));
    y[0] = rf_add(v[0], t);
    y[1] = rf_add(m, r);
    y[2] = rf_sub(m, r);
  }
};

template <int N, int NTOP, int SIGN, int R, int M, int K, typename C>
__host__ __device__ inline void combine(const C (&sub)[R][M], C (&y)[N]) {
  if constexpr (K < M) {
    C v[R], o[R];
    [&]<int... r>(std::integer_sequence<int, r...>) {
      ((v[r] = twiddle<r * K, N, NTOP, SIGN>(sub[r][K])), ...);
    }(std::make_integer_sequence<int, R>{});
    Butterfly<R, SIGN>::run(v, o);
    [&]<int... q>(std::integer_sequence<int, q...>) { ((y[K + M * q] = o[q]), ...); }(
        std::make_integer_sequence<int, R>{});
    combine<N, NTOP, SIGN, R, M, K + 1>(sub, y);
  }
}
}  // namespace detail

template <int N, int NTOP, int SIGN, typename C> struct RegFft {
  static constexpr int R = detail::Radix<N>::R;
  static constexpr int M = N / R;
  __host__ __device__ static void run(const C (&x)[N], C (&y)[N]) {
    C sin_[R][M], sout[R][M];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int m = 0; m < M; ++m) sin_[r][m] = x[R * m + r];
#pragma unroll
    for (int r = 0; r < R; ++r) RegFft<M, NTOP, SIGN, C>::run(sin_[r], sout[r]);
    detail::combine<N, NTOP, SIGN, R, M, 0>(sout, y);
  }
};

// In-place convenience wrapper.
template <int N, int SIGN, typename C> __host__ __device__ inline void reg_fft(C (&x)[N]) {
  C y[N];
  RegFft<N, N, SIGN, C>::run(x, y);
#pragma unroll
  for (int i = 0; i < N; ++i) x[i] = y[i];
}

}  // namespace sgw
