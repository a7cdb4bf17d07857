// Register-resident complex FFTs of compile-time length N (mixed radix 4/2/3,
// decimation in time).  Every index is a compile-time constant after
// unrolling, so the data stay in VGPRs and the twiddles fold into literal
// constants from twiddles.h.  SIGN = +1 computes sum_n x[n] e^{+2 pi i nk/N}
// (QE invfft direction), SIGN = -1 the conjugate transform (fwfft, unscaled).
#pragma once

#include <hip/hip_runtime.h>

#include "twiddles.h"

namespace sgw {

template <int NTOP, int SIGN> struct Root {
  // e^{SIGN 2 pi i e / n} for n dividing NTOP
  template <int E, int n> __host__ __device__ static constexpr double re() {
    return tw::Table<NTOP>::re[((E % n) * (NTOP / n)) % NTOP];
  }
  template <int E, int n> __host__ __device__ static constexpr double im() {
    return SIGN * tw::Table<NTOP>::im[((E % n) * (NTOP / n)) % NTOP];
  }
};

__host__ __device__ inline double2 rf_add(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline double2 rf_sub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// a * (SIGN i)
template <int SIGN> __host__ __device__ inline double2 rf_muli(double2 a) {
  return SIGN > 0 ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}
__host__ __device__ inline double2 rf_mul(double2 a, double c, double s) {
  return make_double2(a.x * c - a.y * s, a.x * s + a.y * c);
}

template <int N, int NTOP, int SIGN> struct RegFft;

template <int NTOP, int SIGN> struct RegFft<1, NTOP, SIGN> {
  __host__ __device__ static void run(const double2 (&x)[1], double2 (&y)[1]) { y[0] = x[0]; }
};

namespace detail {
template <int N> struct Radix {
  static constexpr int R = (N % 4 == 0) ? 4 : (N % 2 == 0) ? 2 : (N % 3 == 0) ? 3 : 0;
  static_assert(R != 0, "register FFT supports lengths 2^a 3^b");
};

// twiddle multiply v * W_N^(RK) with compile-time exponent
template <int RK, int N, int NTOP, int SIGN>
__host__ __device__ inline double2 twiddle(double2 v) {
  if constexpr (RK % N == 0) {
    return v;
  } else {
    return rf_mul(v, Root<NTOP, SIGN>::template re<RK, N>(), Root<NTOP, SIGN>::template im<RK, N>());
  }
}

template <int R, int SIGN> struct Butterfly;
template <int SIGN> struct Butterfly<2, SIGN> {
  __host__ __device__ static void run(const double2 (&v)[2], double2 (&y)[2]) {
    y[0] = rf_add(v[0], v[1]);
    y[1] = rf_sub(v[0], v[1]);
  }
};
template <int SIGN> struct Butterfly<4, SIGN> {
  __host__ __device__ static void run(const double2 (&v)[4], double2 (&y)[4]) {
    double2 t0 = rf_add(v[0], v[2]), t1 = rf_sub(v[0], v[2]);
    double2 t2 = rf_add(v[1], v[3]), t3 = rf_muli<SIGN>(rf_sub(v[1], v[3]));
    y[0] = rf_add(t0, t2);
    y[2] = rf_sub(t0, t2);
    y[1] = rf_add(t1, t3);
    y[3] = rf_sub(t1, t3);
  }
};
template <int SIGN> struct Butterfly<3, SIGN> {
  __host__ __device__ static void run(const double2 (&v)[3], double2 (&y)[3]) {
    constexpr double h = 0.8660254037844386;  // sqrt(3)/2
    double2 t = rf_add(v[1], v[2]), d = rf_sub(v[1], v[2]);
    double2 m = make_double2(v[0].x - 0.5 * t.x, v[0].y - 0.5 * t.y);
    double2 r = rf_muli<SIGN>(make_double2(h * d.x, h * d.y));
    y[0] = rf_add(v[0], t);
    y[1] = rf_add(m, r);
    y[2] = rf_sub(m, r);
  }
};

template <int N, int NTOP, int SIGN, int R, int M, int K>
__host__ __device__ inline void combine(const double2 (&sub)[R][M], double2 (&y)[N]) {
  if constexpr (K < M) {
    double2 v[R], o[R];
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

template <int N, int NTOP, int SIGN> struct RegFft {
  static constexpr int R = detail::Radix<N>::R;
  static constexpr int M = N / R;
  __host__ __device__ static void run(const double2 (&x)[N], double2 (&y)[N]) {
    double2 sin_[R][M], sout[R][M];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int m = 0; m < M; ++m) sin_[r][m] = x[R * m + r];
#pragma unroll
    for (int r = 0; r < R; ++r) RegFft<M, NTOP, SIGN>::run(sin_[r], sout[r]);
    detail::combine<N, NTOP, SIGN, R, M, 0>(sout, y);
  }
};

// In-place convenience wrapper.
template <int N, int SIGN> __host__ __device__ inline void reg_fft(double2 (&x)[N]) {
  double2 y[N];
  RegFft<N, N, SIGN>::run(x, y);
#pragma unroll
  for (int i = 0; i < N; ++i) x[i] = y[i];
}

}  // namespace sgw
