// Host check of the register FFTs (sternheimergw_amd/csrc/regfft.h) against a
// long-double DFT, FP64 and FP32, both signs, every length the kernels use.
// Prints "N sign precision relerr" lines; exit status 1 if a bound fails.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <utility>

#include "regfft.h"
using namespace sgw;

template <typename C> C mkc(double a, double b) { return CplxT<C>::make(a, b); }

template <int N, int SIGN, typename C> double check(unsigned seed) {
  srand(seed);
  C x[N];
  std::complex<long double> ref[N];
  for (int i = 0; i < N; ++i) x[i] = mkc<C>(rand() / (double)RAND_MAX - 0.5, rand() / (double)RAND_MAX - 0.5);
  const long double pi = 3.14159265358979323846264338327950288L;
  for (int k = 0; k < N; ++k) {
    long double sr = 0, si = 0;
    for (int n = 0; n < N; ++n) {
      long double ang = SIGN * 2.0L * pi * (long double)((n * k) % N) / N;
      sr += (long double)x[n].x * cosl(ang) - (long double)x[n].y * sinl(ang);
      si += (long double)x[n].x * sinl(ang) + (long double)x[n].y * cosl(ang);
    }
    ref[k] = std::complex<long double>(sr, si);
  }
  reg_fft<N, SIGN>(x);
  long double err = 0, mag = 0;
  for (int k = 0; k < N; ++k) {
    err = fmaxl(err, std::abs(std::complex<long double>(x[k].x, x[k].y) - ref[k]));
    mag = fmaxl(mag, std::abs(ref[k]));
  }
  return (double)(err / mag);
}

int fails = 0;
template <int N> void both(unsigned seed) {
  const double e64p = check<N, 1, double2>(seed), e64m = check<N, -1, double2>(seed + 1);
  const double e32p = check<N, 1, float2>(seed + 2), e32m = check<N, -1, float2>(seed + 3);
  printf("%d +1 fp64 %.2e\n%d -1 fp64 %.2e\n%d +1 fp32 %.2e\n%d -1 fp32 %.2e\n", N, e64p, N, e64m,
         N, e32p, N, e32m);
  if (e64p > 1e-15 || e64m > 1e-15) ++fails;
  if (e32p > 1e-6 || e32m > 1e-6) ++fails;
}

int main() {
  both<16>(1);
  both<18>(3);
  both<24>(5);
  both<27>(7);
  both<32>(9);
  both<36>(13);
  both<48>(17);
  both<64>(21);
  return fails ? 1 : 0;
}
