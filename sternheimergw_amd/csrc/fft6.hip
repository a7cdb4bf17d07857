// The Green's function of the self-energy on the device: the assembly of
// green_function (phys/green/src/green.f90:196-218) and the 6-D Fourier
// transforms of fft6.f90 (data/fft/src/fft6.f90:84-345).
//
// A two-point function f(r, r') is one matrix per frequency, Fortran layout
// f(nr, nr_p) (first index fastest), nr = n^3 and nr_p = n_p^3 the points of
// the two correlation boxes.  In reciprocal space f(G, G') fills the block
// (num_g, num_g_p) at the top left.  With the box index maps nl (G) and nl_p
// (G'):
//   invfft6: f(r, r') = 1/Omega sum_{G,G'} e^{-iGr} f(G, G') e^{+iG'r'}
//   fwfft6:  f(G, G') = Omega / (nr nr_p) sum_{r,r'} e^{+iGr} f(r, r') e^{-iG'r'}
// The reference gets the e^{-iGr} of the first index by conjugating around
// QE's invfft / fwfft (fft6.f90:148-162, 323-338); here the box FFT runs
// with the opposite sign directly.  Each transform is two batched box-FFT
// passes with one scratch matrix S[g][r'] (num_g x nr_p) between them:
//   invfft6: S[g] = invfft_p(scatter_p(f(g, :) / Omega))     (rows)
//            f(:, r') = fft_(-)(scatter(S[:, r']))            (columns, in place)
//   fwfft6:  f(:, r') = fft_(+)(f(:, r'))                      (columns, in place)
//            S[g][r'] = f(nl[g], r') / nr
//            f(g, g') = Omega / nr_p fft_(-)(S[g])(nl_p[g'])  (rows)
// Index-map kernels put the contiguous index of both operands on adjacent
// lanes where one side is a gather.
#include <algorithm>

#include "boxfft.h"
#include "fft6.h"
#include "ktimer.h"
#include "sgw_internal.h"

namespace sgw {
namespace {

constexpr int kEB = 256;

unsigned blocks(int64_t count) { return (unsigned)((count + kEB - 1) / kEB); }

__global__ void k_fill0(int64_t count, double2 *__restrict__ d) {
  const int64_t e = (int64_t)blockIdx.x * kEB + threadIdx.x;
  if (e < count) d[e] = mk(0.0, 0.0);
}

// S[g][nl_p[j]] = s * f[j * nr + g], g < num_g, j < num_g_p (S zeroed before)
__global__ void k_rows_scatter(int num_g, int num_g_p, int64_t nr, int64_t nr_p,
                               const int32_t *__restrict__ nl_p, const double2 *__restrict__ f,
                               double s, double2 *__restrict__ S) {
  const int64_t e = (int64_t)blockIdx.x * kEB + threadIdx.x;
  if (e >= (int64_t)num_g * num_g_p) return;
  const int64_t g = e % num_g, j = e / num_g;
  const double2 v = f[j * nr + g];
  S[g * nr_p + nl_p[j]] = mk(s * v.x, s * v.y);
}

// f[r' * nr + nl[g]] = S[g][r'], g < num_g, r' < nr_p (f zeroed before)
__global__ void k_cols_scatter(int num_g, int64_t nr, int64_t nr_p, const int32_t *__restrict__ nl,
                               const double2 *__restrict__ S, double2 *__restrict__ f) {
  const int64_t e = (int64_t)blockIdx.x * kEB + threadIdx.x;
  if (e >= (int64_t)num_g * nr_p) return;
  const int64_t rp = e % nr_p, g = e / nr_p;
  f[rp * nr + nl[g]] = S[g * nr_p + rp];
}

// S[g][r'] = s * f[r' * nr + nl[g]]
__global__ void k_cols_gather(int num_g, int64_t nr, int64_t nr_p, const int32_t *__restrict__ nl,
                              const double2 *__restrict__ f, double s, double2 *__restrict__ S) {
  const int64_t e = (int64_t)blockIdx.x * kEB + threadIdx.x;
  if (e >= (int64_t)num_g * nr_p) return;
  const int64_t rp = e % nr_p, g = e / nr_p;
  const double2 v = f[rp * nr + nl[g]];
  S[e] = mk(s * v.x, s * v.y);
}

// f[j * nr + g] = s * S[g][nl_p[j]]
__global__ void k_rows_gather(int num_g, int num_g_p, int64_t nr, int64_t nr_p,
                              const int32_t *__restrict__ nl_p, const double2 *__restrict__ S,
                              double s, double2 *__restrict__ f) {
  const int64_t e = (int64_t)blockIdx.x * kEB + threadIdx.x;
  if (e >= (int64_t)num_g * num_g_p) return;
  const int64_t g = e % num_g, j = e / num_g;
  const double2 v = S[g * nr_p + nl_p[j]];
  f[j * nr + g] = mk(s * v.x, s * v.y);
}

// green[f][c][r] = x[colsys[c]][f][rowmap[r]] where both maps are >= 0 and
// c < ncol, r < nrow; zero elsewhere (green.f90:184-186, 211-213)
__global__ void k_green_assemble(int nfreq, int64_t nr, int64_t nr_p, int nrow, int ncol,
                                 const int32_t *__restrict__ rowmap,
                                 const int32_t *__restrict__ colsys,
                                 const double2 *__restrict__ x, int nshift, int64_t ld,
                                 double2 *__restrict__ green) {
  const int64_t e = (int64_t)blockIdx.x * kEB + threadIdx.x;
  if (e >= (int64_t)nfreq * nr * nr_p) return;
  const int64_t r = e % nr, c = (e / nr) % nr_p, f = e / (nr * nr_p);
  double2 v = mk(0.0, 0.0);
  if (r < nrow && c < ncol) {
    const int s = colsys[c], i = rowmap[r];
    if (s >= 0 && i >= 0) v = x[((int64_t)s * nshift + f) * ld + i];
  }
  green[e] = v;
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

}  // namespace

void fft6_check(const Fft6Args &a, const char *who) {
  SGW_REQUIRE(a.f && a.nl && a.nl_p && a.nmat >= 0, SGW_E_ARG,
              std::string(who) + ": null argument or negative count");
  SGW_REQUIRE(box_supported(a.n) && box_supported(a.n_p), SGW_E_SHAPE,
              std::string(who) + ": box edges must be " + box_sizes_text());
  const int64_t nr = (int64_t)a.n * a.n * a.n, nr_p = (int64_t)a.n_p * a.n_p * a.n_p;
  SGW_REQUIRE(a.num_g >= 1 && a.num_g <= nr && a.num_g_p >= 1 && a.num_g_p <= nr_p, SGW_E_SHAPE,
              std::string(who) + ": need 1 <= num_g <= nr and 1 <= num_g_p <= nr_p");
  SGW_REQUIRE(a.omega > 0.0, SGW_E_ARG, std::string(who) + ": the cell volume must be positive");
}

// invfft6 of matrix m, r' side only: S[g][r'] for the num_g rows
void invfft6_rows(hipStream_t st, const Fft6Args &a, double2 *s, int m) {
  const int n = a.n, n_p = a.n_p, num_g = a.num_g, num_g_p = a.num_g_p;
  const int64_t nr = (int64_t)n * n * n, nr_p = (int64_t)n_p * n_p * n_p;
  double2 *f = a.f + (int64_t)m * nr * nr_p;
  k_fill0<<<blocks((int64_t)num_g * nr_p), kEB, 0, st>>>((int64_t)num_g * nr_p, s);
  k_rows_scatter<<<blocks((int64_t)num_g * num_g_p), kEB, 0, st>>>(
      num_g, num_g_p, nr, nr_p, a.nl_p, f, 1.0 / a.omega, s);
  box_fft(st, n_p, num_g, s, +1);
}

// fwfft6 of matrix m, r' side only, from S[g][r'] into the block of f
void fwfft6_rows(hipStream_t st, const Fft6Args &a, double2 *s, int m) {
  const int n = a.n, n_p = a.n_p, num_g = a.num_g, num_g_p = a.num_g_p;
  const int64_t nr = (int64_t)n * n * n, nr_p = (int64_t)n_p * n_p * n_p;
  double2 *f = a.f + (int64_t)m * nr * nr_p;
  box_fft(st, n_p, num_g, s, -1);
  k_rows_gather<<<blocks((int64_t)num_g * num_g_p), kEB, 0, st>>>(
      num_g, num_g_p, nr, nr_p, a.nl_p, s, a.omega / (double)nr_p, f);
}

// invfft6 of matrix m up to its last pass: f(r, r') in G along r, r' in r'
// space; the r transform is left to the caller
void invfft6_front(hipStream_t st, const Fft6Args &a, double2 *s, int m) {
  const int64_t nr = (int64_t)a.n * a.n * a.n, nr_p = (int64_t)a.n_p * a.n_p * a.n_p;
  double2 *f = a.f + (int64_t)m * nr * nr_p;
  invfft6_rows(st, a, s, m);
  k_fill0<<<blocks(nr * nr_p), kEB, 0, st>>>(nr * nr_p, f);
  k_cols_scatter<<<blocks((int64_t)a.num_g * nr_p), kEB, 0, st>>>(a.num_g, nr, nr_p, a.nl, s, f);
}

// fwfft6 of matrix m from after its first pass (f already transformed along r)
void fwfft6_back(hipStream_t st, const Fft6Args &a, double2 *s, int m) {
  const int64_t nr = (int64_t)a.n * a.n * a.n, nr_p = (int64_t)a.n_p * a.n_p * a.n_p;
  double2 *f = a.f + (int64_t)m * nr * nr_p;
  k_cols_gather<<<blocks((int64_t)a.num_g * nr_p), kEB, 0, st>>>(a.num_g, nr, nr_p, a.nl, f,
                                                                  1.0 / (double)nr, s);
  fwfft6_rows(st, a, s, m);
}

void invfft6_run(hipStream_t st, const Fft6Args &a, double2 *s) {
  const int64_t nr = (int64_t)a.n * a.n * a.n, nr_p = (int64_t)a.n_p * a.n_p * a.n_p;
  for (int m = 0; m < a.nmat; ++m) {
    invfft6_front(st, a, s, m);
    box_fft(st, a.n, nr_p, a.f + (int64_t)m * nr * nr_p, -1);
  }
}

void fwfft6_run(hipStream_t st, const Fft6Args &a, double2 *s) {
  const int64_t nr = (int64_t)a.n * a.n * a.n, nr_p = (int64_t)a.n_p * a.n_p * a.n_p;
  for (int m = 0; m < a.nmat; ++m) {
    box_fft(st, a.n, nr_p, a.f + (int64_t)m * nr * nr_p, +1);
    fwfft6_back(st, a, s, m);
  }
}

}  // namespace sgw

using namespace sgw;

extern "C" int sgw_invfft6(int32_t n, int32_t n_p, int32_t num_g, const int32_t *d_nl,
                           int32_t num_g_p, const int32_t *d_nl_p, double omega, int32_t nmat,
                           sgw_complex *d_f, void *stream) {
  return guarded([&] {
    Fft6Args a{n, n_p, num_g, num_g_p, d_nl, d_nl_p, omega, nmat,
               reinterpret_cast<double2 *>(d_f)};
    fft6_check(a, "invfft6");
    hipStream_t st = (hipStream_t)stream;
    DevBuf<double2> S((size_t)num_g * n_p * n_p * n_p);  // one matrix at a time
    KScope ks("invfft6", st);
    invfft6_run(st, a, S.p);
    SGW_HIP(hipGetLastError());
    SGW_HIP(hipStreamSynchronize(st));  // scratch freed on return
    return 0;
  });
}

extern "C" int sgw_fwfft6(int32_t n, int32_t n_p, int32_t num_g, const int32_t *d_nl,
                          int32_t num_g_p, const int32_t *d_nl_p, double omega, int32_t nmat,
                          sgw_complex *d_f, void *stream) {
  return guarded([&] {
    Fft6Args a{n, n_p, num_g, num_g_p, d_nl, d_nl_p, omega, nmat,
               reinterpret_cast<double2 *>(d_f)};
    fft6_check(a, "fwfft6");
    hipStream_t st = (hipStream_t)stream;
    DevBuf<double2> S((size_t)num_g * n_p * n_p * n_p);
    KScope ks("fwfft6", st);
    fwfft6_run(st, a, S.p);
    SGW_HIP(hipGetLastError());
    SGW_HIP(hipStreamSynchronize(st));
    return 0;
  });
}

extern "C" int sgw_green_assemble(int32_t nfreq, int32_t n, int32_t n_p, int32_t nrow,
                                  const int32_t *d_rowmap, int32_t ncol, const int32_t *d_colsys,
                                  const sgw_complex *d_x, int32_t nshift, int64_t ld,
                                  sgw_complex *d_green, void *stream) {
  return guarded([&] {
    SGW_REQUIRE(d_rowmap && d_colsys && d_x && d_green, SGW_E_ARG,
                "green_assemble: null argument");
    SGW_REQUIRE(nfreq >= 1 && nfreq <= nshift && n >= 1 && n_p >= 1, SGW_E_SHAPE,
                "green_assemble: need 1 <= nfreq <= nshift and positive boxes");
    const int64_t nr = (int64_t)n * n * n, nr_p = (int64_t)n_p * n_p * n_p;
    SGW_REQUIRE(nrow >= 0 && nrow <= nr && ncol >= 0 && ncol <= nr_p, SGW_E_SHAPE,
                "green_assemble: need nrow <= n^3 and ncol <= n_p^3");
    hipStream_t st = (hipStream_t)stream;
    KScope ks("green_assemble", st);
    k_green_assemble<<<blocks((int64_t)nfreq * nr * nr_p), kEB, 0, st>>>(
        nfreq, nr, nr_p, nrow, ncol, d_rowmap, d_colsys, reinterpret_cast<const double2 *>(d_x),
        nshift, ld, reinterpret_cast<double2 *>(d_green));
    SGW_HIP(hipGetLastError());
    return 0;
  });
}
