// The reference form of the correlation self-energy on the device:
//  * analytic continuation of W (algo/analytic/src/analytic.f90:50-348) with
//    the Godby-Needs plasmon-pole model (godby_needs.f90) and the N-point
//    Pade approximant of Vidberg and Serene (pade.f90:41-147), one thread per
//    (G, G') element;
//  * sigma_prod (phys/corr/src/sigma.f90:417-523): W(G, G') -> W(r, r') by
//    invfft6, the product G(r, r') W(r, r'), fwfft6 back, times alpha, and
//    optionally added into Sigma (sigma.f90:741).
#include <algorithm>
#include <vector>

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

extern "C" int sgw_analytic_coeff(int32_t model, int32_t num_g, int32_t nfreq,
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
    } else {
      throw Error{SGW_E_ARG, "analytic_coeff: model must be SGW_GODBY_NEEDS or SGW_PADE_APPROX"};
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
    SGW_REQUIRE(d_coeff && d_gmap && d_gmap_p && d_out && h_z, SGW_E_ARG,
                "analytic_eval: null argument");
    SGW_REQUIRE(model == SGW_GODBY_NEEDS || model == SGW_PADE_APPROX, SGW_E_ARG,
                "analytic_eval: model must be SGW_GODBY_NEEDS or SGW_PADE_APPROX");
    SGW_REQUIRE(model != SGW_GODBY_NEEDS || nfreq == 2, SGW_E_SHAPE,
                "analytic_eval: Godby-Needs uses 2 coefficients");
    const int64_t nr = (int64_t)n * n * n, nr_p = (int64_t)n_p * n_p * n_p;
    SGW_REQUIRE(num_g >= 1 && num_g <= nr && num_g_p >= 1 && num_g_p <= nr_p, SGW_E_SHAPE,
                "analytic_eval: need num_g <= n^3 and num_g_p <= n_p^3");
    const Mesh m = make_mesh(nfreq, h_z, nullptr);
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
