// The self-consistent response around the Sternheimer solves of solve_linter
// (phys/coul/src/solve_linter.f90:279-583), on the device:
//
//  * whole-box FFTs in QE's conventions (invfft no factor, fwfft 1/N^3);
//  * wave <-> box transfers through a sphere's box-index map (QE cft_wave,
//    dffts%nl(igk) scatter/gather);
//  * dV psi products gathered on the k+q sphere (dvqpsi_us.f90:99-132 for the
//    bare potential, solve_linter.f90:393-398 apply_dpot for dvscf);
//  * QE's orthogonalize for insulators (-P_c^+ with S = 1);
//  * the Hartree response potential of QE's dv_of_drho with lrpa
//    (solve_linter.f90:556; main/src/gwq_readin.f90:223 sets lrpa);
//  * the modified Broyden mixing of mix_potential_c (phys/coul/src/mix_pot_c.f90).
//
// Boxes are cubic n^3 (n in {16, 18, 24, 27, 32, 36, 48}), stored
// [vector][i3][i2][i1] with i1 fastest (QE's i1 + n1*(i2 + n2*i3)); the box
// FFTs are those of boxfft.hip.
#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "boxfft.h"
#include "ktimer.h"
#include "sgw_internal.h"

namespace sgw {
namespace {

constexpr int kEB = 256;       // elementwise kernels
constexpr int kRB = 256;       // reduction kernels
constexpr int kRedBlocks = 512;

// SIGN = +1: QE invfft (no factor); SIGN = -1: unscaled fwfft.
template <int SIGN> void box_fft(hipStream_t st, int n, int64_t nvec, double2 *d) {
  ::sgw::box_fft(st, n, nvec, d, SIGN);
}

// ---------------------------------------------------------------------------
// elementwise kernels
// ---------------------------------------------------------------------------
__global__ void k_fill0(int64_t count, double2 *__restrict__ d) {
  const int64_t e = (int64_t)blockIdx.x * kEB + threadIdx.x;
  if (e < count) d[e] = mk(0.0, 0.0);
}

__global__ void k_scale(int64_t count, double s, double2 *__restrict__ d) {
  const int64_t e = (int64_t)blockIdx.x * kEB + threadIdx.x;
  if (e < count) {
    const double2 x = d[e];
    d[e] = mk(s * x.x, s * x.y);
  }
}

// box[v][nl[i]] = psi[v][i], i < npw (box zeroed before)
__global__ void k_scatter(int64_t nvec, int npw, const int32_t *__restrict__ nl,
                          const double2 *__restrict__ psi, int64_t ld, int64_t nbox,
                          double2 *__restrict__ box) {
  const int64_t e = (int64_t)blockIdx.x * kEB + threadIdx.x;
  if (e >= nvec * npw) return;
  const int64_t v = e / npw;
  const int i = (int)(e % npw);
  box[v * nbox + nl[i]] = psi[v * ld + i];
}

// psi[v][i] (+)= scale * box[v][nl[i]] for i < npw; zero padding to ld when
// not accumulating
__global__ void k_gather(int64_t nvec, int npw, const int32_t *__restrict__ nl,
                         const double2 *__restrict__ box, int64_t nbox, double scale,
                         double2 *__restrict__ psi, int64_t ld, int accumulate) {
  const int64_t e = (int64_t)blockIdx.x * kEB + threadIdx.x;
  if (e >= nvec * ld) return;
  const int64_t v = e / ld;
  const int i = (int)(e % ld);
  if (i >= npw) {
    if (!accumulate) psi[e] = mk(0.0, 0.0);
    return;
  }
  const double2 b = box[v * nbox + nl[i]];
  const double2 val = mk(scale * b.x, scale * b.y);
  psi[e] = accumulate ? cadd(psi[e], val) : val;
}

// prod[f][v][r] = psi_r[v][r] * dv[f0 + f][r]
__global__ void k_mul_pot(int64_t nbox, int nbnd, int nf, const double2 *__restrict__ psi_r,
                          const double2 *__restrict__ dv, double2 *__restrict__ prod) {
  const int64_t e = (int64_t)blockIdx.x * kEB + threadIdx.x;
  if (e >= nbox * nbnd * nf) return;
  const int64_t r = e % nbox;
  const int64_t fv = e / nbox;
  const int v = (int)(fv % nbnd), f = (int)(fv / nbnd);
  prod[e] = cmul(psi_r[(int64_t)v * nbox + r], dv[(int64_t)f * nbox + r]);
}

// out[f][v][:] = base[v][:]  (sphere vectors, ld)
__global__ void k_copy_base(int64_t ld, int nbnd, int nf, const double2 *__restrict__ base,
                            double2 *__restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * kEB + threadIdx.x;
  if (e >= ld * nbnd * nf) return;
  out[e] = base[e % (ld * nbnd)];
}

// +/-omega average of solve_linter.f90:462-480 with the wg/wk rescale of
// :373: out[v][f] = scale * x(+omega_f) for f = 0 when zero_freq, else
// scale/2 (x(+omega_f) + x(-omega_f)).  layout 0: x[v][k][ld] over the
// shift list [freq, -freq(2:) or -freq]; layout 1: x[v][f][2][ld] with
// shift 0 = +omega_f and shift 1 = -omega_f.
__global__ void k_average_pm(int nbnd, int nf, int zero_freq, int layout, int64_t ld,
                             double scale, const double2 *__restrict__ x,
                             double2 *__restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * kEB + threadIdx.x;
  if (e >= (int64_t)nbnd * nf * ld) return;
  const int64_t i = e % ld;
  const int f = (int)((e / ld) % nf), v = (int)(e / (ld * nf));
  const int nshift = zero_freq ? 2 * nf - 1 : 2 * nf;
  int64_t ip, im;
  if (layout == 0) {
    ip = ((int64_t)v * nshift + f) * ld;
    im = ((int64_t)v * nshift + (zero_freq ? f + nf - 1 : f + nf)) * ld;
  } else {
    ip = (((int64_t)v * nf + f) * 2) * ld;
    im = ip + ld;
  }
  const double2 a = x[ip + i];
  if (zero_freq && f == 0) {
    out[e] = mk(scale * a.x, scale * a.y);
  } else {
    const double2 b = x[im + i];
    out[e] = mk(0.5 * scale * (a.x + b.x), 0.5 * scale * (a.y + b.y));
  }
}

struct Recip {
  double b[9];  // b1, b2, b3 (Cartesian, bohr^-1)
  double q[3];
};

// Hartree kernel of dv_of_drho in Ry (e2 = 2) on the fwfft'd boxes, with the
// 1/N^3 of fwfft folded in: coefficients with |G|^2 > gcut2, q + G = 0 or
// (zero_mean) G = 0 become zero.
__global__ void k_hartree(int n, int64_t nvec, Recip rc, double gcut2, int zero_mean,
                          double scale, double2 *__restrict__ d) {
  const int64_t nbox = (int64_t)n * n * n;
  const int64_t e = (int64_t)blockIdx.x * kEB + threadIdx.x;
  if (e >= nvec * nbox) return;
  const int64_t r = e % nbox;
  const int i1 = (int)(r % n), i2 = (int)((r / n) % n), i3 = (int)(r / ((int64_t)n * n));
  const int h = n / 2;
  const int m1 = i1 > h ? i1 - n : i1, m2 = i2 > h ? i2 - n : i2, m3 = i3 > h ? i3 - n : i3;
  const double gx = m1 * rc.b[0] + m2 * rc.b[3] + m3 * rc.b[6];
  const double gy = m1 * rc.b[1] + m2 * rc.b[4] + m3 * rc.b[7];
  const double gz = m1 * rc.b[2] + m2 * rc.b[5] + m3 * rc.b[8];
  const double g2 = gx * gx + gy * gy + gz * gz;
  const double qx = gx + rc.q[0], qy = gy + rc.q[1], qz = gz + rc.q[2];
  const double qg2 = qx * qx + qy * qy + qz * qz;
  const bool drop = g2 > gcut2 || qg2 <= 1e-10 || (zero_mean && m1 == 0 && m2 == 0 && m3 == 0);
  const double f = drop ? 0.0 : 2.0 * 4.0 * M_PI / qg2 * scale;
  const double2 x = d[e];
  d[e] = mk(f * x.x, f * x.y);
}

// QE orthogonalize (insulators, S = 1): dvpsi <- -(dvpsi - evq evq^H dvpsi)
// for every vector; one block per vector.
__global__ __launch_bounds__(kRB) void k_orthogonalize(int npw, int64_t ld, int nocc,
                                                       const double2 *__restrict__ evq,
                                                       double2 *__restrict__ dvpsi) {
  __shared__ double2 ps[64];
  __shared__ double2 red[kRB / 64];
  double2 *x = dvpsi + (int64_t)blockIdx.x * ld;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  for (int j = 0; j < nocc; ++j) {
    double2 acc = mk(0.0, 0.0);
    for (int i = t; i < npw; i += kRB) cmacc(acc, evq[(int64_t)j * ld + i], x[i]);
    for (int o = 32; o > 0; o >>= 1) {
      acc.x += __shfl_down(acc.x, o, 64);
      acc.y += __shfl_down(acc.y, o, 64);
    }
    if (lane == 0) red[wv] = acc;
    __syncthreads();
    if (t == 0) {
      double2 s = mk(0.0, 0.0);
      for (int w = 0; w < kRB / 64; ++w) s = cadd(s, red[w]);
      ps[j] = s;
    }
    __syncthreads();
  }
  for (int i = t; i < npw; i += kRB) {
    double2 y = x[i];
    for (int j = 0; j < nocc; ++j) {
      const double2 a = evq[(int64_t)j * ld + i], p = ps[j];
      y.x -= a.x * p.x - a.y * p.y;
      y.y -= a.x * p.y + a.y * p.x;
    }
    x[i] = mk(-y.x, -y.y);
  }
}

// ---------------------------------------------------------------------------
// modified Broyden mixing (mix_pot_c.f90)
// ---------------------------------------------------------------------------
template <int NV>
__device__ void block_reduce_store(double (&v)[NV], double *out) {
  __shared__ double red[kRB / 64][NV];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double a = v[k];
    for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
    if (lane == 0) red[wv][k] = a;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double s = 0.0;
    for (int w = 0; w < kRB / 64; ++w) s += red[w][threadIdx.x];
    out[threadIdx.x] = s;
  }
}

// vout -= vin;  part[b] = (|vout|^2, |df_p|^2) with, when prev != 0,
// df_p = vout - df_p and dv_p = vin - dv_p (mix_pot_c.f90:93-95, 128-131)
__global__ __launch_bounds__(kRB) void k_mix_diff(int64_t ndim, double2 *__restrict__ vout,
                                                  const double2 *__restrict__ vin,
                                                  double2 *__restrict__ dfp,
                                                  double2 *__restrict__ dvp, int prev,
                                                  double *__restrict__ part) {
  double acc[2] = {0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * kRB + threadIdx.x; i < ndim;
       i += (int64_t)gridDim.x * kRB) {
    const double2 a = vin[i];
    const double2 o = csub(vout[i], a);
    vout[i] = o;
    acc[0] += o.x * o.x + o.y * o.y;
    if (prev) {
      const double2 f = csub(o, dfp[i]);
      dfp[i] = f;
      dvp[i] = csub(a, dvp[i]);
      acc[1] += f.x * f.x + f.y * f.y;
    }
  }
  block_reduce_store<2>(acc, part + (int64_t)blockIdx.x * 2);
}

__global__ void k_mix_scale(int64_t ndim, double s, double2 *__restrict__ a,
                            double2 *__restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * kEB + threadIdx.x;
  if (i >= ndim) return;
  const double2 x = a[i], y = b[i];
  a[i] = mk(s * x.x, s * x.y);
  b[i] = mk(s * y.x, s * y.y);
}

// Gram entries <df_j | df_i> for i < j < K and <df_i | vout> for i < K:
// K(K-1)/2 + K complex sums per block (ZDOTC, conjugating the first operand)
template <int K>
__global__ __launch_bounds__(kRB) void k_mix_gram(int64_t ndim, const double2 *__restrict__ df,
                                                  const double2 *__restrict__ vout,
                                                  double *__restrict__ part) {
  constexpr int NP = K * (K - 1) / 2 + K;
  double acc[2 * NP];
#pragma unroll
  for (int k = 0; k < 2 * NP; ++k) acc[k] = 0.0;
  for (int64_t n = (int64_t)blockIdx.x * kRB + threadIdx.x; n < ndim;
       n += (int64_t)gridDim.x * kRB) {
    double2 c[K];
#pragma unroll
    for (int i = 0; i < K; ++i) c[i] = df[(int64_t)i * ndim + n];
    const double2 o = vout[n];
    int p = 0;
#pragma unroll
    for (int i = 0; i < K; ++i) {
#pragma unroll
      for (int j = i + 1; j < K; ++j) {
        // conj(df_j) df_i
        acc[2 * p] += c[j].x * c[i].x + c[j].y * c[i].y;
        acc[2 * p + 1] += c[j].x * c[i].y - c[j].y * c[i].x;
        ++p;
      }
    }
#pragma unroll
    for (int i = 0; i < K; ++i) {
      acc[2 * p] += c[i].x * o.x + c[i].y * o.y;
      acc[2 * p + 1] += c[i].x * o.y - c[i].y * o.x;
      ++p;
    }
  }
  block_reduce_store<2 * NP>(acc, part + (int64_t)blockIdx.x * 2 * NP);
}

// sum the per-block partials: out[k] = sum_b part[b][k]
__global__ __launch_bounds__(kRB) void k_sum_parts(int nblk, int nv, const double *__restrict__ part,
                                                   double *__restrict__ out) {
  for (int k = threadIdx.x; k < nv; k += kRB) {
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += part[(int64_t)b * nv + k];
    out[k] = s;
  }
}

struct MixCoef {
  double2 c[8];  // w_i gamma_i
};

// vin <- vin + alpha vout - sum_i c_i (alpha df_i + dv_i);
// then df_next = vout, dv_next = old vin (mix_pot_c.f90:166-186)
__global__ void k_mix_update(int64_t ndim, int K, MixCoef mc, double alpha,
                             double2 *__restrict__ vin, const double2 *__restrict__ vout,
                             double2 *__restrict__ df, double2 *__restrict__ dv, int inext) {
  const int64_t n = (int64_t)blockIdx.x * kEB + threadIdx.x;
  if (n >= ndim) return;
  const double2 vi = vin[n], vo = vout[n];
  double2 acc = mk(vi.x + alpha * vo.x, vi.y + alpha * vo.y);
  for (int i = 0; i < K; ++i) {
    const double2 f = df[(int64_t)i * ndim + n], g = dv[(int64_t)i * ndim + n];
    const double2 t = mk(alpha * f.x + g.x, alpha * f.y + g.y);
    const double2 c = mc.c[i];
    acc.x -= c.x * t.x - c.y * t.y;
    acc.y -= c.x * t.y + c.y * t.x;
  }
  vin[n] = acc;
  df[(int64_t)inext * ndim + n] = vo;
  dv[(int64_t)inext * ndim + n] = vi;
}

template <int K>
void launch_gram(hipStream_t st, int nblk, int64_t ndim, const double2 *df, const double2 *vout,
                 double *part) {
  k_mix_gram<K><<<nblk, kRB, 0, st>>>(ndim, df, vout, part);
}

void gram(hipStream_t st, int K, int nblk, int64_t ndim, const double2 *df, const double2 *vout,
          double *part) {
  switch (K) {
    case 1: launch_gram<1>(st, nblk, ndim, df, vout, part); break;
    case 2: launch_gram<2>(st, nblk, ndim, df, vout, part); break;
    case 3: launch_gram<3>(st, nblk, ndim, df, vout, part); break;
    case 4: launch_gram<4>(st, nblk, ndim, df, vout, part); break;
    case 5: launch_gram<5>(st, nblk, ndim, df, vout, part); break;
    case 6: launch_gram<6>(st, nblk, ndim, df, vout, part); break;
    case 7: launch_gram<7>(st, nblk, ndim, df, vout, part); break;
    case 8: launch_gram<8>(st, nblk, ndim, df, vout, part); break;
    default: throw Error{SGW_E_ARG, "mix_potential_c: at most 8 stored iterations"};
  }
}

// inverse of a small Hermitian positive matrix (Gauss-Jordan, partial pivoting)
void invert_small(int k, std::vector<std::complex<double>> &a) {
  std::vector<std::complex<double>> inv((size_t)k * k, 0.0);
  for (int i = 0; i < k; ++i) inv[(size_t)i * k + i] = 1.0;
  for (int c = 0; c < k; ++c) {
    int piv = c;
    for (int r = c + 1; r < k; ++r)
      if (std::abs(a[(size_t)r * k + c]) > std::abs(a[(size_t)piv * k + c])) piv = r;
    SGW_REQUIRE(std::abs(a[(size_t)piv * k + c]) > 0.0, SGW_E_STATE,
                "mix_potential_c: singular Broyden matrix (factorization)");
    if (piv != c)
      for (int j = 0; j < k; ++j) {
        std::swap(a[(size_t)c * k + j], a[(size_t)piv * k + j]);
        std::swap(inv[(size_t)c * k + j], inv[(size_t)piv * k + j]);
      }
    const std::complex<double> d = a[(size_t)c * k + c];
    for (int j = 0; j < k; ++j) {
      a[(size_t)c * k + j] /= d;
      inv[(size_t)c * k + j] /= d;
    }
    for (int r = 0; r < k; ++r) {
      if (r == c) continue;
      const std::complex<double> f = a[(size_t)r * k + c];
      if (f == 0.0) continue;
      for (int j = 0; j < k; ++j) {
        a[(size_t)r * k + j] -= f * a[(size_t)c * k + j];
        inv[(size_t)r * k + j] -= f * inv[(size_t)c * k + j];
      }
    }
  }
  a.swap(inv);
}

template <typename F> int guarded(F &&f) {
  try {
    return f();
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::exception &e) {
    set_error(e.what());
    return SGW_E_STATE;
  }
}

unsigned blocks(int64_t count) { return (unsigned)((count + kEB - 1) / kEB); }

}  // namespace
}  // namespace sgw

struct sgw_broyden {
  int64_t ndim = 0;
  int n_iter = 0;
  bool live = false;  // history valid (allocated in the reference's terms)
  sgw::DevBuf<double2> df, dv;
  sgw::DevBuf<double> part, sums;
};

using namespace sgw;

extern "C" int sgw_box_fft(int32_t n, int32_t nvec, int32_t sign, sgw_complex *d_box,
                           void *stream) {
  return guarded([&] {
    SGW_REQUIRE(d_box && nvec >= 0, SGW_E_ARG, "box_fft: null box or negative count");
    SGW_REQUIRE(box_supported(n), SGW_E_SHAPE, "box_fft: n must be 16, 18, 24, 27, 32, 36 or 48");
    SGW_REQUIRE(sign == 1 || sign == -1, SGW_E_ARG, "box_fft: sign must be +1 or -1");
    hipStream_t st = (hipStream_t)stream;
    double2 *d = reinterpret_cast<double2 *>(d_box);
    const int64_t nbox = (int64_t)n * n * n;
    KScope ks("box_fft", st);
    if (sign > 0) {
      box_fft<1>(st, n, nvec, d);
    } else {
      box_fft<-1>(st, n, nvec, d);
      if (nvec > 0)
        k_scale<<<blocks(nbox * nvec), kEB, 0, st>>>(nbox * nvec, 1.0 / (double)nbox, d);
    }
    SGW_HIP(hipGetLastError());
    return 0;
  });
}

extern "C" int sgw_wave_to_box(int32_t n, int32_t nvec, int32_t npw, const int32_t *d_nl,
                               const sgw_complex *d_psi, int64_t ld, sgw_complex *d_box,
                               void *stream) {
  return guarded([&] {
    SGW_REQUIRE(d_nl && d_psi && d_box && nvec >= 0, SGW_E_ARG, "wave_to_box: null argument");
    SGW_REQUIRE(box_supported(n), SGW_E_SHAPE, "wave_to_box: n must be 16, 18, 24, 27, 32, 36 or 48");
    SGW_REQUIRE(npw >= 0 && npw <= ld, SGW_E_SHAPE, "wave_to_box: need 0 <= npw <= ld");
    hipStream_t st = (hipStream_t)stream;
    const int64_t nbox = (int64_t)n * n * n;
    double2 *box = reinterpret_cast<double2 *>(d_box);
    KScope ks("wave_to_box", st);
    if (nvec > 0) {
      k_fill0<<<blocks(nbox * nvec), kEB, 0, st>>>(nbox * nvec, box);
      if (npw > 0)
        k_scatter<<<blocks((int64_t)nvec * npw), kEB, 0, st>>>(
            nvec, npw, d_nl, reinterpret_cast<const double2 *>(d_psi), ld, nbox, box);
      box_fft<1>(st, n, nvec, box);
    }
    SGW_HIP(hipGetLastError());
    return 0;
  });
}

extern "C" int sgw_box_to_wave(int32_t n, int32_t nvec, int32_t npw, const int32_t *d_nl,
                               sgw_complex *d_box, sgw_complex *d_psi, int64_t ld,
                               int32_t accumulate, void *stream) {
  return guarded([&] {
    SGW_REQUIRE(d_nl && d_psi && d_box && nvec >= 0, SGW_E_ARG, "box_to_wave: null argument");
    SGW_REQUIRE(box_supported(n), SGW_E_SHAPE, "box_to_wave: n must be 16, 18, 24, 27, 32, 36 or 48");
    SGW_REQUIRE(npw >= 0 && npw <= ld, SGW_E_SHAPE, "box_to_wave: need 0 <= npw <= ld");
    hipStream_t st = (hipStream_t)stream;
    const int64_t nbox = (int64_t)n * n * n;
    double2 *box = reinterpret_cast<double2 *>(d_box);
    KScope ks("box_to_wave", st);
    if (nvec > 0) {
      box_fft<-1>(st, n, nvec, box);
      k_gather<<<blocks((int64_t)nvec * ld), kEB, 0, st>>>(
          nvec, npw, d_nl, box, nbox, 1.0 / (double)nbox, reinterpret_cast<double2 *>(d_psi),
          ld, accumulate ? 1 : 0);
    }
    SGW_HIP(hipGetLastError());
    return 0;
  });
}

extern "C" int sgw_dvscf_psi(int32_t n, int32_t nbnd, int32_t nfreq, const sgw_complex *d_psi_r,
                             const sgw_complex *d_dvscf, int32_t npwq, const int32_t *d_nlq,
                             const sgw_complex *d_base, sgw_complex *d_dvpsi, int64_t ld,
                             void *stream) {
  return guarded([&] {
    SGW_REQUIRE(d_psi_r && d_dvscf && d_nlq && d_dvpsi, SGW_E_ARG, "dvscf_psi: null argument");
    SGW_REQUIRE(box_supported(n), SGW_E_SHAPE, "dvscf_psi: n must be 16, 18, 24, 27, 32, 36 or 48");
    SGW_REQUIRE(nbnd >= 1 && nfreq >= 1 && npwq >= 1 && npwq <= ld, SGW_E_SHAPE,
                "dvscf_psi: need nbnd, nfreq >= 1 and 1 <= npwq <= ld");
    hipStream_t st = (hipStream_t)stream;
    const int64_t nbox = (int64_t)n * n * n;
    // frequencies per pass: at most 512 boxes of scratch
    const int fchunk = std::max(1, std::min<int>(nfreq, 512 / nbnd));
    DevBuf<double2> prod((size_t)nbox * nbnd * fchunk);
    const double2 *psi_r = reinterpret_cast<const double2 *>(d_psi_r);
    const double2 *dv = reinterpret_cast<const double2 *>(d_dvscf);
    double2 *out = reinterpret_cast<double2 *>(d_dvpsi);
    KScope ks("dvscf_psi", st);
    if (d_base)
      k_copy_base<<<blocks(ld * nbnd * nfreq), kEB, 0, st>>>(
          ld, nbnd, nfreq, reinterpret_cast<const double2 *>(d_base), out);
    for (int f0 = 0; f0 < nfreq; f0 += fchunk) {
      const int nf = std::min(fchunk, nfreq - f0);
      const int64_t nvec = (int64_t)nf * nbnd;
      k_mul_pot<<<blocks(nbox * nvec), kEB, 0, st>>>(nbox, nbnd, nf, psi_r, dv + (int64_t)f0 * nbox,
                                                      prod.p);
      box_fft<-1>(st, n, nvec, prod.p);
      k_gather<<<blocks(nvec * ld), kEB, 0, st>>>(nvec, npwq, d_nlq, prod.p, nbox,
                                                   1.0 / (double)nbox,
                                                   out + (int64_t)f0 * nbnd * ld, ld,
                                                   d_base ? 1 : 0);
    }
    SGW_HIP(hipGetLastError());
    SGW_HIP(hipStreamSynchronize(st));  // scratch freed on return
    return 0;
  });
}

extern "C" int sgw_orthogonalize(int32_t nvec, int32_t npw, int32_t nocc, const sgw_complex *d_evq,
                                 sgw_complex *d_dvpsi, int64_t ld, void *stream) {
  return guarded([&] {
    SGW_REQUIRE(d_evq && d_dvpsi && nvec >= 0, SGW_E_ARG, "orthogonalize: null argument");
    SGW_REQUIRE(nocc >= 0 && nocc <= 64 && npw >= 0 && npw <= ld, SGW_E_SHAPE,
                "orthogonalize: need 0 <= nocc <= 64 and 0 <= npw <= ld");
    hipStream_t st = (hipStream_t)stream;
    KScope ks("orthogonalize", st);
    if (nvec > 0)
      k_orthogonalize<<<nvec, kRB, 0, st>>>(npw, ld, nocc, reinterpret_cast<const double2 *>(d_evq),
                                            reinterpret_cast<double2 *>(d_dvpsi));
    SGW_HIP(hipGetLastError());
    return 0;
  });
}

extern "C" int sgw_average_pm(int32_t nbnd, int32_t nf, int32_t zero_freq, int32_t layout,
                              const sgw_complex *d_x, int64_t ld, double scale,
                              sgw_complex *d_out, void *stream) {
  return guarded([&] {
    SGW_REQUIRE(d_x && d_out, SGW_E_ARG, "average_pm: null argument");
    SGW_REQUIRE(nbnd >= 0 && nf >= 1 && ld >= 1 && (layout == 0 || layout == 1), SGW_E_ARG,
                "average_pm: need nbnd >= 0, nf >= 1, ld >= 1 and layout 0 or 1");
    hipStream_t st = (hipStream_t)stream;
    const int64_t cnt = (int64_t)nbnd * nf * ld;
    if (cnt > 0)
      k_average_pm<<<blocks(cnt), kEB, 0, st>>>(nbnd, nf, zero_freq ? 1 : 0, layout, ld, scale,
                                                 reinterpret_cast<const double2 *>(d_x),
                                                 reinterpret_cast<double2 *>(d_out));
    SGW_HIP(hipGetLastError());
    return 0;
  });
}

extern "C" int sgw_dv_of_drho(int32_t n, int32_t nvec, const double *bg, const double *xq,
                              double gcut2, int32_t zero_mean, sgw_complex *d_box, void *stream) {
  return guarded([&] {
    SGW_REQUIRE(bg && xq && d_box && nvec >= 0, SGW_E_ARG, "dv_of_drho: null argument");
    SGW_REQUIRE(box_supported(n), SGW_E_SHAPE, "dv_of_drho: n must be 16, 18, 24, 27, 32, 36 or 48");
    hipStream_t st = (hipStream_t)stream;
    const int64_t nbox = (int64_t)n * n * n;
    double2 *d = reinterpret_cast<double2 *>(d_box);
    Recip rc;
    for (int i = 0; i < 9; ++i) rc.b[i] = bg[i];
    for (int i = 0; i < 3; ++i) rc.q[i] = xq[i];
    KScope ks("dv_of_drho", st);
    if (nvec > 0) {
      box_fft<-1>(st, n, nvec, d);
      k_hartree<<<blocks(nbox * nvec), kEB, 0, st>>>(n, nvec, rc, gcut2, zero_mean ? 1 : 0,
                                                      1.0 / (double)nbox, d);
      box_fft<1>(st, n, nvec, d);
    }
    SGW_HIP(hipGetLastError());
    return 0;
  });
}

extern "C" sgw_broyden *sgw_broyden_create(int64_t ndim, int32_t n_iter) {
  try {
    SGW_REQUIRE(ndim >= 1, SGW_E_ARG, "broyden_create: ndim must be positive");
    SGW_REQUIRE(n_iter >= 1 && n_iter <= 8, SGW_E_ARG,
                "broyden_create: n_iter must be in [1, 8] (maxter = 8)");
    auto *b = new sgw_broyden;
    b->ndim = ndim;
    b->n_iter = n_iter;
    b->df.alloc((size_t)ndim * n_iter);
    b->dv.alloc((size_t)ndim * n_iter);
    const int nv = 2 * (8 * 7 / 2 + 8);
    b->part.alloc((size_t)kRedBlocks * nv);
    b->sums.alloc(nv);
    return b;
  } catch (const Error &e) {
    set_error(e.msg);
    return nullptr;
  } catch (const std::exception &e) {
    set_error(e.what());
    return nullptr;
  }
}

extern "C" void sgw_broyden_destroy(sgw_broyden *b) { delete b; }

extern "C" int sgw_mix_potential_c(sgw_broyden *b, int64_t ndim, sgw_complex *d_vout,
                                   sgw_complex *d_vin, double alphamix, double *dr2, double tr2,
                                   int32_t iter, int32_t n_iter, int32_t *conv, void *stream) {
  return guarded([&] {
    SGW_REQUIRE(b && d_vout && d_vin && dr2 && conv, SGW_E_ARG, "mix_potential_c: null argument");
    SGW_REQUIRE(ndim == b->ndim && n_iter == b->n_iter, SGW_E_SHAPE,
                "mix_potential_c: ndim and n_iter must match the Broyden state");
    SGW_REQUIRE(iter >= 1, SGW_E_ARG, "mix_potential_c: iter must be >= 1");
    SGW_REQUIRE(iter == 1 || b->live, SGW_E_STATE,
                "mix_potential_c: iter > 1 without the history of earlier iterations");
    hipStream_t st = (hipStream_t)stream;
    double2 *vout = reinterpret_cast<double2 *>(d_vout);
    double2 *vin = reinterpret_cast<double2 *>(d_vin);
    const int64_t N = ndim;
    const int nblk = (int)std::min<int64_t>(kRedBlocks, (N + kRB - 1) / kRB);
    KScope ks("mix_potential_c", st);
    // slots are 1-based in the reference; stored 0-based here
    const int iter_used = std::min(iter - 1, n_iter);
    const int ipos = iter - 1 - ((iter - 2) / n_iter) * n_iter;     // valid for iter > 1
    const int inext = iter - ((iter - 1) / n_iter) * n_iter;
    const bool prev = iter > 1;
    double2 *dfp = prev ? b->df.p + (int64_t)(ipos - 1) * N : nullptr;
    double2 *dvp = prev ? b->dv.p + (int64_t)(ipos - 1) * N : nullptr;
    k_mix_diff<<<nblk, kRB, 0, st>>>(N, vout, vin, dfp, dvp, prev ? 1 : 0, b->part.p);
    k_sum_parts<<<1, kRB, 0, st>>>(nblk, 2, b->part.p, b->sums.p);
    double h2[2];
    SGW_HIP(hipMemcpyAsync(h2, b->sums.p, sizeof(h2), hipMemcpyDeviceToHost, st));
    SGW_HIP(hipStreamSynchronize(st));
    const double r = std::sqrt(h2[0]) / (double)N;
    *dr2 = r * r;
    *conv = *dr2 < tr2 ? 1 : 0;
    if (*conv) {  // the reference deallocates its history and returns
      b->live = false;
      return 0;
    }
    if (prev) {
      const double nrm = std::sqrt(h2[1]);
      SGW_REQUIRE(nrm > 0.0, SGW_E_STATE, "mix_potential_c: zero change of the residual");
      k_mix_scale<<<blocks(N), kEB, 0, st>>>(N, 1.0 / nrm, dfp, dvp);
    }
    MixCoef mc{};
    const double w0 = 0.01, wi = 1.0;  // mix_pot_c.f90:90 (w0 = 0.01, w = 1)
    if (iter_used > 0) {
      const int K = iter_used;
      const int np = K * (K - 1) / 2 + K;
      gram(st, K, nblk, N, b->df.p, vout, b->part.p);
      k_sum_parts<<<1, kRB, 0, st>>>(nblk, 2 * np, b->part.p, b->sums.p);
      std::vector<double> hs((size_t)2 * np);
      SGW_HIP(hipMemcpyAsync(hs.data(), b->sums.p, hs.size() * sizeof(double),
                             hipMemcpyDeviceToHost, st));
      SGW_HIP(hipStreamSynchronize(st));
      // a(i, j) = w_i w_j <df_j|df_i> (i < j), a(i, i) = w0^2 + w_i^2, Hermitian
      std::vector<std::complex<double>> a((size_t)K * K);
      int p = 0;
      for (int i = 0; i < K; ++i) {
        a[(size_t)i * K + i] = w0 * w0 + wi * wi;
        for (int j = i + 1; j < K; ++j, ++p) {
          const std::complex<double> v(hs[2 * p], hs[2 * p + 1]);
          a[(size_t)i * K + j] = wi * wi * v;
          a[(size_t)j * K + i] = std::conj(wi * wi * v);
        }
      }
      std::vector<std::complex<double>> work(K);
      for (int i = 0; i < K; ++i, ++p) work[i] = {hs[2 * p], hs[2 * p + 1]};
      invert_small(K, a);  // beta
      for (int i = 0; i < K; ++i) {
        std::complex<double> gamma = 0.0;
        for (int j = 0; j < K; ++j) gamma += a[(size_t)j * K + i] * wi * work[j];
        const std::complex<double> c = wi * gamma;
        mc.c[i] = make_double2(c.real(), c.imag());
      }
    }
    k_mix_update<<<blocks(N), kEB, 0, st>>>(N, iter_used, mc, alphamix, vin, vout, b->df.p,
                                            b->dv.p, inext - 1);
    SGW_HIP(hipGetLastError());
    b->live = true;
    return 0;
  });
}
