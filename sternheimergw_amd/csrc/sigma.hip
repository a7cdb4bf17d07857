// Sigma_c frequency convolution as a batched complex GEMM on FP64 MFMA:
//   Sigma[i] (=|+=) alpha * sum_j w_j G[i][j] W[j],   i < nout, j < nin,
// all matrices n x n complex, column-major with leading dimension ld.
//
// The reference accumulates Sigma(G, G'; w_sigma) over the Green's-function
// frequencies with weights alpha * w (phys/corr/src/sigma.f90:705-745; the
// product itself is formed in real space by sigma_prod, :417-523).  Here the
// same frequency sum is one long-K GEMM per output frequency:
//   Sigma[i] = alpha [w_0 G[i][0] | ... | w_{nin-1} G[i][nin-1]] x [W[0]; ...; W[nin-1]].
//
// Kernel: one 128 x 64 complex tile of Sigma[i] per block, four waves of
// 64 x 32; K staged through LDS 16 at a time; each complex 16 x 16 x 4 step
// is four real v_mfma_f64_16x16x4f64 (re*re - im*im, re*im + im*re) into
// FP64 accumulators.
#include "sgw.h"

#include <cmath>
#include <string>
#include <vector>

#include "ktimer.h"
#include "sgw_internal.h"

namespace sgw {
namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kSBM = 128, kSBN = 64, kSBK = 16, kSPad = 1;

struct SigmaArgs {
  int n;
  int64_t ld;
  int nout, nin;
  int64_t gsi, gsj, wsj, ssi;
  const double2 *G, *W;
  const double *wt;
  double2 alpha;
  double2 *S;
  int accumulate;
  int tiles_m, tiles_n, nblocks;
};

// Block b runs on XCD b % 8; blocks of one XCD take consecutive tiles so the
// tiles sharing a row panel of G[i][j] or a column panel of W[j] meet in the
// same L2.
__device__ inline int xcd_tile(int b, int nblocks) {
  const int per = (nblocks + 7) / 8;
  return (b % 8) * per + b / 8;
}

__global__ __launch_bounds__(256) void k_sigma_f64(SigmaArgs a) {
  __shared__ double2 As[kSBK][kSBM + kSPad];
  __shared__ double2 Bs[kSBK][kSBN + kSPad];
  const int t = xcd_tile(blockIdx.x, a.nblocks);
  const int tiles = a.tiles_m * a.tiles_n;
  if (t >= tiles * a.nout) return;
  const int i = t / tiles, tt = t % tiles;
  const int m0 = (tt % a.tiles_m) * kSBM, n0 = (tt / a.tiles_m) * kSBN;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wm = (wv & 1) * 64, wn = (wv >> 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
  const int n = a.n;
  const int64_t ld = a.ld;
  v4d cre[4][2], cim[4][2];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      cre[r][c] = v4d{0, 0, 0, 0};
      cim[r][c] = v4d{0, 0, 0, 0};
    }
  const int arow = tid & 127, ak = tid >> 7;  // A staging: 8 (row, k) per thread
  const int bk = tid & 15, bc = tid >> 4;     // B staging: 4 (k, col) per thread
  // stages t = (j, kk): K runs over the G'' index of every convolution
  // frequency in turn; the next stage's tiles are fetched into registers
  // while the MFMAs of the current one run
  const int nk = (n + kSBK - 1) / kSBK, nstage = a.nin * nk;
  double2 ra[8], rb[4];
  double rw = 0;
  auto fetch = [&](int t) {
    const int j = t / nk, kk = (t % nk) * kSBK;
    const double2 *Gij = a.G + (int64_t)i * a.gsi + (int64_t)j * a.gsj;
    const double2 *Wj = a.W + (int64_t)j * a.wsj;
    rw = a.wt[j];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int gm = m0 + arow, gk = kk + ak + 2 * r;
      ra[r] = (gm < n && gk < n) ? Gij[gm + (int64_t)gk * ld] : mk(0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int gk = kk + bk, gn = n0 + bc + 16 * r;
      rb[r] = (gk < n && gn < n) ? Wj[gk + (int64_t)gn * ld] : mk(0, 0);
    }
  };
  fetch(0);
  for (int t = 0; t < nstage; ++t) {
    {
      __syncthreads();
#pragma unroll
      for (int r = 0; r < 8; ++r) As[ak + 2 * r][arow] = mk(ra[r].x * rw, ra[r].y * rw);
#pragma unroll
      for (int r = 0; r < 4; ++r) Bs[bk][bc + 16 * r] = rb[r];
      __syncthreads();
      if (t + 1 < nstage) fetch(t + 1);
#pragma unroll
      for (int k4 = 0; k4 < kSBK / 4; ++k4) {
        const int kr = 4 * k4 + lk;
        double2 av[4], bv[2];
#pragma unroll
        for (int r = 0; r < 4; ++r) av[r] = As[kr][wm + 16 * r + lr];
#pragma unroll
        for (int c = 0; c < 2; ++c) bv[c] = Bs[kr][wn + 16 * c + lr];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int c = 0; c < 2; ++c) {
            cre[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[r].x, bv[c].x, cre[r][c], 0, 0, 0);
            cre[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(-av[r].y, bv[c].y, cre[r][c], 0, 0, 0);
            cim[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[r].x, bv[c].y, cim[r][c], 0, 0, 0);
            cim[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[r].y, bv[c].x, cim[r][c], 0, 0, 0);
          }
      }
    }
  }
  // v_mfma_f64_16x16x4f64 leaves D(lane / 16 + 4 q, lane % 16) in element q
  // of a lane's accumulator (A(lane % 16, lane / 16), B(lane / 16, lane % 16)
  // in its operands); checked on the device
  double2 *Si = a.S + (int64_t)i * a.ssi;
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = m0 + wm + 16 * r + lk + 4 * q, col = n0 + wn + 16 * c + lr;
        if (row < n && col < n) {
          const double2 acc = mk(cre[r][c][q], cim[r][c][q]);
          double2 v = cmul(a.alpha, acc);
          double2 *p = Si + row + (int64_t)col * ld;
          if (a.accumulate) v = cadd(*p, v);
          *p = v;
        }
      }
}

// ---------------------------------------------------------------------------
// FP32 split precision: the same tiling with operands rounded to FP32 when
// staged and v_mfma_f32_16x16x4f32 products (twice the FP64 MFMA rate); the
// FP32 accumulators hold the sum over one frequency j only (K = n) and are
// flushed into FP64 accumulators before the next, so the rounding of the
// 64-frequency sum stays FP64 and the error is that of an n-term FP32 dot.
// ---------------------------------------------------------------------------
typedef float v4f __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_sigma_f32s(SigmaArgs a) {
  __shared__ float2 As[kSBK][kSBM + kSPad];
  __shared__ float2 Bs[kSBK][kSBN + kSPad];
  const int t = xcd_tile(blockIdx.x, a.nblocks);
  const int tiles = a.tiles_m * a.tiles_n;
  if (t >= tiles * a.nout) return;
  const int i = t / tiles, tt = t % tiles;
  const int m0 = (tt % a.tiles_m) * kSBM, n0 = (tt / a.tiles_m) * kSBN;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wm = (wv & 1) * 64, wn = (wv >> 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
  const int n = a.n;
  const int64_t ld = a.ld;
  v4f fre[4][2], fim[4][2];   // FP32 partial sums of the current frequency
  double dre[4][2][4], dim_[4][2][4];  // FP64 running sums
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      fre[r][c] = v4f{0, 0, 0, 0};
      fim[r][c] = v4f{0, 0, 0, 0};
#pragma unroll
      for (int q = 0; q < 4; ++q) dre[r][c][q] = dim_[r][c][q] = 0.0;
    }
  const int arow = tid & 127, ak = tid >> 7;
  const int bk = tid & 15, bc = tid >> 4;
  const int nk = (n + kSBK - 1) / kSBK, nstage = a.nin * nk;
  double2 ra[8], rb[4];
  double rw = 0;
  auto fetch = [&](int s) {
    const int j = s / nk, kk = (s % nk) * kSBK;
    const double2 *Gij = a.G + (int64_t)i * a.gsi + (int64_t)j * a.gsj;
    const double2 *Wj = a.W + (int64_t)j * a.wsj;
    rw = a.wt[j];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int gm = m0 + arow, gk = kk + ak + 2 * r;
      ra[r] = (gm < n && gk < n) ? Gij[gm + (int64_t)gk * ld] : mk(0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int gk = kk + bk, gn = n0 + bc + 16 * r;
      rb[r] = (gk < n && gn < n) ? Wj[gk + (int64_t)gn * ld] : mk(0, 0);
    }
  };
  fetch(0);
  for (int s = 0; s < nstage; ++s) {
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 8; ++r)
      As[ak + 2 * r][arow] = make_float2((float)(ra[r].x * rw), (float)(ra[r].y * rw));
#pragma unroll
    for (int r = 0; r < 4; ++r) Bs[bk][bc + 16 * r] = make_float2((float)rb[r].x, (float)rb[r].y);
    __syncthreads();
    if (s + 1 < nstage) fetch(s + 1);
#pragma unroll
    for (int k4 = 0; k4 < kSBK / 4; ++k4) {
      const int kr = 4 * k4 + lk;
      float2 av[4], bv[2];
#pragma unroll
      for (int r = 0; r < 4; ++r) av[r] = As[kr][wm + 16 * r + lr];
#pragma unroll
      for (int c = 0; c < 2; ++c) bv[c] = Bs[kr][wn + 16 * c + lr];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          fre[r][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[r].x, bv[c].x, fre[r][c], 0, 0, 0);
          fre[r][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(-av[r].y, bv[c].y, fre[r][c], 0, 0, 0);
          fim[r][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[r].x, bv[c].y, fim[r][c], 0, 0, 0);
          fim[r][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[r].y, bv[c].x, fim[r][c], 0, 0, 0);
        }
    }
    if ((s + 1) % nk == 0) {  // end of frequency j: flush the FP32 partials
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            dre[r][c][q] += (double)fre[r][c][q];
            dim_[r][c][q] += (double)fim[r][c][q];
          }
          fre[r][c] = v4f{0, 0, 0, 0};
          fim[r][c] = v4f{0, 0, 0, 0};
        }
    }
  }
  // v_mfma_f32_16x16x4f32 leaves D(4 (lane / 16) + q, lane % 16) in element q
  double2 *Si = a.S + (int64_t)i * a.ssi;
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = m0 + wm + 16 * r + 4 * lk + q, col = n0 + wn + 16 * c + lr;
        if (row < n && col < n) {
          double2 v = cmul(a.alpha, mk(dre[r][c][q], dim_[r][c][q]));
          double2 *p = Si + row + (int64_t)col * ld;
          if (a.accumulate) v = cadd(*p, v);
          *p = v;
        }
      }
}

// ---------------------------------------------------------------------------
// Synthetic seeded inputs of the convolution benchmark (generated on the
// device because the full Green's-function stack is 147 GB at n = 1500).
// ---------------------------------------------------------------------------
__device__ inline uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ inline double unit_sym(uint64_t z) {  // uniform in [-1, 1)
  return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

// G[i][j](a, b) = (u1 + i u2) / ((1 + nu_ij) sqrt(n)), nu_ij = omega_i + omega'_j (Ha),
// u1, u2 = splitmix64 of the element's counter ((((seed * nout + i) * nin + j) * n + b) * n + a).
__global__ void k_sigma_gen_g(int n, int64_t ld, int i0, int ni, int nout, int nin, uint64_t seed,
                              const double *nu, double2 *G, int64_t gsi, int64_t gsj) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nn = (int64_t)n * n;
  const int ij = blockIdx.y;  // local (i, j)
  const int il = ij / nin, j = ij % nin;
  if (e >= nn || il >= ni) return;
  const int i = i0 + il;
  const int64_t arow = e % n, bcol = e / n;
  const uint64_t ctr = ((((seed * nout + i) * (uint64_t)nin + j) * n + bcol) * n + arow);
  const double sc = 1.0 / ((1.0 + nu[(int64_t)i * nin + j]) * sqrt((double)n));
  G[(int64_t)il * gsi + (int64_t)j * gsj + arow + bcol * ld] =
      mk(unit_sym(splitmix64(2 * ctr)) * sc, unit_sym(splitmix64(2 * ctr + 1)) * sc);
}

// W[j](a, b) = c_j s_a s_b exp(-|g_a - g_b|^2 / (2 L^2)) e^{i (th_a - th_b)},
// s_a = 1 / (1 + |g_a|^2), c_j = 1 / (1 + omega'_j^2): Hermitian positive
// semidefinite (a Gaussian kernel) and decaying as 1 / (1 + |G|^2).
__global__ void k_sigma_gen_w(int n, int64_t ld, int nin, const double *gvec, const double *theta,
                              const double *cw, double inv2l2, double2 *W, int64_t wsj) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nn = (int64_t)n * n;
  const int j = blockIdx.y;
  if (e >= nn || j >= nin) return;
  const int64_t ia = e % n, ib = e / n;
  const double *ga = gvec + 3 * ia, *gb = gvec + 3 * ib;
  const double dx = ga[0] - gb[0], dy = ga[1] - gb[1], dz = ga[2] - gb[2];
  const double g2a = ga[0] * ga[0] + ga[1] * ga[1] + ga[2] * ga[2];
  const double g2b = gb[0] * gb[0] + gb[1] * gb[1] + gb[2] * gb[2];
  const double mag = cw[j] / ((1.0 + g2a) * (1.0 + g2b)) * exp(-(dx * dx + dy * dy + dz * dz) * inv2l2);
  const double ph = theta[ia] - theta[ib];
  W[(int64_t)j * wsj + ia + ib * ld] = mk(mag * cos(ph), mag * sin(ph));
}

}  // namespace
}  // namespace sgw

using sgw::Error;

namespace {
template <typename F> int sigma_guarded(F &&f) {
  try {
    return f();
  } catch (const Error &e) {
    sgw::set_error(e.msg);
    return e.code;
  } catch (const std::exception &e) {
    sgw::set_error(e.what());
    return SGW_E_STATE;
  }
}
}  // namespace

extern "C" {

int sgw_sigma_conv_device(const sgw_sigma_conv_desc *desc, const sgw_complex *d_g,
                          const sgw_complex *d_w, const double *d_weight, sgw_complex alpha,
                          sgw_complex *d_sigma, void *stream) {
  return sigma_guarded([&] {
    SGW_REQUIRE(desc && d_g && d_w && d_weight && d_sigma, SGW_E_ARG,
                "sigma_conv: null argument");
    const sgw_sigma_conv_desc &c = *desc;
    SGW_REQUIRE(c.n >= 1 && c.ld >= c.n && c.nout >= 1 && c.nin >= 1, SGW_E_SHAPE,
                "sigma_conv: need n >= 1, ld >= n, nout >= 1, nin >= 1");
    SGW_REQUIRE(c.precision == SGW_SIGMA_FP64 || c.precision == SGW_SIGMA_FP32_SPLIT, SGW_E_ARG,
                "sigma_conv: unsupported precision mode");
    const int64_t mat = c.ld * (int64_t)c.n;
    SGW_REQUIRE(c.w_stride >= mat && c.s_stride >= mat, SGW_E_SHAPE,
                "sigma_conv: matrix strides smaller than ld * n");
    SGW_REQUIRE((c.g_stride_j >= mat && c.g_stride_i >= c.g_stride_j * c.nin) ||
                    (c.g_stride_i >= mat && c.g_stride_j >= c.g_stride_i * c.nout),
                SGW_E_SHAPE, "sigma_conv: G strides overlap");
    sgw::SigmaArgs a;
    a.n = c.n;
    a.ld = c.ld;
    a.nout = c.nout;
    a.nin = c.nin;
    a.gsi = c.g_stride_i;
    a.gsj = c.g_stride_j;
    a.wsj = c.w_stride;
    a.ssi = c.s_stride;
    a.G = reinterpret_cast<const double2 *>(d_g);
    a.W = reinterpret_cast<const double2 *>(d_w);
    a.wt = d_weight;
    a.alpha = make_double2(alpha.re, alpha.im);
    a.S = reinterpret_cast<double2 *>(d_sigma);
    a.accumulate = c.accumulate ? 1 : 0;
    a.tiles_m = (c.n + sgw::kSBM - 1) / sgw::kSBM;
    a.tiles_n = (c.n + sgw::kSBN - 1) / sgw::kSBN;
    const int64_t blocks = (int64_t)a.tiles_m * a.tiles_n * c.nout;
    SGW_REQUIRE(blocks < (1ll << 31) - 8, SGW_E_SHAPE, "sigma_conv: too many tiles");
    a.nblocks = (int)(((blocks + 7) / 8) * 8);
    hipStream_t st = (hipStream_t)stream;
    {
      if (c.precision == SGW_SIGMA_FP64) {
        sgw::KScope ks("sigma_f64", st);
        sgw::k_sigma_f64<<<a.nblocks, 256, 0, st>>>(a);
      } else {
        sgw::KScope ks("sigma_f32s", st);
        sgw::k_sigma_f32s<<<a.nblocks, 256, 0, st>>>(a);
      }
    }
    SGW_HIP(hipGetLastError());
    return 0;
  });
}

int sgw_sigma_synthetic_device(int32_t n, int64_t ld, int32_t i0, int32_t ni, int32_t nout,
                               int32_t nin, uint64_t seed, const double *d_nu,
                               sgw_complex *d_g, int64_t g_stride_i, int64_t g_stride_j,
                               const double *d_gvec, const double *d_theta, const double *d_cw,
                               double length, sgw_complex *d_w, int64_t w_stride, void *stream) {
  return sigma_guarded([&] {
    SGW_REQUIRE(n >= 1 && ld >= n && nout >= 1 && nin >= 1 && i0 >= 0 && ni >= 0 &&
                    i0 + ni <= nout,
                SGW_E_SHAPE, "sigma_synthetic: bad sizes");
    hipStream_t st = (hipStream_t)stream;
    const int64_t nn = (int64_t)n * n;
    const unsigned gx = (unsigned)((nn + 255) / 256);
    if (d_g && ni > 0) {
      SGW_REQUIRE(d_nu != nullptr, SGW_E_ARG, "sigma_synthetic: frequencies missing");
      SGW_REQUIRE((int64_t)ni * nin < 65536, SGW_E_SHAPE, "sigma_synthetic: too many matrices");
      sgw::k_sigma_gen_g<<<dim3(gx, (unsigned)(ni * nin)), 256, 0, st>>>(
          n, ld, i0, ni, nout, nin, seed, d_nu, reinterpret_cast<double2 *>(d_g), g_stride_i,
          g_stride_j);
      SGW_HIP(hipGetLastError());
    }
    if (d_w) {
      SGW_REQUIRE(d_gvec && d_theta && d_cw && length > 0 && nin < 65536, SGW_E_ARG,
                  "sigma_synthetic: W inputs missing");
      sgw::k_sigma_gen_w<<<dim3(gx, (unsigned)nin), 256, 0, st>>>(
          n, ld, nin, d_gvec, d_theta, d_cw, 0.5 / (length * length),
          reinterpret_cast<double2 *>(d_w), w_stride);
      SGW_HIP(hipGetLastError());
    }
    return 0;
  });
}

}  // extern "C"
