// Whole-box FFTs: two passes per transform.
//  * k_box_xy: one workgroup per (box, z-plane); the n x n plane is staged
//    through LDS (rows padded to n+1 against bank conflicts) and n lanes do
//    the x lines, then the y lines, from registers.
//  * k_box_z: one thread per (box, y, x) column with the n points of its z
//    line in registers; adjacent lanes take adjacent x, so every load and
//    store of a wave is one contiguous run.
#include "boxfft.h"
#include "regfft.h"

#include <type_traits>

namespace sgw {
namespace {

template <int N, int SIGN>
__global__ __launch_bounds__(64) void k_box_xy(double2 *__restrict__ d) {
  constexpr int P = N + 1;
  __shared__ double2 s[N * P];
  double2 *g = d + (int64_t)blockIdx.x * N * N;
  for (int e = threadIdx.x; e < N * N; e += 64) s[(e / N) * P + e % N] = g[e];
  __syncthreads();
  const int t = threadIdx.x;
  if (t < N) {  // x line y = t
    double2 x[N];
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] = s[t * P + i];
    reg_fft<N, SIGN>(x);
#pragma unroll
    for (int i = 0; i < N; ++i) s[t * P + i] = x[i];
  }
  __syncthreads();
  if (t < N) {  // y line x = t
    double2 x[N];
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] = s[i * P + t];
    reg_fft<N, SIGN>(x);
#pragma unroll
    for (int i = 0; i < N; ++i) s[i * P + t] = x[i];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < N * N; e += 64) g[e] = s[(e / N) * P + e % N];
}

template <int N, int SIGN>
__global__ __launch_bounds__(64) void k_box_z(int64_t nvec, double2 *__restrict__ d) {
  const int64_t t = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (t >= nvec * N * N) return;
  const int64_t v = t / (N * N), c = t % (N * N);
  double2 *p = d + v * N * N * N + c;
  double2 x[N];
#pragma unroll
  for (int i = 0; i < N; ++i) x[i] = p[(int64_t)i * N * N];
  reg_fft<N, SIGN>(x);
#pragma unroll
  for (int i = 0; i < N; ++i) p[(int64_t)i * N * N] = x[i];
}

// The middle of a real-space product between two box transforms: for every
// z line, d <- Z_{+1}(g * Z_{-1}(d)).  With the xy passes around it,
// xy(-1), this, xy(+1) is the full inverse transform, the pointwise product
// with g and the full forward transform: the 3-D DFT is a product of
// commuting passes along the axes.  g has the layout of d.
template <int N>
__global__ __launch_bounds__(64) void k_box_z_mul_z(int64_t nvec, double2 *__restrict__ d,
                                                    const double2 *__restrict__ g) {
  const int64_t t = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (t >= nvec * N * N) return;
  const int64_t v = t / (N * N), c = t % (N * N);
  double2 *p = d + v * N * N * N + c;
  const double2 *q = g + v * N * N * N + c;
  double2 x[N];
#pragma unroll
  for (int i = 0; i < N; ++i) x[i] = p[(int64_t)i * N * N];
  reg_fft<N, -1>(x);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const double2 w = q[(int64_t)i * N * N];
    x[i] = make_double2(x[i].x * w.x - x[i].y * w.y, x[i].x * w.y + x[i].y * w.x);
  }
  reg_fft<N, +1>(x);
#pragma unroll
  for (int i = 0; i < N; ++i) p[(int64_t)i * N * N] = x[i];
}

// xy pass of box r' = blockIdx.x / N, plane z = blockIdx.x % N, whose input
// holds only the values S[g][r'] at the box positions of the plane's list
// (pos[k], k in [start[z], start[z + 1])) and zeros elsewhere: the plane is
// built in LDS, so the zero-filled box never passes through HBM.
template <int N, int SIGN>
__global__ __launch_bounds__(64) void k_box_xy_scatter(const double2 *__restrict__ S,
                                                       int64_t nr_p,
                                                       const int32_t *__restrict__ start,
                                                       const int32_t *__restrict__ gl,
                                                       const int32_t *__restrict__ pos,
                                                       double2 *__restrict__ d) {
  constexpr int P = N + 1;
  __shared__ double2 s[N * P];
  const int64_t box = blockIdx.x / N;
  const int z = blockIdx.x % N;
  double2 *g = d + (int64_t)blockIdx.x * N * N;
  for (int e = threadIdx.x; e < N * P; e += 64) s[e] = make_double2(0, 0);
  __syncthreads();
  for (int k = start[z] + threadIdx.x; k < start[z + 1]; k += 64) {
    const int e = pos[k];
    s[(e / N) * P + e % N] = S[(int64_t)gl[k] * nr_p + box];
  }
  __syncthreads();
  const int t = threadIdx.x;
  if (t < N) {  // x line y = t
    double2 x[N];
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] = s[t * P + i];
    reg_fft<N, SIGN>(x);
#pragma unroll
    for (int i = 0; i < N; ++i) s[t * P + i] = x[i];
  }
  __syncthreads();
  if (t < N) {  // y line x = t
    double2 x[N];
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] = s[i * P + t];
    reg_fft<N, SIGN>(x);
#pragma unroll
    for (int i = 0; i < N; ++i) s[i * P + t] = x[i];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < N * N; e += 64) g[e] = s[(e / N) * P + e % N];
}

// xy pass of a plane whose result is wanted only at the plane's list
// positions: S[g][r'] = scale * (transformed plane)[pos]; the plane is not
// written back.
template <int N, int SIGN>
__global__ __launch_bounds__(64) void k_box_xy_gather(const double2 *__restrict__ d,
                                                      int64_t nr_p,
                                                      const int32_t *__restrict__ start,
                                                      const int32_t *__restrict__ gl,
                                                      const int32_t *__restrict__ pos,
                                                      double scale, double2 *__restrict__ S) {
  constexpr int P = N + 1;
  __shared__ double2 s[N * P];
  const int64_t box = blockIdx.x / N;
  const int z = blockIdx.x % N;
  const double2 *g = d + (int64_t)blockIdx.x * N * N;
  for (int e = threadIdx.x; e < N * N; e += 64) s[(e / N) * P + e % N] = g[e];
  __syncthreads();
  const int t = threadIdx.x;
  if (t < N) {  // x line y = t
    double2 x[N];
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] = s[t * P + i];
    reg_fft<N, SIGN>(x);
#pragma unroll
    for (int i = 0; i < N; ++i) s[t * P + i] = x[i];
  }
  __syncthreads();
  if (t < N) {  // y line x = t
    double2 x[N];
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] = s[i * P + t];
    reg_fft<N, SIGN>(x);
#pragma unroll
    for (int i = 0; i < N; ++i) s[i * P + t] = x[i];
  }
  __syncthreads();
  for (int k = start[z] + threadIdx.x; k < start[z + 1]; k += 64) {
    const int e = pos[k];
    const double2 v = s[(e / N) * P + e % N];
    S[(int64_t)gl[k] * nr_p + box] = make_double2(scale * v.x, scale * v.y);
  }
}

template <int N, int SIGN> void run(hipStream_t st, int64_t nvec, double2 *d) {
  if (nvec <= 0) return;
  k_box_xy<N, SIGN><<<(unsigned)(nvec * N), 64, 0, st>>>(d);
  k_box_z<N, SIGN><<<(unsigned)((nvec * N * N + 63) / 64), 64, 0, st>>>(nvec, d);
}

// f(std::integral_constant<int, n>) for a supported box side n
template <typename F> void with_side(int n, F &&f) {
  switch (n) {
    case 16: f(std::integral_constant<int, 16>{}); break;
    case 18: f(std::integral_constant<int, 18>{}); break;
    case 24: f(std::integral_constant<int, 24>{}); break;
    case 27: f(std::integral_constant<int, 27>{}); break;
    case 32: f(std::integral_constant<int, 32>{}); break;
    case 36: f(std::integral_constant<int, 36>{}); break;
    case 48: f(std::integral_constant<int, 48>{}); break;
    default:
      throw Error{SGW_E_SHAPE, std::string("box FFT: cubic boxes of ") + box_sizes_text() +
                                   " points only"};
  }
}

template <int SIGN> void dispatch(hipStream_t st, int n, int64_t nvec, double2 *d) {
  with_side(n, [&](auto side) { run<decltype(side)::value, SIGN>(st, nvec, d); });
}

}  // namespace

bool box_supported(int n) {
  return n == 16 || n == 18 || n == 24 || n == 27 || n == 32 || n == 36 || n == 48;
}

const char *box_sizes_text() { return "16, 18, 24, 27, 32, 36 or 48"; }

void box_fft(hipStream_t st, int n, int64_t nvec, double2 *d, int sign) {
  if (sign > 0)
    dispatch<1>(st, n, nvec, d);
  else
    dispatch<-1>(st, n, nvec, d);
}

void box_fft_xy(hipStream_t st, int n, int64_t nvec, double2 *d, int sign) {
  if (nvec <= 0) return;
  with_side(n, [&](auto side) {
    constexpr int N = decltype(side)::value;
    if (sign > 0)
      k_box_xy<N, 1><<<(unsigned)(nvec * N), 64, 0, st>>>(d);
    else
      k_box_xy<N, -1><<<(unsigned)(nvec * N), 64, 0, st>>>(d);
  });
}

void box_xy_scatter(hipStream_t st, int n, int64_t nvec, const double2 *S, const int32_t *start,
                    const int32_t *gl, const int32_t *pos, double2 *d, int sign) {
  if (nvec <= 0) return;
  with_side(n, [&](auto side) {
    constexpr int N = decltype(side)::value;
    if (sign > 0)
      k_box_xy_scatter<N, 1><<<(unsigned)(nvec * N), 64, 0, st>>>(S, nvec, start, gl, pos, d);
    else
      k_box_xy_scatter<N, -1><<<(unsigned)(nvec * N), 64, 0, st>>>(S, nvec, start, gl, pos, d);
  });
}

void box_xy_gather(hipStream_t st, int n, int64_t nvec, const double2 *d, const int32_t *start,
                   const int32_t *gl, const int32_t *pos, double scale, double2 *S, int sign) {
  if (nvec <= 0) return;
  with_side(n, [&](auto side) {
    constexpr int N = decltype(side)::value;
    if (sign > 0)
      k_box_xy_gather<N, 1><<<(unsigned)(nvec * N), 64, 0, st>>>(d, nvec, start, gl, pos, scale,
                                                                 S);
    else
      k_box_xy_gather<N, -1><<<(unsigned)(nvec * N), 64, 0, st>>>(d, nvec, start, gl, pos, scale,
                                                                  S);
  });
}

void box_z_product(hipStream_t st, int n, int64_t nvec, double2 *d, const double2 *g) {
  if (nvec <= 0) return;
  with_side(n, [&](auto side) {
    constexpr int N = decltype(side)::value;
    k_box_z_mul_z<N><<<(unsigned)((nvec * N * N + 63) / 64), 64, 0, st>>>(nvec, d, g);
  });
}

}  // namespace sgw
