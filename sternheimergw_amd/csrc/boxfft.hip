// Whole-box FFTs: two passes per transform.
//  * k_box_xy: one workgroup per (box, z-plane); the n x n plane is staged
//    through LDS (rows padded to n+1 against bank conflicts) and n lanes do
//    the x lines, then the y lines, from registers.
//  * k_box_z: one thread per (box, y, x) column with the n points of its z
//    line in registers; adjacent lanes take adjacent x, so every load and
//    store of a wave is one contiguous run.
#include "boxfft.h"
#include "regfft.h"

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

template <int N, int SIGN> void run(hipStream_t st, int64_t nvec, double2 *d) {
  if (nvec <= 0) return;
  k_box_xy<N, SIGN><<<(unsigned)(nvec * N), 64, 0, st>>>(d);
  k_box_z<N, SIGN><<<(unsigned)((nvec * N * N + 63) / 64), 64, 0, st>>>(nvec, d);
}

template <int SIGN> void dispatch(hipStream_t st, int n, int64_t nvec, double2 *d) {
  switch (n) {
    case 16: run<16, SIGN>(st, nvec, d); break;
    case 18: run<18, SIGN>(st, nvec, d); break;
    case 24: run<24, SIGN>(st, nvec, d); break;
    case 27: run<27, SIGN>(st, nvec, d); break;
    case 32: run<32, SIGN>(st, nvec, d); break;
    case 36: run<36, SIGN>(st, nvec, d); break;
    case 48: run<48, SIGN>(st, nvec, d); break;
    default:
      throw Error{SGW_E_SHAPE, std::string("box FFT: cubic boxes of ") + box_sizes_text() +
                                   " points only"};
  }
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

}  // namespace sgw
