// Internal definitions shared by the HIP translation units of libsgw_hip.
#pragma once

#include <hip/hip_runtime.h>

#include "sgw.h"

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace sgw {

// ---------------------------------------------------------------------------
// error state
// ---------------------------------------------------------------------------
void set_error(const std::string &msg);
const char *last_error();

struct Error {
  int code;
  std::string msg;
};

#define SGW_HIP(expr)                                                        \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess) {                                                  \
      throw ::sgw::Error{-3, std::string(#expr) + ": " + hipGetErrorString(e_)}; \
    }                                                                        \
  } while (0)

#define SGW_REQUIRE(cond, code, msg)                                         \
  do {                                                                       \
    if (!(cond)) throw ::sgw::Error{(code), (msg)};                          \
  } while (0)

// API status codes (SGW_E_*, negative) come from the public header;
// non-negative values are solver ierr codes.

// ---------------------------------------------------------------------------
// complex arithmetic on double2 / float2 (x = re, y = im)
// ---------------------------------------------------------------------------
template <typename T> struct cplx_t;
template <> struct cplx_t<double> { using type = double2; };
template <> struct cplx_t<float> { using type = float2; };
template <typename T> using cplx = typename cplx_t<T>::type;

__host__ __device__ inline double2 mk(double re, double im) { return make_double2(re, im); }
__host__ __device__ inline float2 mkf(float re, float im) { return make_float2(re, im); }
__host__ __device__ inline double2 cadd(double2 a, double2 b) { return mk(a.x + b.x, a.y + b.y); }
__host__ __device__ inline double2 csub(double2 a, double2 b) { return mk(a.x - b.x, a.y - b.y); }
__host__ __device__ inline double2 cmul(double2 a, double2 b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline double2 cneg(double2 a) { return mk(-a.x, -a.y); }
__host__ __device__ inline double2 cconj(double2 a) { return mk(a.x, -a.y); }
// Smith's algorithm
__host__ __device__ inline double2 cdiv(double2 a, double2 b) {
  if (fabs(b.x) >= fabs(b.y)) {
    double r = b.y / b.x, d = b.x + b.y * r;
    return mk((a.x + a.y * r) / d, (a.y - a.x * r) / d);
  } else {
    double r = b.x / b.y, d = b.y + b.x * r;
    return mk((a.x * r + a.y) / d, (a.y * r - a.x) / d);
  }
}
__host__ __device__ inline float2 cadd(float2 a, float2 b) { return mkf(a.x + b.x, a.y + b.y); }
__host__ __device__ inline float2 csub(float2 a, float2 b) { return mkf(a.x - b.x, a.y - b.y); }
__host__ __device__ inline float2 cmul(float2 a, float2 b) {
  return mkf(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline float2 cconj(float2 a) { return mkf(a.x, -a.y); }

// acc += a * b (complex multiply-accumulate written as explicit FMAs)
__host__ __device__ inline void cmac(double2 &acc, double2 a, double2 b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
}
__host__ __device__ inline void cmac(float2 &acc, float2 a, float2 b) {
  acc.x = fmaf(a.x, b.x, acc.x);
  acc.x = fmaf(-a.y, b.y, acc.x);
  acc.y = fmaf(a.x, b.y, acc.y);
  acc.y = fmaf(a.y, b.x, acc.y);
}
// acc += conj(a) * b
__host__ __device__ inline void cmacc(double2 &acc, double2 a, double2 b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(-a.y, b.x, acc.y);
}

// ---------------------------------------------------------------------------
// device buffers (RAII)
// ---------------------------------------------------------------------------
template <typename T> struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    if (count == 0) return;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      throw Error{SGW_E_ALLOC, std::string("hipMalloc of ") + std::to_string(count * sizeof(T)) +
                                   " bytes failed: " + hipGetErrorString(e)};
    }
    n = count;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void upload(const T *h, size_t count, hipStream_t st = 0) {
    SGW_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, st));
  }
  void zero(hipStream_t st = 0) {
    if (n) SGW_HIP(hipMemsetAsync(p, 0, n * sizeof(T), st));
  }
};

// ---------------------------------------------------------------------------
// Operator interface: out_v = (A + sigma_v) in_v for every active vector v.
// Vectors v live at in + v*in_stride (length n, padded length ld).  group[v]
// selects the operator instance (k+q point) of vector v.
// ---------------------------------------------------------------------------
struct Operator {
  int64_t n = 0;   // logical vector length (single-group operators)
  int64_t ld = 0;  // storage length per vector
  virtual ~Operator() {}
  // every pointer is device memory; h_sigma is a host mirror of d_sigma
  // (operators that run on the host need it).
  virtual void apply(hipStream_t st, int nvec, const double2 *in, int64_t in_stride,
                     double2 *out, int64_t out_stride, const double2 *d_sigma,
                     const double2 *h_sigma, const int32_t *d_active,
                     const int32_t *h_active, const int32_t *d_group) = 0;
  // number of valid entries of vector v's group (padding beyond is zero)
  // Called with the host copy of the per-vector group indices before a
  // sequence of apply() calls on nvec vectors (a solve); operators that work
  // on groups of vectors build their plans here.
  virtual void prepare(const int32_t *h_group, int nvec) { (void)h_group; (void)nvec; }
  virtual int64_t group_n(int g) const { (void)g; return n; }
  virtual int ngroups() const { return 1; }
  virtual bool host_side() const { return false; }
};

}  // namespace sgw
