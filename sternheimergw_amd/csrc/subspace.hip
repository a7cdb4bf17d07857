// The SternheimerGW subspace solver on the GPU (priority 3 of select_solver).
//
// algo/linear_solver/src/linear_solver.f90:82-503: a Krylov subspace W = (A + sigma) V
// is grown for one shift, kept orthonormal by modified Gram-Schmidt
// (data/algebra/src/gram_schmidt.f90:35-134) and re-orthonormalised for every
// further shift after W += (sigma_new - sigma_old) V.  The sequential
// Gram-Schmidt sweeps run inside one workgroup each (they are latency bound,
// one block reduction per pair); the operator is the solver's Operator.
#include "subspace.h"

#include <algorithm>
#include <cmath>

namespace sgw {

namespace {

constexpr int NT = 512;

__device__ inline double2 wave_sum2(double2 v) {
  for (int off = 32; off > 0; off >>= 1) {
    v.x += __shfl_xor(v.x, off, 64);
    v.y += __shfl_xor(v.y, off, 64);
  }
  return v;
}

__device__ inline double2 bsum(double2 v, double2 *lds) {
  v = wave_sum2(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  if (threadIdx.x < 64) {
    double2 t = threadIdx.x < NT / 64 ? lds[threadIdx.x] : mk(0, 0);
    t = wave_sum2(t);
    if (threadIdx.x == 0) lds[0] = t;
  }
  __syncthreads();
  return lds[0];
}

// w_i += diff * v_i for all columns (linear_solver.f90:289-292)
__global__ void k_shift_basis(int64_t total, double2 diff, const double2 *__restrict__ vv,
                              double2 *__restrict__ ww) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  double2 w = ww[i];
  cmac(w, diff, vv[i]);
  ww[i] = w;
}

// gram_schmidt(first, basis, vector), first 1-based; one block.
__global__ __launch_bounds__(NT) void k_gram_schmidt(int first, int nb, int64_t n,
                                                     double2 *__restrict__ W,
                                                     double2 *__restrict__ V) {
  __shared__ double2 lds[NT / 64];
  for (int ib = first; ib <= nb; ++ib) {  // step 1
    double2 *wi = W + (int64_t)(ib - 1) * n, *vi = V + (int64_t)(ib - 1) * n;
    for (int jb = 1; jb <= first - 1; ++jb) {
      const double2 *wj = W + (int64_t)(jb - 1) * n, *vj = V + (int64_t)(jb - 1) * n;
      double2 acc = mk(0, 0);
      for (int64_t e = threadIdx.x; e < n; e += NT) cmacc(acc, wj[e], wi[e]);
      acc = bsum(acc, lds);
      const double2 m = cneg(acc);
      for (int64_t e = threadIdx.x; e < n; e += NT) {
        double2 a = wi[e];
        cmac(a, m, wj[e]);
        wi[e] = a;
        double2 b = vi[e];
        cmac(b, m, vj[e]);
        vi[e] = b;
      }
      __syncthreads();
    }
  }
  for (int ib = first; ib <= nb; ++ib) {  // steps 2-4
    double2 *wi = W + (int64_t)(ib - 1) * n, *vi = V + (int64_t)(ib - 1) * n;
    double2 acc = mk(0, 0);
    for (int64_t e = threadIdx.x; e < n; e += NT) {
      double2 a = wi[e];
      acc.x = fma(a.x, a.x, acc.x);
      acc.x = fma(a.y, a.y, acc.x);
    }
    acc = bsum(acc, lds);
    const double2 sc = cdiv(mk(1, 0), mk(sqrt(acc.x), 0.0));
    for (int64_t e = threadIdx.x; e < n; e += NT) {
      wi[e] = cmul(sc, wi[e]);
      vi[e] = cmul(sc, vi[e]);
    }
    __syncthreads();
    for (int jb = ib + 1; jb <= nb; ++jb) {
      double2 *wj = W + (int64_t)(jb - 1) * n, *vj = V + (int64_t)(jb - 1) * n;
      double2 d = mk(0, 0);
      for (int64_t e = threadIdx.x; e < n; e += NT) cmacc(d, wi[e], wj[e]);
      d = bsum(d, lds);
      const double2 m = cneg(d);
      for (int64_t e = threadIdx.x; e < n; e += NT) {
        double2 a = wj[e];
        cmac(a, m, wi[e]);
        wj[e] = a;
        double2 b = vj[e];
        cmac(b, m, vi[e]);
        vj[e] = b;
      }
      __syncthreads();
    }
  }
}

// overlap_i = (w_i, b) for all i; r = b - sum_i overlap_i w_i; norm(r)
// (linear_solver.f90:359-374).  One block.
__global__ __launch_bounds__(NT) void k_residual(int nb, int64_t n, const double2 *__restrict__ W,
                                                 const double2 *__restrict__ b,
                                                 double2 *__restrict__ r,
                                                 double2 *__restrict__ overlap,
                                                 double *__restrict__ norm) {
  __shared__ double2 lds[NT / 64];
  for (int64_t e = threadIdx.x; e < n; e += NT) r[e] = b[e];
  for (int ib = 0; ib < nb; ++ib) {
    const double2 *wi = W + (int64_t)ib * n;
    double2 acc = mk(0, 0);
    for (int64_t e = threadIdx.x; e < n; e += NT) cmacc(acc, wi[e], b[e]);
    acc = bsum(acc, lds);
    if (threadIdx.x == 0) overlap[ib] = acc;
    for (int64_t e = threadIdx.x; e < n; e += NT) r[e] = csub(r[e], cmul(acc, wi[e]));
  }
  double2 s = mk(0, 0);
  for (int64_t e = threadIdx.x; e < n; e += NT) {
    double2 a = r[e];
    s.x = fma(a.x, a.x, s.x);
    s.x = fma(a.y, a.y, s.x);
  }
  s = bsum(s, lds);
  if (threadIdx.x == 0) *norm = sqrt(s.x);
}

// x = sum_i (w_i, b) v_i  (linear_solver.f90:485-501).  One block.
__global__ __launch_bounds__(NT) void k_result(int nb, int64_t n, const double2 *__restrict__ W,
                                               const double2 *__restrict__ V,
                                               const double2 *__restrict__ b,
                                               double2 *__restrict__ x, int32_t *__restrict__ nan) {
  __shared__ double2 lds[NT / 64];
  for (int64_t e = threadIdx.x; e < n; e += NT) x[e] = mk(0, 0);
  for (int ib = 0; ib < nb; ++ib) {
    const double2 *wi = W + (int64_t)ib * n, *vi = V + (int64_t)ib * n;
    double2 acc = mk(0, 0);
    for (int64_t e = threadIdx.x; e < n; e += NT) cmacc(acc, wi[e], b[e]);
    acc = bsum(acc, lds);
    for (int64_t e = threadIdx.x; e < n; e += NT) {
      double2 a = x[e];
      cmac(a, acc, vi[e]);
      x[e] = a;
    }
  }
  int f = 0;
  for (int64_t e = threadIdx.x; e < n; e += NT)
    if (isnan(x[e].x) || isnan(x[e].y)) f = 1;
  if (f) *nan = 1;
}

}  // namespace

int SubspaceSolver::solve(Operator &op, hipStream_t st, int32_t group, const double2 *d_b,
                          int64_t n, const double2 *h_sigma, int nshift, int max_iter,
                          double threshold, double *abs_threshold_out, double2 *d_x,
                          int64_t x_ld, int *basis_out) {
  SGW_REQUIRE(x_ld >= n, SGW_E_SHAPE, "linear_solver: solution leading dimension < n");
  int cap = 16, nb = 0, ierr = 0;
  DevBuf<double2> V((size_t)n * cap), W((size_t)n * cap), r(n), newv(n), ov(std::max(cap, 1));
  DevBuf<double> dnorm(1);
  DevBuf<double2> dsig(1);
  DevBuf<int32_t> dgroup(1), dactive(1), dnan(1);
  int32_t one = 1;
  dgroup.upload(&group, 1, st);
  op.prepare(&group, 1);
  dactive.upload(&one, 1, st);
  dnan.zero(st);
  // abs_threshold = threshold * norm(b)   (linear_solver.f90:200-215)
  double bn = 0;
  {
    DevBuf<double2> ovb(1);
    k_residual<<<1, NT, 0, st>>>(0, n, W.p, d_b, r.p, ovb.p, dnorm.p);
    SGW_HIP(hipMemcpyAsync(&bn, dnorm.p, sizeof(double), hipMemcpyDeviceToHost, st));
    SGW_HIP(hipStreamSynchronize(st));
  }
  const double abs_thr = threshold * bn;
  if (abs_threshold_out) *abs_threshold_out = abs_thr;
  double2 sub_sigma = mk(0, 0);
  for (int ishift = 0; ishift < nshift; ++ishift) {
    const double2 sg = h_sigma[ishift];
    // orthogonal_subspace (:272-300)
    const double2 diff = csub(sg, sub_sigma);
    if (nb > 0) {
      int64_t total = (int64_t)n * nb;
      k_shift_basis<<<(unsigned)((total + 255) / 256), 256, 0, st>>>(total, diff, V.p, W.p);
      k_gram_schmidt<<<1, NT, 0, st>>>(1, nb, n, W.p, V.p);
    }
    sub_sigma = sg;
    dsig.upload(&sg, 1, st);
    bool conv = false;
    for (int iter = 1; iter <= max_iter; ++iter) {
      if ((size_t)std::max(nb, 1) > ov.n) ov.alloc(2 * ov.n);
      k_residual<<<1, NT, 0, st>>>(nb, n, W.p, d_b, r.p, ov.p, dnorm.p);
      double nr = 0;
      SGW_HIP(hipMemcpyAsync(&nr, dnorm.p, sizeof(double), hipMemcpyDeviceToHost, st));
      SGW_HIP(hipStreamSynchronize(st));
      conv = nr < abs_thr;
      if (!conv && n == nb) {
        std::printf("WARNING: SternheimerGW linear solver right-hand side cannot be mapped by "
                    "complete basis\n");
        ierr = 3;
      } else {
        ierr = 0;
      }
      if (conv || ierr != 0) break;
      op.apply(st, 1, r.p, n, newv.p, n, dsig.p, &sg, dactive.p, &one, dgroup.p);  // :168
      if (nb + 1 > cap) {  // expand storage (MOVE_ALLOC, :424-433)
        int ncap = cap * 2;
        DevBuf<double2> V2((size_t)n * ncap), W2((size_t)n * ncap);
        SGW_HIP(hipMemcpyAsync(V2.p, V.p, sizeof(double2) * n * nb, hipMemcpyDeviceToDevice, st));
        SGW_HIP(hipMemcpyAsync(W2.p, W.p, sizeof(double2) * n * nb, hipMemcpyDeviceToDevice, st));
        SGW_HIP(hipStreamSynchronize(st));
        V = std::move(V2);
        W = std::move(W2);
        cap = ncap;
      }
      SGW_HIP(hipMemcpyAsync(W.p + (int64_t)nb * n, newv.p, sizeof(double2) * n,
                             hipMemcpyDeviceToDevice, st));
      SGW_HIP(hipMemcpyAsync(V.p + (int64_t)nb * n, r.p, sizeof(double2) * n,
                             hipMemcpyDeviceToDevice, st));
      ++nb;
      k_gram_schmidt<<<1, NT, 0, st>>>(nb, nb, n, W.p, V.p);  // :438
    }
    if (!conv) {  // :176-180
      std::printf("WARNING: SternheimerGW linear solver did not converge\n");
      ierr = 1;
      break;
    }
    k_result<<<1, NT, 0, st>>>(nb, n, W.p, V.p, d_b, d_x + (int64_t)ishift * x_ld, dnan.p);
    int32_t hn = 0;
    SGW_HIP(hipMemcpyAsync(&hn, dnan.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    SGW_HIP(hipStreamSynchronize(st));
    if (hn) {  // :186-190
      std::printf("WARNING: SternheimerGW linear solver produced at least one NaN\n");
      ierr = 2;
      break;
    }
  }
  SGW_HIP(hipStreamSynchronize(st));
  if (basis_out) *basis_out = nb;
  return ierr;
}

}  // namespace sgw
