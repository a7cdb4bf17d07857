// Mixed-precision multishift solve: FP32 local-potential FFTs in the inner
// solves, FP64 iterative refinement (no counterpart in the reference, whose
// solver is FP64 throughout; the FP64 steps follow bicgstab.f90:154-275).
//
//   1. x = multishift BiCGStab(l) of (A + sigma_k) x_k = b with the FP32 FFT
//      H-apply, to the inner threshold (the shifts share the seed's Krylov
//      space because every shift starts from the residual b);
//   2. r_k = b - (A + sigma_k) x_k in FP64 for every shift;
//   3. if some r_k is above the target: the shifts' residuals are no longer
//      collinear (their FP32 errors differ), so each (system, shift) becomes
//      its own single-shift system (A + sigma_k) d_k = r_k, solved with the
//      FP32 FFT H-apply (streamed through the solver's slots), and
//      x_k += d_k; back to 2.
//
// Vectors, recurrence scalars, the projector terms (beta and alpha_pv P_v)
// and every residual are FP64; only the local-potential FFTs are FP32.
#include "handles.h"
#include "ktimer.h"
#include "operators.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace sgw {

namespace {

constexpr int kB = 256;

// r[v] = rhs[s] - ax[v] and ||r[v]|| / ||rhs[s]||, v = s*nshift + k
__global__ __launch_bounds__(kB) void k_resid_vec(int nshift, int64_t ld,
                                                  const double2 *__restrict__ rhs,
                                                  const double2 *__restrict__ ax,
                                                  double2 *__restrict__ r,
                                                  double *__restrict__ res) {
  __shared__ double2 lds[kB / 64];
  const int v = blockIdx.x;
  const int s = v / nshift;
  const double2 *b = rhs + (int64_t)s * ld, *a = ax + (int64_t)v * ld;
  double2 *rv = r + (int64_t)v * ld;
  double2 acc = mk(0, 0);
  for (int64_t i = threadIdx.x; i < ld; i += kB) {
    const double2 d = csub(b[i], a[i]);
    rv[i] = d;
    acc.x = fma(d.x, d.x, acc.x);
    acc.x = fma(d.y, d.y, acc.x);
    acc.y = fma(b[i].x, b[i].x, acc.y);
    acc.y = fma(b[i].y, b[i].y, acc.y);
  }
  for (int off = 32; off > 0; off >>= 1) {
    acc.x += __shfl_xor(acc.x, off, 64);
    acc.y += __shfl_xor(acc.y, off, 64);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) lds[wid] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double2 t = mk(0, 0);
    for (int w = 0; w < kB / 64; ++w) t = cadd(t, lds[w]);
    res[v] = t.y > 0 ? sqrt(t.x / t.y) : sqrt(t.x);
  }
}

// x[v] += d[v] for the listed vectors (v = s*nshift + k)
__global__ __launch_bounds__(kB) void k_add_corr(const int32_t *__restrict__ list, int64_t ld,
                                                 const double2 *__restrict__ d,
                                                 double2 *__restrict__ x) {
  const int i = blockIdx.y;
  const int64_t e = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (e >= ld) return;
  const int v = list[i];
  double2 t = x[(int64_t)v * ld + e];
  t = cadd(t, d[(int64_t)i * ld + e]);
  x[(int64_t)v * ld + e] = t;
}

// b[i] <- r[list[i]] (the single-shift right-hand sides), sigma likewise
__global__ __launch_bounds__(kB) void k_pick(const int32_t *__restrict__ list, int64_t ld,
                                             const double2 *__restrict__ r,
                                             const double2 *__restrict__ sig_all,
                                             double2 *__restrict__ b, double2 *__restrict__ sig) {
  const int i = blockIdx.y;
  const int64_t e = (int64_t)blockIdx.x * kB + threadIdx.x;
  const int v = list[i];
  if (blockIdx.x == 0 && threadIdx.x == 0) sig[i] = sig_all[v];
  if (e < ld) b[(int64_t)i * ld + e] = r[(int64_t)v * ld + e];
}

}  // namespace

}  // namespace sgw

using namespace sgw;

extern "C" int sgw_sternheimer_solve_refined(sgw_batch_solver *solver, const int32_t *d_group,
                                             const sgw_complex *d_rhs, const sgw_complex *d_sigma,
                                             const sgw_complex *h_sigma, sgw_complex *d_x,
                                             double inner_threshold, int32_t max_refine,
                                             int32_t correction_slots, int32_t *h_ierr,
                                             int64_t *h_stats, double *h_res, void *stream) {
  try {
    SGW_REQUIRE(solver && solver->op && d_rhs && d_sigma && h_sigma && d_x, SGW_E_ARG,
                "solve_refined: null argument");
    auto *pw = dynamic_cast<PwOperator *>(solver->op->op.get());
    SGW_REQUIRE(pw != nullptr, SGW_E_ARG, "solve_refined: plane-wave operators only");
    SGW_REQUIRE(inner_threshold > 0 && max_refine >= 0 && correction_slots >= 1, SGW_E_ARG,
                "solve_refined: need inner_threshold > 0, max_refine >= 0, slots >= 1");
    const sgw_batch_config &c = solver->cfg;
    const int S = c.nsys, K = c.nshift;
    const int64_t ld = c.ld, nv = (int64_t)S * K;
    hipStream_t st = (hipStream_t)stream;
    const double2 *rhs = reinterpret_cast<const double2 *>(d_rhs);
    const double2 *sig = reinterpret_cast<const double2 *>(d_sigma);
    double2 *x = reinterpret_cast<double2 *>(d_x);
    const int mode0 = pw->fft_precision();
    struct Restore {
      PwOperator *pw;
      int mode;
      ~Restore() { pw->set_fft_precision(mode, nullptr); }
    } restore{pw, mode0};
    std::vector<int32_t> hg(S, 0), it(S), e(S);
    if (d_group) {
      SGW_HIP(hipMemcpyAsync(hg.data(), d_group, sizeof(int32_t) * S, hipMemcpyDeviceToHost, st));
      SGW_HIP(hipStreamSynchronize(st));
    }
    int64_t stats[6] = {0, 0, 0, 0, 0, 0};  // FP32 H-applies (multishift, corrections),
                                            // FP64 residual applications, refinements,
                                            // multishift iterations, correction iterations
    // 1. the multishift solve with FP32 FFTs
    pw->set_fft_precision(1, nullptr);
    solver->engine->solve(*pw, st, rhs, ld, sig, reinterpret_cast<const double2 *>(h_sigma),
                          d_group, inner_threshold, c.max_iter, x, ld, e.data(), it.data(), 1,
                          c.shift_test, 0.0);
    for (int s = 0; s < S; ++s) stats[4] += it[s];
    stats[0] += 2LL * c.lmax * stats[4];
    // 2-3. residuals in FP64, single-shift corrections with FP32 FFTs
    DevBuf<double2> ax((size_t)nv * ld), r((size_t)nv * ld);
    DevBuf<double> res(nv);
    DevBuf<int32_t> gv(nv);
    std::vector<int32_t> hgv(nv);
    for (int64_t v = 0; v < nv; ++v) hgv[v] = hg[v / K];
    gv.upload(hgv.data(), nv, st);
    std::vector<double> hres(nv);
    // the target on every shift's relative residual; the corrections aim a
    // tenth below it (absolute, on the smallest right-hand side)
    const double target = c.threshold;
    std::vector<double> bnorm(S, 1.0);
    std::unique_ptr<BatchBicgstab> corr;
    int rounds = 0;
    for (;; ++rounds) {
      pw->set_fft_precision(0, nullptr);
      pw->prepare(hgv.data(), (int)nv);
      pw->apply(st, (int)nv, x, ld, ax.p, ld, sig, nullptr, nullptr, nullptr, gv.p);
      k_resid_vec<<<(unsigned)nv, kB, 0, st>>>(K, ld, rhs, ax.p, r.p, res.p);
      SGW_HIP(hipGetLastError());
      SGW_HIP(hipMemcpyAsync(hres.data(), res.p, sizeof(double) * nv, hipMemcpyDeviceToHost, st));
      SGW_HIP(hipStreamSynchronize(st));
      stats[2] += nv;
      std::vector<int32_t> todo;
      for (int64_t v = 0; v < nv; ++v)
        if (!(hres[v] <= target)) todo.push_back((int32_t)v);
      if (todo.empty() || rounds >= max_refine) break;
      // the right-hand sides' norms (once): ||b|| = ||r|| / res at x = 0 is
      // not available, so take them from the device once
      if (rounds == 0) {
        std::vector<double2> hb((size_t)S * ld);
        SGW_HIP(hipMemcpyAsync(hb.data(), rhs, sizeof(double2) * S * ld, hipMemcpyDeviceToHost,
                               st));
        SGW_HIP(hipStreamSynchronize(st));
        for (int s = 0; s < S; ++s) {
          double a = 0;
          for (int64_t i = 0; i < ld; ++i)
            a += hb[(size_t)s * ld + i].x * hb[(size_t)s * ld + i].x +
                 hb[(size_t)s * ld + i].y * hb[(size_t)s * ld + i].y;
          bnorm[s] = std::sqrt(a);
        }
      }
      double bmin = 1e300;
      for (int s = 0; s < S; ++s) bmin = std::min(bmin, bnorm[s]);
      const int nq = (int)todo.size();
      DevBuf<int32_t> list(nq);
      list.upload(todo.data(), nq, st);
      DevBuf<double2> b1((size_t)nq * ld), s1(nq), d1((size_t)nq * ld);
      const unsigned tiles = (unsigned)((ld + kB - 1) / kB);
      k_pick<<<dim3(tiles, nq), kB, 0, st>>>(list.p, ld, r.p, sig, b1.p, s1.p);
      std::vector<double2> hs1(nq);
      SGW_HIP(hipMemcpyAsync(hs1.data(), s1.p, sizeof(double2) * nq, hipMemcpyDeviceToHost, st));
      SGW_HIP(hipStreamSynchronize(st));
      std::vector<int32_t> hq(nq), qe(nq), qi(nq);
      for (int i = 0; i < nq; ++i) hq[i] = hg[todo[i] / K];
      const int slots = std::min(nq, (int)correction_slots);
      if (!corr) corr.reset(new BatchBicgstab(slots, 1, c.lmax, ld));
      pw->set_fft_precision(1, nullptr);
      corr->solve_stream(*pw, st, nq, b1.p, ld, s1.p, hs1.data(), hq.data(),
                         0.1 * target * bmin, c.max_iter, d1.p, ld, qe.data(), qi.data(), 0);
      int64_t ci = 0;
      for (int i = 0; i < nq; ++i) ci += qi[i];
      stats[5] += ci;
      stats[1] += 2LL * c.lmax * ci;
      k_add_corr<<<dim3(tiles, nq), kB, 0, st>>>(list.p, ld, d1.p, x);
      SGW_HIP(hipGetLastError());
    }
    stats[3] = rounds;
    int worst = 0;
    for (int s = 0; s < S; ++s) {
      int ierr = 0;
      for (int k = 0; k < K; ++k) {
        const double rv = hres[(size_t)s * K + k];
        if (std::isnan(rv)) ierr = 2;
        else if (!(rv <= target) && ierr == 0) ierr = 1;
      }
      if (h_ierr) h_ierr[s] = ierr;
      worst = std::max(worst, ierr);
    }
    if (h_stats) std::copy(stats, stats + 6, h_stats);
    if (h_res) std::copy(hres.begin(), hres.end(), h_res);
    return worst;
  } catch (const Error &err) {
    set_error(err.msg);
    return err.code;
  }
}
