// Adjoint of the batched Sternheimer solve: the gradient of
//   L = sum_{s,k} w_sk Re <c_s | x_sk>,   (H_{g(s)} + sigma_sk) x_sk = b_s,
// with respect to the local potential V_loc(r) of the plane-wave operator.
//
// dL = Re c^H dx = -Re y^H dH x with (H + sigma)^H y = c, and H is Hermitian,
// so the backward problem is the same batched multishift solve at the
// conjugate shifts conj(sigma_sk) on right-hand sides c_s.  The local term of
// H is fwfft(V(r) invfft(x)) with fwfft carrying 1/N (QE convention), hence
//   dL/dV(r) = -(1/N) sum_{s,k} w_sk Re[ conj(y_sk(r)) x_sk(r) ],
//   x(r) = invfft(x(G)) without a factor.
// The forward solutions x_sk are the stored ones of the caller's forward solve.
#include <algorithm>
#include <vector>

#include "handles.h"
#include "ktimer.h"
#include "operators.h"

namespace sgw {
namespace {

constexpr int kAB = 256;

__global__ void k_adj_conj(int64_t n, const double2 *__restrict__ in, double2 *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * kAB + threadIdx.x;
  if (i < n) out[i] = cconj(in[i]);
}

__global__ void k_adj_pair_group(int nsys, int nshift, const int32_t *__restrict__ g,
                                 int32_t *__restrict__ gp) {
  const int p = blockIdx.x * kAB + threadIdx.x;
  if (p < nsys * nshift) gp[p] = g ? g[p / nshift] : 0;
}

// part[p] = w_p Re <c_s | x_p> for pair p = s * nshift + k (one block per pair)
__global__ __launch_bounds__(kAB) void k_adj_objective(int nshift, int64_t ld,
                                                       const double2 *__restrict__ c,
                                                       const double2 *__restrict__ x,
                                                       const double *__restrict__ w,
                                                       double *__restrict__ part) {
  __shared__ double red[kAB];
  const int p = blockIdx.x, s = p / nshift;
  const double2 *cs = c + (int64_t)s * ld, *xp = x + (int64_t)p * ld;
  double acc = 0;
  for (int64_t e = threadIdx.x; e < ld; e += kAB) acc += cs[e].x * xp[e].x + cs[e].y * xp[e].y;
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int off = kAB / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[p] = w[p] * red[0];
}

// grad[x + N (y + N z)] = -(1/N^3) sum_b G[b][z][x][y]
__global__ void k_adj_grad_out(int n, int nb, const double *__restrict__ G,
                               double *__restrict__ grad) {
  const int64_t e = (int64_t)blockIdx.x * kAB + threadIdx.x;
  const int64_t nbox = (int64_t)n * n * n;
  if (e >= nbox) return;
  double acc = 0;
  for (int b = 0; b < nb; ++b) acc += G[(int64_t)b * nbox + e];
  const int64_t y = e % n, x = (e / n) % n, z = e / ((int64_t)n * n);
  grad[x + n * (y + (int64_t)n * z)] = -acc / (double)nbox;
}

}  // namespace
}  // namespace sgw

namespace {
template <typename F> int adj_guarded(F &&f) {
  try {
    return f();
  } catch (const sgw::Error &e) {
    sgw::set_error(e.msg);
    return e.code;
  } catch (const std::exception &e) {
    sgw::set_error(e.what());
    return SGW_E_STATE;
  }
}
}  // namespace

extern "C" int sgw_sternheimer_adjoint(sgw_batch_solver *solver, const int32_t *d_group,
                                       const sgw_complex *d_c, const sgw_complex *d_sigma,
                                       const sgw_complex *h_sigma, const double *d_weight,
                                       const sgw_complex *d_x, sgw_complex *d_y, double *d_grad,
                                       double *h_objective, int32_t *h_ierr, void *stream) {
  using namespace sgw;
  return adj_guarded([&] {
    SGW_REQUIRE(solver && solver->op && d_c && d_sigma && h_sigma && d_weight && d_x && d_y,
                SGW_E_ARG, "sternheimer_adjoint: null argument");
    auto *pw = dynamic_cast<PwOperator *>(solver->op->op.get());
    SGW_REQUIRE(pw != nullptr || d_grad == nullptr, SGW_E_ARG,
                "sternheimer_adjoint: the V_loc gradient needs a plane-wave operator");
    const sgw_batch_config &c = solver->cfg;
    const int S = c.nsys, K = c.nshift;
    const int64_t ld = c.ld, P = (int64_t)S * K;
    hipStream_t st = (hipStream_t)stream;
    const double2 *x = reinterpret_cast<const double2 *>(d_x);
    double2 *y = reinterpret_cast<double2 *>(d_y);
    const double2 *cv = reinterpret_cast<const double2 *>(d_c);
    // backward solve at the conjugate shifts
    DevBuf<double2> sig_c(P);
    k_adj_conj<<<(unsigned)((P + kAB - 1) / kAB), kAB, 0, st>>>(
        P, reinterpret_cast<const double2 *>(d_sigma), sig_c.p);
    std::vector<double2> h_conj(P);
    for (int64_t i = 0; i < P; ++i) h_conj[i] = mk(h_sigma[i].re, -h_sigma[i].im);
    std::vector<int32_t> ierr(S), it(S);
    {
      KScope ks("adjoint_solve", st);
      solver->engine->solve(*solver->op->op, st, cv, ld, sig_c.p, h_conj.data(), d_group,
                            c.threshold, c.max_iter, y, ld, ierr.data(), it.data(), 1,
                            c.shift_test, c.relax_threshold);
    }
    int worst = 0;
    for (int s = 0; s < S; ++s) worst = std::max(worst, ierr[s]);
    if (h_ierr) std::copy(ierr.begin(), ierr.end(), h_ierr);
    // objective
    if (h_objective) {
      DevBuf<double> part(P);
      k_adj_objective<<<(unsigned)P, kAB, 0, st>>>(K, ld, cv, x, d_weight, part.p);
      std::vector<double> hp(P);
      SGW_HIP(hipMemcpyAsync(hp.data(), part.p, sizeof(double) * P, hipMemcpyDeviceToHost, st));
      SGW_HIP(hipStreamSynchronize(st));
      double L = 0;
      for (double v : hp) L += v;
      *h_objective = L;
    }
    // gradient on the box: pairs in chunks whose z-transformed columns fit in
    // about 1 GiB, partial planes G[b] accumulated over the chunks
    if (d_grad) {
      KScope ks("adjoint_grad", st);
      const int N = pw->box_n();
      const int64_t nbox = pw->box_size();
      DevBuf<int32_t> gp(P);
      k_adj_pair_group<<<(unsigned)((P + kAB - 1) / kAB), kAB, 0, st>>>(S, K, d_group, gp.p);
      const int64_t per_pair = 2 * pw->column_slots() * N * (int64_t)sizeof(double2);
      const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(P, (1ll << 30) / per_pair));
      const int nb = 64;
      DevBuf<double> G((size_t)nb * nbox);
      for (int64_t p0 = 0; p0 < P; p0 += chunk) {
        const int np = (int)std::min<int64_t>(chunk, P - p0);
        pw->grad_accumulate(st, np, x + p0 * ld, y + p0 * ld, ld, gp.p + p0, d_weight + p0, nb,
                            G.p, p0 > 0);
      }
      k_adj_grad_out<<<(unsigned)((nbox + kAB - 1) / kAB), kAB, 0, st>>>(N, nb, G.p, d_grad);
      SGW_HIP(hipGetLastError());
      SGW_HIP(hipStreamSynchronize(st));
    }
    return worst;
  });
}
