// C ABI of libsgw_hip.so (declared in include/sgw.h).
#include "sgw.h"

#include <cstring>
#include <memory>
#include <mutex>
#include <string>

#include "bicgstab.h"
#include "handles.h"
#include "operators.h"
#include "subspace.h"

namespace sgw {

static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }
const char *last_error() { return g_last_error.c_str(); }

template <typename F> static int guarded(F &&f) {
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

template <typename F> static auto guarded_ptr(F &&f) -> decltype(f()) {
  try {
    return f();
  } catch (const Error &e) {
    set_error(e.msg);
    return nullptr;
  } catch (const std::exception &e) {
    set_error(e.what());
    return nullptr;
  }
}

static const double2 *D2(const sgw_complex *p) { return reinterpret_cast<const double2 *>(p); }
static double2 *D2(sgw_complex *p) { return reinterpret_cast<double2 *>(p); }

namespace {

constexpr int kB = 256;

// res[v] = || rhs[s] - ax[v] || / || rhs[s] ||, v = s*nshift + k
__global__ __launch_bounds__(kB) void k_resid(int nshift, int64_t ld,
                                              const double2 *__restrict__ rhs,
                                              const double2 *__restrict__ ax,
                                              double *__restrict__ res) {
  __shared__ double2 lds[kB / 64];
  const int v = blockIdx.x;
  const int s = v / nshift;
  const double2 *b = rhs + (int64_t)s * ld, *a = ax + (int64_t)v * ld;
  double2 acc = mk(0, 0);
  for (int64_t i = threadIdx.x; i < ld; i += kB) {
    double2 d = csub(b[i], a[i]);
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
    for (int w = 0; w < kB / 64; ++w) { t.x += lds[w].x; t.y += lds[w].y; }
    res[v] = sqrt(t.x) / sqrt(t.y);
  }
}

__global__ void k_expand_group(int nsys, int nshift, const int32_t *__restrict__ g,
                               int32_t *__restrict__ gv) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nsys * nshift) return;
  gv[v] = g ? g[v / nshift] : 0;
}

}  // namespace

// Run the multishift BiCGStab on one system with host arrays.
static int bicg_single(Operator &op, int32_t group, const sgw_bicgstab_config &cfg,
                       const sgw_complex *bb, int64_t n, const sgw_complex *sigma,
                       int32_t nshift, sgw_complex *xx, int32_t *iters) {
  SGW_REQUIRE(bb && sigma && xx, SGW_E_ARG, "bicgstab: null array");
  SGW_REQUIRE(n >= 1 && nshift >= 1, SGW_E_SHAPE, "bicgstab: empty problem");
  SGW_REQUIRE(op.ld == n || op.group_n(group) <= n, SGW_E_SHAPE,
              "bicgstab: right-hand side and operator sizes differ");
  const int64_t ld = op.ld;
  SGW_REQUIRE(n <= ld, SGW_E_SHAPE, "bicgstab: vector longer than the operator");
  hipStream_t st = nullptr;
  DevBuf<double2> d_b(ld), d_sig(nshift), d_x((size_t)nshift * ld);
  DevBuf<int32_t> d_g(1);
  d_b.zero(st);
  d_b.upload(D2(bb), n, st);
  d_sig.upload(D2(sigma), nshift, st);
  d_g.upload(&group, 1, st);
  BatchBicgstab eng(1, nshift, cfg.lmax, ld);
  int32_t ierr = 0, it = 0;
  eng.solve(op, st, d_b.p, ld, d_sig.p, D2(sigma), d_g.p, cfg.threshold, cfg.max_iter, d_x.p, ld,
            &ierr, &it, 1);
  if (ierr == 1 && it > cfg.max_iter)
    std::printf("WARNING: BiCGstab algorithm did not converge in %d iterations.\n", cfg.max_iter);
  if (ierr == 2) std::printf(" WARNING: BiCGstab algorithm resulted in one or more NaN\n");
  SGW_HIP(hipMemcpy2D(xx, sizeof(sgw_complex) * n, d_x.p, sizeof(double2) * ld,
                      sizeof(double2) * n, nshift, hipMemcpyDeviceToHost));
  if (iters) *iters = it;
  return ierr;
}

static int subspace_single(Operator &op, int32_t group, sgw_linear_solver_config &cfg,
                           const sgw_complex *bb, int64_t n, const sgw_complex *sigma,
                           int32_t nshift, sgw_complex *xx) {
  SGW_REQUIRE(bb && sigma && xx, SGW_E_ARG, "linear_solver: null array");
  SGW_REQUIRE(n >= 1 && nshift >= 1, SGW_E_SHAPE, "linear_solver: empty problem");
  SGW_REQUIRE(cfg.recover_subspace == 1, SGW_E_ARG,
              "linear_solver: unknown option for recover subspace");
  const int64_t ld = op.ld;
  SGW_REQUIRE(n <= ld, SGW_E_SHAPE, "linear_solver: vector longer than the operator");
  hipStream_t st = nullptr;
  DevBuf<double2> d_b(ld), d_x((size_t)nshift * ld);
  d_b.zero(st);
  d_b.upload(D2(bb), n, st);
  d_x.zero(st);
  // the solver works on the operator's padded length; padding stays zero
  SubspaceSolver ss;
  int ierr = ss.solve(op, st, group, d_b.p, ld, D2(sigma), nshift, cfg.max_iter, cfg.threshold,
                      &cfg.abs_threshold, d_x.p, ld, nullptr);
  // columns not reached (ierr /= 0 exits early) keep the values xx had on entry
  // in the reference (INTENT(OUT), undefined); they are returned as zero here.
  SGW_HIP(hipMemcpy2D(xx, sizeof(sgw_complex) * n, d_x.p, sizeof(double2) * ld,
                      sizeof(double2) * n, nshift, hipMemcpyDeviceToHost));
  return ierr;
}

static int select_single(Operator &op, int32_t group, const sgw_select_solver_config &cfg,
                         const sgw_complex *bb, int64_t n, const sgw_complex *sigma,
                         int32_t nshift, sgw_complex *xx) {
  SGW_REQUIRE(cfg.priority != nullptr && cfg.npriority > 0, SGW_E_ARG,
              "select_solver: priority of the solvers not specified");
  int ierr = 1;
  for (int is = 0; is < cfg.npriority; ++is) {
    if (is > 0) std::printf("First choice of solver did not converge, try a different one\n");
    switch (cfg.priority[is]) {
      case SGW_BICGSTAB_MULTI:
      case SGW_BICGSTAB_NO_MULTI: {
        // priority 2 repeats the identical multishift solve SIZE(sigma) times
        // in the reference (select_solver.f90:144-146); one solve gives the
        // same x and ierr.
        sgw_bicgstab_config bc{cfg.bicg_lmax, cfg.threshold, cfg.max_iter};
        ierr = bicg_single(op, group, bc, bb, n, sigma, nshift, xx, nullptr);
        break;
      }
      case SGW_LINEAR_SOLVER: {
        sgw_linear_solver_config lc{1, cfg.max_iter, cfg.threshold, 0.0};
        ierr = subspace_single(op, group, lc, bb, n, sigma, nshift, xx);
        break;
      }
      default:
        break;
    }
    if (ierr == 0) break;
  }
  return ierr;
}

}  // namespace sgw

using namespace sgw;

extern "C" {

const char *sgw_last_error(void) { return sgw::last_error(); }
const char *sgw_version(void) { return "sgw-mi355x 0.1 (gfx950)"; }

int sgw_select_solver(const sgw_select_solver_config *config, sgw_op_fn AA, void *ctx,
                      const sgw_complex *bb, int64_t n, const sgw_complex *sigma,
                      int32_t nshift, sgw_complex *xx) {
  return guarded([&] {
    SGW_REQUIRE(config && AA, SGW_E_ARG, "select_solver: null config or operator");
    CallbackOperator op(reinterpret_cast<host_op_fn>(AA), ctx, n);
    return select_single(op, 0, *config, bb, n, sigma, nshift, xx);
  });
}

int sgw_bicgstab(const sgw_bicgstab_config *config, sgw_op_fn AA, void *ctx,
                 const sgw_complex *bb, int64_t n, const sgw_complex *sigma, int32_t nshift,
                 sgw_complex *xx, int32_t *iters) {
  return guarded([&] {
    SGW_REQUIRE(config && AA, SGW_E_ARG, "bicgstab: null config or operator");
    CallbackOperator op(reinterpret_cast<host_op_fn>(AA), ctx, n);
    return bicg_single(op, 0, *config, bb, n, sigma, nshift, xx, iters);
  });
}

int sgw_linear_solver(sgw_linear_solver_config *config, sgw_op_fn AA, void *ctx,
                      const sgw_complex *bb, int64_t n, const sgw_complex *sigma,
                      int32_t nshift, sgw_complex *xx) {
  return guarded([&] {
    SGW_REQUIRE(config && AA, SGW_E_ARG, "linear_solver: null config or operator");
    CallbackOperator op(reinterpret_cast<host_op_fn>(AA), ctx, n);
    return subspace_single(op, 0, *config, bb, n, sigma, nshift, xx);
  });
}

sgw_operator *sgw_operator_dense(const sgw_complex *a, int64_t n) {
  return guarded_ptr([&]() -> sgw_operator * {
    SGW_REQUIRE(a && n >= 1, SGW_E_ARG, "operator_dense: bad arguments");
    auto *h = new sgw_operator;
    h->op.reset(new DenseOperator(D2(a), n));
    return h;
  });
}

sgw_operator *sgw_operator_callback(sgw_op_fn AA, void *ctx, int64_t n) {
  return guarded_ptr([&]() -> sgw_operator * {
    SGW_REQUIRE(AA && n >= 1, SGW_E_ARG, "operator_callback: bad arguments");
    auto *h = new sgw_operator;
    h->op.reset(new CallbackOperator(reinterpret_cast<host_op_fn>(AA), ctx, n));
    return h;
  });
}

sgw_operator *sgw_operator_pw(const sgw_pw_desc *desc) {
  return guarded_ptr([&]() -> sgw_operator * {
    SGW_REQUIRE(desc && desc->vloc && desc->npw && desc->mill && desc->kin, SGW_E_ARG,
                "operator_pw: null descriptor field");
    SGW_REQUIRE(desc->nproj == 0 || (desc->proj && desc->dcoef), SGW_E_ARG,
                "operator_pw: projectors missing");
    PwDesc d{desc->n1, desc->n2, desc->n3, desc->vloc, desc->ngroup, desc->npw, desc->ld,
             desc->mill, desc->kin, desc->nproj, D2(desc->proj), desc->dcoef};
    auto *h = new sgw_operator;
    h->op.reset(new PwOperator(d));
    return h;
  });
}

void sgw_operator_destroy(sgw_operator *op) { delete op; }

int sgw_operator_set_fft_precision(sgw_operator *op, int32_t mode) {
  return guarded([&] {
    SGW_REQUIRE(op && op->op, SGW_E_ARG, "operator_set_fft_precision: null operator");
    SGW_REQUIRE(mode == 0 || mode == 1, SGW_E_ARG, "operator_set_fft_precision: mode 0 or 1");
    auto *pw = dynamic_cast<PwOperator *>(op->op.get());
    SGW_REQUIRE(pw != nullptr, SGW_E_ARG, "operator_set_fft_precision: plane-wave operators only");
    pw->set_fft_precision(mode, nullptr);
    return 0;
  });
}

int sgw_operator_apply_device(sgw_operator *op, int32_t nvec, const int32_t *d_group,
                              const sgw_complex *d_sigma, const sgw_complex *d_x, int64_t ldx,
                              sgw_complex *d_ax, void *stream) {
  return guarded([&] {
    SGW_REQUIRE(op && op->op, SGW_E_ARG, "operator_apply: null operator");
    SGW_REQUIRE(!op->op->host_side(), SGW_E_ARG, "operator_apply_device: host operator");
    SGW_REQUIRE(ldx >= op->op->ld, SGW_E_SHAPE, "operator_apply: ldx < operator ld");
    std::vector<int32_t> hg(nvec, 0);
    if (d_group)
      SGW_HIP(hipMemcpy(hg.data(), d_group, sizeof(int32_t) * nvec, hipMemcpyDeviceToHost));
    op->op->prepare(hg.data(), nvec);
    op->op->apply((hipStream_t)stream, nvec, D2(d_x), ldx, D2(d_ax), ldx, D2(d_sigma), nullptr,
                  nullptr, nullptr, d_group);
    return 0;
  });
}

int sgw_operator_apply_device_noalloc(sgw_operator *op, int32_t nvec, const int32_t *d_group,
                                      const sgw_complex *d_sigma, const sgw_complex *d_x,
                                      int64_t ldx, sgw_complex *d_ax, void *stream) {
  return guarded([&] {
    SGW_REQUIRE(op && op->op, SGW_E_ARG, "operator_apply_noalloc: null operator");
    auto *pw = dynamic_cast<PwOperator *>(op->op.get());
    SGW_REQUIRE(pw != nullptr, SGW_E_ARG,
                "operator_apply_noalloc: plane-wave operators only");
    pw->apply_noalloc((hipStream_t)stream, nvec, d_group, D2(d_sigma), D2(d_x), ldx, D2(d_ax));
    return 0;
  });
}

int sgw_operator_apply(sgw_operator *op, int32_t nvec, const int32_t *group,
                       const sgw_complex *sigma, const sgw_complex *x, int64_t ldx,
                       sgw_complex *ax) {
  return guarded([&] {
    SGW_REQUIRE(op && op->op && x && ax && sigma, SGW_E_ARG, "operator_apply: null argument");
    SGW_REQUIRE(nvec >= 1, SGW_E_ARG, "operator_apply: nvec < 1");
    Operator &o = *op->op;
    SGW_REQUIRE(ldx >= o.ld, SGW_E_SHAPE, "operator_apply: ldx < operator ld");
    if (group)
      for (int v = 0; v < nvec; ++v)
        SGW_REQUIRE(group[v] >= 0 && group[v] < o.ngroups(), SGW_E_ARG,
                    "operator_apply: group index out of range");
    hipStream_t st = nullptr;
    DevBuf<double2> dx((size_t)nvec * ldx), dax((size_t)nvec * ldx), ds(nvec);
    DevBuf<int32_t> dg(nvec), da(nvec);
    std::vector<int32_t> ones(nvec, 1);
    dx.upload(D2(x), (size_t)nvec * ldx, st);
    dax.zero(st);
    ds.upload(D2(sigma), nvec, st);
    da.upload(ones.data(), nvec, st);
    std::vector<int32_t> hg(nvec, 0);
    if (group) std::copy(group, group + nvec, hg.begin());
    dg.upload(hg.data(), nvec, st);
    o.prepare(hg.data(), nvec);
    o.apply(st, nvec, dx.p, ldx, dax.p, ldx, ds.p, D2(sigma), da.p, ones.data(), dg.p);
    SGW_HIP(hipMemcpyAsync(ax, dax.p, sizeof(double2) * nvec * ldx, hipMemcpyDeviceToHost, st));
    SGW_HIP(hipStreamSynchronize(st));
    return 0;
  });
}

int sgw_select_solver_op(const sgw_select_solver_config *config, sgw_operator *op,
                         int32_t group, const sgw_complex *bb, int64_t n,
                         const sgw_complex *sigma, int32_t nshift, sgw_complex *xx) {
  return guarded([&] {
    SGW_REQUIRE(config && op && op->op, SGW_E_ARG, "select_solver_op: null argument");
    SGW_REQUIRE(group >= 0 && group < op->op->ngroups(), SGW_E_ARG,
                "select_solver_op: group index out of range");
    SGW_REQUIRE(n == op->op->group_n(group), SGW_E_SHAPE,
                "select_solver_op: right-hand side size differs from the operator's");
    return select_single(*op->op, group, *config, bb, n, sigma, nshift, xx);
  });
}

sgw_batch_solver *sgw_batch_solver_create(sgw_operator *op, const sgw_batch_config *cfg) {
  return guarded_ptr([&]() -> sgw_batch_solver * {
    SGW_REQUIRE(op && op->op && cfg, SGW_E_ARG, "batch_solver_create: null argument");
    SGW_REQUIRE(cfg->nsys >= 1 && cfg->nshift >= 1, SGW_E_ARG, "batch_solver_create: empty batch");
    SGW_REQUIRE(cfg->ld >= op->op->ld, SGW_E_SHAPE, "batch_solver_create: ld < operator ld");
    auto *s = new sgw_batch_solver;
    s->op = op;
    s->cfg = *cfg;
    s->engine.reset(new BatchBicgstab(cfg->nsys, cfg->nshift, cfg->lmax, op->op->ld));
    return s;
  });
}

void sgw_batch_solver_destroy(sgw_batch_solver *solver) { delete solver; }

int sgw_batch_solver_run(sgw_batch_solver *solver, const int32_t *d_group,
                         const sgw_complex *d_rhs, const sgw_complex *d_sigma,
                         const sgw_complex *h_sigma, sgw_complex *d_x, int32_t *h_ierr,
                         int32_t *h_iters, void *stream) {
  return guarded([&] {
    SGW_REQUIRE(solver && d_rhs && d_sigma && h_sigma && d_x, SGW_E_ARG,
                "batch_solver_run: null argument");
    const sgw_batch_config &c = solver->cfg;
    std::vector<int32_t> ierr(c.nsys), it(c.nsys);
    solver->engine->solve(*solver->op->op, (hipStream_t)stream, D2(d_rhs), c.ld, D2(d_sigma),
                          D2(h_sigma), d_group, c.threshold, c.max_iter, D2(d_x), c.ld,
                          ierr.data(), it.data(), 1, c.shift_test, c.relax_threshold);
    int worst = 0;
    for (int s = 0; s < c.nsys; ++s) worst = std::max(worst, ierr[s]);
    if (h_ierr) std::memcpy(h_ierr, ierr.data(), sizeof(int32_t) * c.nsys);
    if (h_iters) std::memcpy(h_iters, it.data(), sizeof(int32_t) * c.nsys);
    return worst;
  });
}

int sgw_batch_solver_run_stream(sgw_batch_solver *solver, int32_t nq, const int32_t *h_group,
                                const sgw_complex *d_rhs, const sgw_complex *d_sigma,
                                const sgw_complex *h_sigma, sgw_complex *d_x, int32_t *h_ierr,
                                int32_t *h_iters, void *stream) {
  return guarded([&] {
    SGW_REQUIRE(solver && h_group && d_rhs && d_sigma && h_sigma && d_x, SGW_E_ARG,
                "batch_solver_run_stream: null argument");
    SGW_REQUIRE(nq >= 0, SGW_E_ARG, "batch_solver_run_stream: negative queue length");
    const sgw_batch_config &c = solver->cfg;
    std::vector<int32_t> ierr(std::max(nq, 1)), it(std::max(nq, 1));
    solver->engine->solve_stream(*solver->op->op, (hipStream_t)stream, nq, D2(d_rhs), c.ld,
                                 D2(d_sigma), D2(h_sigma), h_group, c.threshold, c.max_iter,
                                 D2(d_x), c.ld, ierr.data(), it.data(), c.shift_test,
                                 c.relax_threshold);
    int worst = 0;
    for (int q = 0; q < nq; ++q) worst = std::max(worst, ierr[q]);
    if (h_ierr && nq > 0) std::memcpy(h_ierr, ierr.data(), sizeof(int32_t) * nq);
    if (h_iters && nq > 0) std::memcpy(h_iters, it.data(), sizeof(int32_t) * nq);
    return worst;
  });
}

int sgw_sternheimer_solve_batched(sgw_operator *op, const sgw_batch_config *cfg,
                                  const int32_t *group, const sgw_complex *rhs,
                                  const sgw_complex *sigma, sgw_complex *x, int32_t *ierr,
                                  int32_t *iters) {
  return guarded([&] {
    SGW_REQUIRE(op && op->op && cfg && rhs && sigma && x, SGW_E_ARG,
                "sternheimer_solve_batched: null argument");
    SGW_REQUIRE(cfg->ld >= op->op->ld, SGW_E_SHAPE, "sternheimer_solve_batched: ld < operator ld");
    const int S = cfg->nsys, K = cfg->nshift;
    const int64_t ld = cfg->ld;
    if (group)
      for (int s = 0; s < S; ++s)
        SGW_REQUIRE(group[s] >= 0 && group[s] < op->op->ngroups(), SGW_E_ARG,
                    "sternheimer_solve_batched: group index out of range");
    hipStream_t st = nullptr;
    DevBuf<double2> d_rhs((size_t)S * ld), d_sig((size_t)S * K), d_x((size_t)S * K * ld);
    DevBuf<int32_t> d_g(S);
    d_rhs.upload(D2(rhs), (size_t)S * ld, st);
    d_sig.upload(D2(sigma), (size_t)S * K, st);
    if (group) d_g.upload(group, S, st);
    else d_g.zero(st);
    d_x.zero(st);
    BatchBicgstab eng(S, K, cfg->lmax, ld);
    std::vector<int32_t> e(S), it(S);
    eng.solve(*op->op, st, d_rhs.p, ld, d_sig.p, D2(sigma), d_g.p, cfg->threshold, cfg->max_iter,
              d_x.p, ld, e.data(), it.data(), 1, cfg->shift_test, cfg->relax_threshold);
    SGW_HIP(hipMemcpy(x, d_x.p, sizeof(double2) * S * K * ld, hipMemcpyDeviceToHost));
    int worst = 0;
    for (int s = 0; s < S; ++s) worst = std::max(worst, e[s]);
    if (ierr) std::memcpy(ierr, e.data(), sizeof(int32_t) * S);
    if (iters) std::memcpy(iters, it.data(), sizeof(int32_t) * S);
    return worst;
  });
}

int sgw_residual_norms_device(sgw_operator *op, int32_t nsys, int32_t nshift, int64_t ld,
                              const int32_t *d_group, const sgw_complex *d_rhs,
                              const sgw_complex *d_sigma, const sgw_complex *d_x,
                              double *d_res, void *stream) {
  return guarded([&] {
    SGW_REQUIRE(op && op->op && !op->op->host_side(), SGW_E_ARG,
                "residual_norms: need a device operator");
    SGW_REQUIRE(ld >= op->op->ld, SGW_E_SHAPE, "residual_norms: ld < operator ld");
    hipStream_t st = (hipStream_t)stream;
    const int nv = nsys * nshift;
    DevBuf<double2> ax((size_t)nv * ld);
    DevBuf<int32_t> gv(nv), act(nv);
    k_expand_group<<<(nv + 255) / 256, 256, 0, st>>>(nsys, nshift, d_group, gv.p);
    std::vector<int32_t> hsys(nsys, 0), hv(nv);
    if (d_group)
      SGW_HIP(hipMemcpyAsync(hsys.data(), d_group, sizeof(int32_t) * nsys, hipMemcpyDeviceToHost,
                             st));
    SGW_HIP(hipStreamSynchronize(st));
    for (int v = 0; v < nv; ++v) hv[v] = hsys[v / nshift];
    op->op->prepare(hv.data(), nv);
    std::vector<int32_t> ones(nv, 1);
    act.upload(ones.data(), nv, st);
    op->op->apply(st, nv, D2(d_x), ld, ax.p, ld, D2(d_sigma), nullptr, act.p, nullptr, gv.p);
    k_resid<<<nv, kB, 0, st>>>(nshift, ld, D2(d_rhs), ax.p, d_res);
    SGW_HIP(hipGetLastError());
    SGW_HIP(hipStreamSynchronize(st));
    return 0;
  });
}

}  // extern "C"
