// Density response of the Sternheimer solutions, the step after the solve in
// solve_linter (phys/coul/src/solve_linter.f90:486-497 calls QE's
// incdrhoscf for every frequency): for one k+q group and every shift k
//   drho_k(r) (+)= weight * sum_v conj(psi_v(r)) dpsi_{v,k}(r),
// psi_v(r), dpsi(r) = invfft of the plane-wave coefficients (QE sign, no
// factor).  QE's incdrhoscf uses weight = 2 wk / Omega (spin, cell volume);
// the caller passes the weight.
#include <vector>

#include "handles.h"
#include "ktimer.h"
#include "operators.h"

namespace sgw {
namespace {

constexpr int kDB = 256;

// [v][z][y][x] -> [v][z][x][y]
__global__ void k_box_to_zxy(int n, int64_t nvec, const double2 *__restrict__ in,
                             double2 *__restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * kDB + threadIdx.x;
  const int64_t nbox = (int64_t)n * n * n;
  if (e >= nbox * nvec) return;
  const int64_t v = e / nbox, r = e % nbox;
  const int64_t x = r % n, y = (r / n) % n, z = r / ((int64_t)n * n);
  out[v * nbox + (z * n + x) * n + y] = in[e];
}

// out[k][z][y][x] (+)= in[k][z][x][y]
__global__ void k_box_from_zxy(int n, int64_t nvec, const double2 *__restrict__ in,
                               double2 *__restrict__ out, int accumulate) {
  const int64_t e = (int64_t)blockIdx.x * kDB + threadIdx.x;
  const int64_t nbox = (int64_t)n * n * n;
  if (e >= nbox * nvec) return;
  const int64_t v = e / nbox, r = e % nbox;
  const int64_t x = r % n, y = (r / n) % n, z = r / ((int64_t)n * n);
  const double2 val = in[v * nbox + (z * n + x) * n + y];
  out[e] = accumulate ? cadd(out[e], val) : val;
}

}  // namespace
}  // namespace sgw

extern "C" int sgw_drho_accumulate(sgw_operator *op, int32_t group, int32_t nbnd, int32_t nshift,
                                   const sgw_complex *d_dpsi, int64_t ld,
                                   const sgw_complex *d_psi_r, double weight,
                                   sgw_complex *d_drho, int32_t accumulate, void *stream) {
  using namespace sgw;
  try {
    SGW_REQUIRE(op && op->op && d_dpsi && d_psi_r && d_drho, SGW_E_ARG,
                "drho_accumulate: null argument");
    auto *pw = dynamic_cast<PwOperator *>(op->op.get());
    SGW_REQUIRE(pw != nullptr, SGW_E_ARG, "drho_accumulate: plane-wave operators only");
    SGW_REQUIRE(nbnd >= 1 && nshift >= 1 && ld >= pw->ld, SGW_E_SHAPE,
                "drho_accumulate: need nbnd, nshift >= 1 and ld >= the operator's ld");
    hipStream_t st = (hipStream_t)stream;
    const int n = pw->box_n();
    const int64_t nbox = pw->box_size();
    DevBuf<double2> pr((size_t)nbnd * nbox), dz((size_t)nshift * nbox);
    {
      KScope ks("drho", st);
      k_box_to_zxy<<<(unsigned)((nbox * nbnd + kDB - 1) / kDB), kDB, 0, st>>>(
          n, nbnd, reinterpret_cast<const double2 *>(d_psi_r), pr.p);
      pw->drho_accumulate(st, group, nbnd, nshift, reinterpret_cast<const double2 *>(d_dpsi), ld,
                          pr.p, weight, dz.p, false);
      k_box_from_zxy<<<(unsigned)((nbox * nshift + kDB - 1) / kDB), kDB, 0, st>>>(
          n, nshift, dz.p, reinterpret_cast<double2 *>(d_drho), accumulate ? 1 : 0);
    }
    SGW_HIP(hipGetLastError());
    SGW_HIP(hipStreamSynchronize(st));
    return 0;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::exception &e) {
    set_error(e.what());
    return SGW_E_STATE;
  }
}
