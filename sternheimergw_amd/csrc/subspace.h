// SternheimerGW subspace linear solver (select_solver priority 3) on the GPU.
#pragma once

#include "sgw_internal.h"

namespace sgw {

struct SubspaceSolver {
  // Solves (A + sigma_k) x_k = b for k = 0..nshift-1 (one system, group
  // `group` of the operator).  d_b, d_x: device; h_sigma: host.  Returns ierr
  // with the reference's semantics (0, 1 not converged, 2 NaN).
  int solve(Operator &op, hipStream_t st, int32_t group, const double2 *d_b, int64_t n,
            const double2 *h_sigma, int nshift, int max_iter, double threshold,
            double *abs_threshold_out, double2 *d_x, int64_t x_ld, int *basis_out);
};

}  // namespace sgw
