// Batched multishift BiCGStab(l) (Frommer 2003) on the GPU.
#pragma once

#include "sgw_internal.h"

namespace sgw {

constexpr int kMaxL = 15;  // the reference's tests exercise l = 1..15

class BatchBicgstab {
 public:
  // nsys systems, each with nshift shifts (the first is the seed), vectors of
  // padded length ld.
  BatchBicgstab(int nsys, int nshift, int lmax, int64_t ld);
  ~BatchBicgstab();
  BatchBicgstab(const BatchBicgstab &) = delete;
  BatchBicgstab &operator=(const BatchBicgstab &) = delete;

  // d_rhs: [nsys][rhs_ld]; d_sigma/h_sigma: [nsys][nshift]; d_group: [nsys]
  // (operator group of each system, may be null for single-group operators);
  // d_x: [nsys][nshift][x_ld].  h_ierr/h_iters: per system, host.
  void solve(Operator &op, hipStream_t st, const double2 *d_rhs, int64_t rhs_ld,
             const double2 *d_sigma, const double2 *h_sigma, const int32_t *d_group,
             double threshold, int max_iter, double2 *d_x, int64_t x_ld, int32_t *h_ierr,
             int32_t *h_iters, int check_every = 1, int shift_test = 0,
             double relax_thr = 0.0);

  // Streaming solve of a queue of nq systems through the nsys slots: d_rhs
  // [nq][rhs_ld], d_sigma/h_sigma [nq][nshift], h_group [nq] (operator group of
  // each queued system, host), d_x [nq][nshift][x_ld].  Each group owns a fixed
  // share of the slots; a slot whose system has finished takes the next queued
  // system of its group at the next outer iteration, so the batch stays full
  // until the queue runs dry.  Every system runs exactly the iterations of the
  // lock-step solve (the iterates do not depend on the other slots).
  // Device-resident operators only; no relaxed precision.
  void solve_stream(Operator &op, hipStream_t st, int nq, const double2 *d_rhs, int64_t rhs_ld,
                    const double2 *d_sigma, const double2 *h_sigma, const int32_t *h_group,
                    double threshold, int max_iter, double2 *d_x, int64_t x_ld,
                    int32_t *h_ierr, int32_t *h_iters, int shift_test,
                    double relax_thr = 0.0);

 private:
  struct Impl;
  Impl *impl_;
};

}  // namespace sgw
