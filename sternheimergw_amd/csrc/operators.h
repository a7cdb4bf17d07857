// Operators (A + sigma) applied by the GPU solvers.
#pragma once

#include "sgw_internal.h"

namespace sgw {

using host_op_fn = void (*)(const double2 *sigma, const double2 *x, double2 *Ax, int64_t n,
                            void *ctx);

// A given as a dense column-major n x n matrix: Ax = A x + sigma x.
class DenseOperator : public Operator {
 public:
  DenseOperator(const double2 *h_a, int64_t n);
  void apply(hipStream_t st, int nvec, const double2 *in, int64_t in_stride, double2 *out,
             int64_t out_stride, const double2 *d_sigma, const double2 *h_sigma,
             const int32_t *d_active, const int32_t *h_active, const int32_t *d_group) override;

 private:
  DevBuf<double2> a_;
};

// Host procedure AA(sigma, x, Ax) of the reference interface.
class CallbackOperator : public Operator {
 public:
  CallbackOperator(host_op_fn fn, void *ctx, int64_t n_) : fn_(fn), ctx_(ctx) { n = n_; ld = n_; }
  void apply(hipStream_t st, int nvec, const double2 *in, int64_t in_stride, double2 *out,
             int64_t out_stride, const double2 *d_sigma, const double2 *h_sigma,
             const int32_t *d_active, const int32_t *h_active, const int32_t *d_group) override;
  bool host_side() const override { return true; }

 private:
  host_op_fn fn_;
  void *ctx_;
  std::vector<double2> hx_, hax_;
};

struct PwDesc {
  int n1, n2, n3;
  const double *vloc;      // n1*n2*n3, x fastest
  int ngroup;
  const int32_t *npw;      // [ngroup]
  int64_t ld;
  const int32_t *mill;     // [ngroup][ld][3]
  const double *kin;       // [ngroup][ld] (Ry)
  int nproj;
  const double2 *proj;     // [ngroup][nproj][ld]
  const double *dcoef;     // [ngroup][nproj] (Ry)
};

// Device memory one plane-wave application works in beyond its inputs and
// outputs: z-stick and plane buffers, the local-potential result, projector
// partial sums and overlaps, and the per-group vector lists.
struct PwWork {
  double2 *T, *S, *loc, *part, *becp;
  const int32_t *plan_vec, *plan_cnt;
  int plan_max, plan_tiles;
  int32_t *vlist;  // [nvec]: the running vectors in order, then -1
};

class PwOperator : public Operator {
 public:
  explicit PwOperator(const PwDesc &desc);
  void apply(hipStream_t st, int nvec, const double2 *in, int64_t in_stride, double2 *out,
             int64_t out_stride, const double2 *d_sigma, const double2 *h_sigma,
             const int32_t *d_active, const int32_t *h_active, const int32_t *d_group) override;
  // The same application with every work buffer carved from the library's
  // fixed per-device scratch array (no device allocation): vectors at
  // in + v ldx, results at out + v ldx.  Rejects batches whose work buffers
  // exceed the array.  Calls that overlap on one device share the array and
  // race.
  void apply_noalloc(hipStream_t st, int nvec, const int32_t *d_group, const double2 *d_sigma,
                     const double2 *in, int64_t ldx, double2 *out);
  int64_t group_n(int g) const override;
  int ngroups() const override { return ngroup_; }
  void prepare(const int32_t *h_group, int nvec) override;

  // Adjoint gradient partials: G[b][z][x][y] (+)= sum over pairs p = b, b+nb, ...
  // (p < np) of w[p] Re(conj(y_p(r)) x_p(r)); G has nb * box_size() doubles.
  void grad_accumulate(hipStream_t st, int np, const double2 *x, const double2 *y,
                       int64_t stride, const int32_t *d_group, const double *d_w, int nb,
                       double *G, bool accumulate);
  // Density response (incdrhoscf): D[k][z][x][y] (+)= wgt sum_v conj(psi_v(r))
  // dpsi_{v,k}(r) for the nbnd x nshift vectors dpsi + (v nshift + k) stride
  // of group `group`; Pr holds psi_v(r) as [v][z][x][y]; D has nshift boxes.
  void drho_accumulate(hipStream_t st, int group, int nbnd, int nshift, const double2 *dpsi,
                       int64_t stride, const double2 *Pr, double wgt, double2 *D,
                       bool accumulate);
  // Precision of the local-potential FFT pipeline: 0 FP64, 1 FP32, 2 per
  // vector, FP32 where d_relax[v] != 0 (a device array of the caller, read at
  // every apply; the cubic FFT path only, else FP64).
  void set_fft_precision(int mode, const int32_t *d_relax) {
    fft_mode_ = mode;
    relax_ = d_relax;
  }
  int fft_precision() const { return fft_mode_; }
  int box_n() const { return n1_; }
  int64_t column_slots() const { return nslot_; }
  int64_t box_size() const { return (int64_t)n1_ * n2_ * n3_; }

 private:
  int fft_mode_ = 0;
  const int32_t *relax_ = nullptr;
  // FP64 cubic path: the projector stages share launches with the z inverse
  // and plane transforms (SGW_PW_FUSE=0 in the environment keeps them apart)
  bool fuse_stages_ = true;
  // fused path: the plane and z transforms split over lane pairs
  // (SGW_PW_SPLIT=0 keeps one lane per row or column)
  bool split_planes_ = true;
  // unsplit plane transforms: two vectors per plane block where they fit
  // (SGW_PW_PAIR=0 keeps one)
  bool pair_planes_ = true;
  // split plane transforms specialised to the sphere's half-width when it is
  // one of the usual ones for the box (SGW_PW_BAND=0 keeps the general kernel)
  bool band_planes_ = true;
  void ensure_scratch(int nvec);
  void apply_work(hipStream_t st, int nvec, const double2 *in, int64_t in_stride, double2 *out,
                  int64_t out_stride, const double2 *d_sigma, const int32_t *d_active,
                  const int32_t *d_group, const PwWork &w);
  // vectors of each group: plan_vec_[g * plan_max_ + j], j < plan_cnt_[g]
  int plan_nvec_ = -1, plan_max_ = 0, plan_tiles_ = 0;
  std::vector<int32_t> plan_key_;
  DevBuf<int32_t> plan_vec_, plan_cnt_;
  DevBuf<double2> part_, becp_;
  int n1_, n2_, n3_, ngroup_, nproj_;
  int M1_, M2_, M3_ = 0, nc1_, nc2_, nslot_;
  std::vector<int32_t> npw_;
  DevBuf<double> vloc_, vloc_zxy_, kin_, dcoef_;
  DevBuf<int32_t> d_npw_, col_start_, col_cnt_, col_lo_, sidx_, sm3_, slot_of_, m3_of_;
  DevBuf<double2> proj_, tw_, S_, T_, loc_;
  DevBuf<int32_t> vlist_;
  bool columns_contiguous_ = true;
  // column layouts of the split z and plane transforms' work arrays: the
  // packed one (per group, each row of the column grid from its first to its
  // last column holding plane waves; ncs_ columns per plane) and the full
  // grid (nslot_ columns, the layout of every other kernel)
  int ncs_ = 0;
  DevBuf<int4> pk_info_, gr_info_;
  DevBuf<int32_t> pk_row_start_, pk_row_lo_, pk_row_cnt_;
  DevBuf<int32_t> gr_row_start_, gr_row_lo_, gr_row_cnt_;
};

}  // namespace sgw
