// Whole-box 3-D FFTs of cubic n^3 boxes in QE's index order
// (i1 + n*(i2 + n*i3), i1 fastest), nvec boxes back to back, in place.
#pragma once

#include "sgw_internal.h"

namespace sgw {

// Box edges with a register FFT (2^a 3^b): 16, 18, 24, 27, 32, 36, 48.
bool box_supported(int n);
const char *box_sizes_text();

// SIGN = +1: sum e^{+iGr} (QE invfft, no factor); SIGN = -1: sum e^{-iGr}
// without the 1/n^3 of QE's fwfft.  Throws for unsupported n.
void box_fft(hipStream_t st, int n, int64_t nvec, double2 *d, int sign);
// The xy pass of box_fft alone (box_fft = this, then the z pass).
void box_fft_xy(hipStream_t st, int n, int64_t nvec, double2 *d, int sign);
// d <- Z_{+1}(g * Z_{-1}(d)) along z for every line: between box_fft_xy(-1)
// and box_fft_xy(+1), the product of the inverse transform with g followed
// by the forward transform.
void box_z_product(hipStream_t st, int n, int64_t nvec, double2 *d, const double2 *g);
// nvec boxes d[r'][r] (r' < nvec) that hold S[g][r'] at the box positions of a
// list of (g, position) entries grouped by plane z (entries start[z] ..
// start[z + 1] - 1: g in gl, position x + n y in pos) and zeros elsewhere:
// box_xy_scatter writes the xy pass of such boxes without reading them;
// box_xy_gather runs the xy pass of d and keeps only S[g][r'] = scale * value
// at the list's positions (d is not written).
void box_xy_scatter(hipStream_t st, int n, int64_t nvec, const double2 *S, const int32_t *start,
                    const int32_t *gl, const int32_t *pos, double2 *d, int sign);
void box_xy_gather(hipStream_t st, int n, int64_t nvec, const double2 *d, const int32_t *start,
                   const int32_t *gl, const int32_t *pos, double scale, double2 *S, int sign);

}  // namespace sgw
