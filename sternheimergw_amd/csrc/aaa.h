// AAA continuation of W (vendor/analytic/src/aaa.f90 through
// algo/analytic/src/analytic.f90:162-199): coefficients of every element.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sgw {

// d_w[f * ne + e] holds W of element e at mesh point z[f] taken from slice
// src[f] (NULL: f); it receives model 4's [position, value, weight] thirds
// (N / 3 each) or model 5's [pole, residue] halves.  Synchronises; throws
// when an element needs more than N / 2 support points (model 5) or has
// more relevant poles than N / 2.
void aaa_coeffs(int model, double thres, int64_t ne, int nfreq, const double2 *z,
                const int32_t *src, double2 *d_w, hipStream_t st);

}  // namespace sgw
