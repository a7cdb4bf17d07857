// 6-D Fourier transforms of two-point functions (fft6.f90), callable from
// other translation units of the library.
#pragma once

#include "sgw_internal.h"

namespace sgw {

struct Fft6Args {
  int n, n_p, num_g, num_g_p;
  const int32_t *nl, *nl_p;  // device box-index maps of the two G lists
  double omega;              // cell volume
  int nmat;
  double2 *f;                // nmat matrices f(nr, nr_p), Fortran layout
};

void fft6_check(const Fft6Args &a, const char *who);

// One matrix at a time on `st`; `scratch` holds num_g * n_p^3 elements.
void invfft6_run(hipStream_t st, const Fft6Args &a, double2 *scratch);
void fwfft6_run(hipStream_t st, const Fft6Args &a, double2 *scratch);
// The parts of matrix m: invfft6 = front, then box_fft(n, n_p^3, f_m, -1);
// fwfft6 = box_fft(n, n_p^3, f_m, +1), then back.
void invfft6_front(hipStream_t st, const Fft6Args &a, double2 *scratch, int m);
void fwfft6_back(hipStream_t st, const Fft6Args &a, double2 *scratch, int m);
// The r' sides alone: invfft6_front = invfft6_rows, then the zero-filled
// box scatter of the rows; fwfft6_back = the box gather, then fwfft6_rows.
void invfft6_rows(hipStream_t st, const Fft6Args &a, double2 *scratch, int m);
void fwfft6_rows(hipStream_t st, const Fft6Args &a, double2 *scratch, int m);

}  // namespace sgw
