// Optional per-kernel timing with HIP events recorded on the launch stream.
#pragma once

#include "sgw_internal.h"

namespace sgw {

// Timing level set by sgw_kernel_timing(level): 0 off, 1 one event pair per
// operator application ("happly") and per solver phase, 2 also one pair per
// kernel launch inside the operator.  A scope whose level exceeds the current
// one costs one branch.
int ktimer_level();

class KScope {
 public:
  KScope(const char *name, hipStream_t st, int level = 1);
  ~KScope();

 private:
  int slot_ = -1;
  hipStream_t st_;
};

}  // namespace sgw
