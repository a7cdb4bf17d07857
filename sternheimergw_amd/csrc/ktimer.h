// Optional per-kernel timing with HIP events recorded on the launch stream.
#pragma once

#include "sgw_internal.h"

namespace sgw {

// Enabled by sgw_kernel_timing(1); when disabled the scope costs one branch.
bool ktimer_enabled();

class KScope {
 public:
  KScope(const char *name, hipStream_t st);
  ~KScope();

 private:
  int slot_ = -1;
  hipStream_t st_;
};

}  // namespace sgw
