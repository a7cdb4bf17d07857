// Launch-shape autotuner.  For a key (a kernel and a size class of its
// problem), the first call times every candidate launch once on the caller's
// stream (HIP events, then a wait) and keeps the fastest; later calls with the
// key run only that one.  Candidates must compute the same results from the
// same inputs and must not read what they write, because the first call runs
// all of them.  SGW_AUTOTUNE=0 in the environment keeps the first candidate
// (the default shape), SGW_AUTOTUNE_FORCE=i runs candidate i (tests of each
// shape); a stream under graph capture uses the choice made so far, or the
// default, and never times.
#pragma once

#include "sgw_internal.h"

#include <functional>
#include <string>

namespace sgw {

int autotune(const std::string &key, int ncand, hipStream_t st,
             const std::function<void(int)> &launch);

// the candidate chosen for `key` so far, -1 if none (for tests and reports)
int autotune_choice(const std::string &key);

// next power of two >= n (size classes of the keys)
inline int size_class(int n) {
  int c = 1;
  while (c < n) c <<= 1;
  return c;
}

}  // namespace sgw
