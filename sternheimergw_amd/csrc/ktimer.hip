// Per-kernel timing: each timed launch records a start and a stop event on
// its own stream; totals are summed at query time (after a device sync).
#include "ktimer.h"

#include <cstring>
#include <map>
#include <mutex>
#include <string>

namespace sgw {

namespace {

struct Rec {
  std::string name;
  hipEvent_t a, b;
};

std::mutex g_mu;
int g_level = 0;
std::vector<Rec> g_recs;       // events of the current measurement
std::vector<hipEvent_t> g_pool;  // recycled events
std::map<std::string, std::pair<double, long long>> g_done;  // resolved totals

hipEvent_t take_event() {
  if (!g_pool.empty()) {
    hipEvent_t e = g_pool.back();
    g_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  SGW_HIP(hipEventCreate(&e));
  return e;
}

// Sum elapsed times of recorded pairs into g_done and recycle the events.
void resolve_locked() {
  if (g_recs.empty()) return;
  SGW_HIP(hipDeviceSynchronize());
  for (auto &r : g_recs) {
    float ms = 0.f;
    SGW_HIP(hipEventElapsedTime(&ms, r.a, r.b));
    auto &d = g_done[r.name];
    d.first += ms;
    d.second += 1;
    g_pool.push_back(r.a);
    g_pool.push_back(r.b);
  }
  g_recs.clear();
}

}  // namespace

int ktimer_level() { return g_level; }

KScope::KScope(const char *name, hipStream_t st, int level) : st_(st) {
  if (level > g_level) return;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_recs.size() > 20000) resolve_locked();  // bound the number of live events
  Rec r{name, take_event(), take_event()};
  SGW_HIP(hipEventRecord(r.a, st));
  g_recs.push_back(r);
  slot_ = (int)g_recs.size() - 1;
}

KScope::~KScope() {
  if (slot_ < 0) return;
  std::lock_guard<std::mutex> lk(g_mu);
  if (slot_ < (int)g_recs.size()) (void)hipEventRecord(g_recs[slot_].b, st_);
}

}  // namespace sgw

extern "C" {

// Set the timing level (0 off, 1 operator applications and solver phases,
// 2 every kernel launch); any call resets the accumulated totals.
int sgw_kernel_timing(int enable) {
  try {
    std::lock_guard<std::mutex> lk(sgw::g_mu);
    sgw::resolve_locked();
    sgw::g_done.clear();
    sgw::g_level = enable < 0 ? 0 : (enable > 2 ? 2 : enable);
    return 0;
  } catch (const sgw::Error &e) {
    sgw::set_error(e.msg);
    return e.code;
  }
}

// Total milliseconds and number of timed launches of kernel `name`.
int sgw_kernel_timing_query(const char *name, double *total_ms, int64_t *launches) {
  try {
    std::lock_guard<std::mutex> lk(sgw::g_mu);
    sgw::resolve_locked();
    auto it = sgw::g_done.find(name ? name : "");
    if (total_ms) *total_ms = it == sgw::g_done.end() ? 0.0 : it->second.first;
    if (launches) *launches = it == sgw::g_done.end() ? 0 : it->second.second;
    return 0;
  } catch (const sgw::Error &e) {
    sgw::set_error(e.msg);
    return e.code;
  }
}

}  // extern "C"
