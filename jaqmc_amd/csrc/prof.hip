// HIP-event timing of launch classes (benchmark instrumentation).  When
// enabled, each instrumented launch is bracketed by two events recorded on its
// own stream; jq_profile_read sums the elapsed times.
#include <vector>

#include "../../include/jaqmc_amd_prof.h"
#include "prof.h"

namespace {
struct Pool {
  std::vector<hipEvent_t> ev;
  int used = 0;
  int open = 0;
};
bool g_on = false;
Pool g_pool[PROF_NCLASS];

hipEvent_t get(Pool& p, int i) {
  while ((int)p.ev.size() <= i) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    p.ev.push_back(e);
  }
  return p.ev[i];
}
}  // namespace

void prof_begin(int cls, hipStream_t s) {
  if (!g_on || cls < 0 || cls >= PROF_NCLASS) return;
  Pool& p = g_pool[cls];
  hipEvent_t e = get(p, 2 * p.used);
  if (e) {
    (void)hipEventRecord(e, s);
    p.open = 1;
  }
}

void prof_end(int cls, hipStream_t s) {
  if (!g_on || cls < 0 || cls >= PROF_NCLASS) return;
  Pool& p = g_pool[cls];
  if (!p.open) return;
  hipEvent_t e = get(p, 2 * p.used + 1);
  if (e) {
    (void)hipEventRecord(e, s);
    p.used++;
  }
  p.open = 0;
}

extern "C" int jq_profile_enable(int on) {
  g_on = on != 0;
  if (g_on)
    for (auto& p : g_pool) p.used = 0, p.open = 0;
  return 0;
}

extern "C" int jq_profile_read(int cls, double* total_ms, int* count) {
  if (cls < 0 || cls >= PROF_NCLASS || !total_ms || !count) return -1;
  Pool& p = g_pool[cls];
  double t = 0.0;
  for (int i = 0; i < p.used; ++i) {
    float ms = 0.f;
    if (hipEventSynchronize(p.ev[2 * i + 1]) != hipSuccess) return -2;
    if (hipEventElapsedTime(&ms, p.ev[2 * i], p.ev[2 * i + 1]) != hipSuccess) return -2;
    t += ms;
  }
  *total_ms = t;
  *count = p.used;
  return 0;
}
