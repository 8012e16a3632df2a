// Per-stage HIP-event timing registry (see timing.hpp).
#include <hip/hip_runtime.h>
#include <mutex>
#include <vector>
#include "timing.hpp"
#include "common.hpp"
#include "../../include/smx.h"

namespace smx {

namespace {
struct Rec {
  hipEvent_t a, b;
  int stage;
  long launches;
};
std::mutex mu;
bool on = false;
std::vector<Rec> recs;
hipEvent_t open_ev[ST_COUNT];
bool open_set[ST_COUNT] = {};
}  // namespace

bool timing_on() { return on; }

void stage_begin(Stage s, hipStream_t st) {
  if (!on) return;
  std::lock_guard<std::mutex> lk(mu);
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return;
  (void)hipEventRecord(e, st);
  open_ev[s] = e;
  open_set[s] = true;
}

void stage_end(Stage s, hipStream_t st, long launches) {
  if (!on) return;
  std::lock_guard<std::mutex> lk(mu);
  if (!open_set[s]) return;
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return;
  (void)hipEventRecord(e, st);
  recs.push_back({open_ev[s], e, (int)s, launches});
  open_set[s] = false;
}

}  // namespace smx

using namespace smx;

extern "C" {

int smx_timing_enable(int enable) {
  std::lock_guard<std::mutex> lk(mu);
  on = enable != 0;
  for (auto& r : recs) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  recs.clear();
  return 0;
}

int smx_timing_read(double* ms, int64_t* launches, int32_t max_stages) {
  std::lock_guard<std::mutex> lk(mu);
  for (int i = 0; i < max_stages; ++i) {
    ms[i] = 0.0;
    launches[i] = 0;
  }
  for (auto& r : recs) {
    SMX_HIP(hipEventSynchronize(r.b));
    float t = 0.f;
    SMX_HIP(hipEventElapsedTime(&t, r.a, r.b));
    if (r.stage < max_stages) {
      ms[r.stage] += t;
      launches[r.stage] += r.launches;
    }
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  recs.clear();
  return 0;
}

}  // extern "C"
