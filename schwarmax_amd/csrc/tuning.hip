// Registry of the launch-shape knobs (tuning.hpp) and their C ABI.
#include <atomic>
#include "common.hpp"
#include "tuning.hpp"
#include "../../include/smx.h"

namespace smx {

namespace {
// Defaults, as measured at the plain 20,000-orbit and the 40,000-orbit NNLS
// shapes: occupancy 2, 8-wave Gram (25.8 against 29.8 ms for 4 waves at
// plain-20k shapes), syrkx from 16,384 rows, the thin FP64 GEMM, and a
// split-K target of 2,048 workgroups for the tiled FP64 GEMM (the right-hand
// sides and one anchor: 5.9 ms against 6.5 ms at 1,024 at plain-20k shapes).
constexpr int64_t kDefault[KN_COUNT] = {2, 8, 16384, 2048, 0, 1, 1};
std::atomic<int64_t> g_knob[KN_COUNT] = {{kDefault[0]}, {kDefault[1]}, {kDefault[2]},
                                         {kDefault[3]}, {kDefault[4]}, {kDefault[5]},
                                         {kDefault[6]}};

bool valid(int k, int64_t v) {
  switch (k) {
    case KN_ADMM_OCC: return v >= 1 && v <= 8;
    case KN_GRAM_WAVES: return v == 4 || v == 8;
    case KN_INV_SYRKX_MIN: return v >= 256;
    case KN_GEMM64_SPLIT: return v >= 256 && v <= 4096;
    case KN_ADMM_LOADS: return v >= 0 && v <= 2;
    case KN_GEMM64_KERNEL: return v == 1 || v == 2;
    case KN_LIB_BINNING: return v == 1 || v == 2;
    default: return false;
  }
}
}  // namespace

int64_t knob(Knob k) { return g_knob[k].load(std::memory_order_relaxed); }

}  // namespace smx

using namespace smx;

extern "C" {

int smx_tune_set(int32_t k, int64_t value) {
  SMX_REQUIRE(k >= 0 && k < KN_COUNT, "smx_tune_set: unknown knob %d", (int)k);
  const int64_t v = value == 0 ? kDefault[k] : value;
  SMX_REQUIRE(valid(k, v), "smx_tune_set: value %lld out of range for knob %d", (long long)value,
              (int)k);
  g_knob[k].store(v, std::memory_order_relaxed);
  return 0;
}

int smx_tune_get(int32_t k, int64_t* value) {
  SMX_REQUIRE(k >= 0 && k < KN_COUNT && value, "smx_tune_get: bad arguments");
  *value = knob(static_cast<Knob>(k));
  return 0;
}

}  // extern "C"
