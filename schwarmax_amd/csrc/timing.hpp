// Optional per-stage timing with HIP events recorded on the launch stream.
// Enabled with smx_timing_enable(1); read with smx_timing_read().
#pragma once
#include <hip/hip_runtime.h>

namespace smx {

enum Stage {
  ST_ICS = 0, ST_INTEGRATE, ST_FINALIZE, ST_GRAM, ST_FACTOR, ST_RHS, ST_ADMM, ST_MAPS,
  ST_DENSITY, ST_ADMM_PRODUCT, ST_ADJOINT, ST_JACOBIAN, ST_COUNT
};

bool timing_on();
// Record the start / end of a stage on stream st (no-ops when timing is off).
void stage_begin(Stage s, hipStream_t st);
void stage_end(Stage s, hipStream_t st, long launches);

}  // namespace smx
