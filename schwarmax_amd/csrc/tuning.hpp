// Launch-shape knobs that the autotuner sets (smx_tune_set / smx_tune_get).
// Process-wide atomics; a kernel launch reads the value current at launch.
// None of them changes what a kernel computes, only how the work is laid out
// (the split-K knobs change the order of FP32 / FP64 partial sums).
#pragma once
#include <cstdint>

namespace smx {

enum Knob {
  KN_ADMM_OCC = 0,       // ADMM product workgroups per CU targeted by its split-K rule (1..8)
  KN_GRAM_WAVES,         // waves per Gram workgroup (4 or 8)
  KN_INV_SYRKX_MIN,      // inverse nodes of at least this many rows use rocBLAS syrkx updates
  KN_GEMM64_SPLIT,       // workgroups the FP64 GEMM's split-K rule fills (256..4096)
  KN_ADMM_LOADS,         // ADMM product's loads of M: 0 auto, 1 non-temporal, 2 cached
  KN_GEMM64_KERNEL,      // FP64 GEMMs with few rows: 1 thin kernel, 2 tiled kernel with split-K
  KN_LIB_BINNING,        // smx_build_library: 1 samples + LDS histograms, 2 fused integrator (same bits)
  KN_COUNT
};

int64_t knob(Knob k);

}  // namespace smx
