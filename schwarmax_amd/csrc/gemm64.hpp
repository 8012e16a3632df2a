// Thin FP64 GEMM on MFMA (gemm64.hip): C (M x N) = alpha A (M x Kr) B (N x Kr)^T
// with few rows M (the NNLS right-hand sides), both operands k-contiguous.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace smx {

constexpr int GEMM_THIN_MAX_RB = 7;                           // 16-row blocks per wave
constexpr int64_t GEMM_THIN_MAX_ROWS = 8 * 16 * GEMM_THIN_MAX_RB;   // rows M it takes (8 passes)

// Requirements: 0 < M <= GEMM_THIN_MAX_ROWS, N > 0, Kr > 0, lda and ldb even
// and A, B 16-byte aligned (two-double loads), ldc >= N.  work (work_elems
// doubles, may be null) holds the split partials; the split is chosen so that
// they fit.  Returns 0, or nonzero after recording an error.
int gemm_thin64(hipStream_t st, int64_t M, int64_t N, int64_t Kr, double alpha, const double* A,
                int64_t lda, const double* B, int64_t ldb, double* C, int64_t ldc, double* work,
                int64_t work_elems);

}  // namespace smx
