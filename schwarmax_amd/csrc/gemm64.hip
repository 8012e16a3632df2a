// Thin FP64 GEMM on the matrix cores (declared in gemm64.hpp).
// This file is compiled with -mllvm -amdgpu-mfma-vgpr-form: the kernel keeps
// its FP64 accumulators in VGPRs across the reduction loop.  In the default
// form the compiler gives the MFMAs AGPR accumulators and copies all of them
// to VGPRs and back on every loop iteration (an accvgpr read and write per
// accumulator register and step pair), and its anchor-shaped product
// (100 x 20,000 x 20,000) ran 2.13 ms against 1.45 ms.
#include <hip/hip_runtime.h>
#include <cstdint>
#include <mutex>

#include "common.hpp"
#include "gemm64.hpp"

namespace smx {

typedef double f64x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// Thin FP64 GEMM on MFMA:  C (M x N) = alpha A (M x Kr) B (N x Kr)^T for few
// rows (M = the K bootstrap right-hand sides): c = -Y U and the ADMM anchors,
// where B (the library rows or the n x n inverse) streams through once and A
// stays in L2.  One wave per 64 output columns x 16 NRB rows x one split of
// the reduction, one wave per SIMD: its 4 NRB 16 x 16 FP64 accumulators of
// v_mfma_f64_16x16x4f64 take up to 224 VGPRs (the operand sets partly live
// in AGPRs).  No LDS and no barrier:
// lane (g = lane / 16, c = lane % 16) loads row c of every 16-row block of A
// and column c of every 16-column block of B at the reduction indices
// k + 4g .. k + 4g + 3 of each 16-index step (two 16-byte loads per block),
// and MFMA step s of the step takes index k + 4g + s from lane group g, so
// the four steps cover k .. k + 15.  The next step's operands are in flight
// while the current step's 16 NRB MFMAs run.  Rows past M and columns past N
// load row M - 1 / column N - 1 and are not stored; a reduction tail of fewer
// than 16 indices is loaded element by element, zero past its end.
// ---------------------------------------------------------------------------
constexpr int TH_KT = 16;     // reduction indices per step
constexpr int TH_MAXRB = GEMM_THIN_MAX_RB;   // 16-row blocks per wave at most
constexpr int TH_COLS = 64;   // output columns per wave

typedef double f64x2 __attribute__((ext_vector_type(2)));

template <int NRB>
struct ThinOps {
  f64x2 a[NRB][2], b[4][2];
};

template <bool NT>
__device__ __forceinline__ f64x2 thin_ld(const double* p) {
  if constexpr (NT) return __builtin_nontemporal_load(reinterpret_cast<const f64x2*>(p));
  else return *reinterpret_cast<const f64x2*>(p);
}

template <int NRB, bool NT>
__device__ __forceinline__ void thin_load(ThinOps<NRB>& o, const double* const (&pa)[NRB],
                                          const double* const (&pb)[4], int64_t k) {
#pragma unroll
  for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
    for (int h = 0; h < 2; ++h) o.a[rb][h] = *reinterpret_cast<const f64x2*>(pa[rb] + k + 2 * h);
#pragma unroll
  for (int cb = 0; cb < 4; ++cb)
#pragma unroll
    for (int h = 0; h < 2; ++h) o.b[cb][h] = thin_ld<NT>(pb[cb] + k + 2 * h);
}

template <int NRB>
__device__ __forceinline__ void thin_mfma(const ThinOps<NRB>& o, f64x4 (&acc)[NRB][4]) {
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int rb = 0; rb < NRB; ++rb) {
      const double a = (s & 1) ? o.a[rb][s >> 1].y : o.a[rb][s >> 1].x;
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) {
        const double b = (s & 1) ? o.b[cb][s >> 1].y : o.b[cb][s >> 1].x;
        acc[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[rb][cb], 0, 0, 0);
      }
    }
}

// Waves in order (row pass fastest, then 64-column group, then split); the
// workgroups' XCD-contiguous order (xcd_block) keeps one XCD on the A range of
// one or two splits.  out + split * split_stride receives alpha times the
// split's partial product (rows of this pass only).
template <int NRB, bool NT>
__global__ __launch_bounds__(256, 1) void gemm_thin64_kernel(
    int64_t M, int64_t N, int64_t Kr, double alpha, const double* __restrict__ A, int64_t lda,
    const double* __restrict__ B, int64_t ldb, double* __restrict__ out, int64_t ld_out,
    int64_t kchunk, int64_t split_stride, int64_t rows_per_pass, int64_t passes, int64_t ncw,
    int64_t nwaves) {
  const int64_t wv = xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
  if (wv >= nwaves) return;   // whole wave; the kernel has no barrier
  const int64_t pass = wv % passes, cw = (wv / passes) % ncw, split = wv / (passes * ncw);
  const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  const int64_t r0 = pass * rows_per_pass;
  const int64_t rend = r0 + rows_per_pass < M ? r0 + rows_per_pass : M;
  const int64_t j0 = cw * TH_COLS;
  const int64_t kb = split * kchunk;
  const int64_t len = (kb + kchunk < Kr ? kb + kchunk : Kr) - kb;
  const double* pa[NRB];
  const double* pb[4];
#pragma unroll
  for (int rb = 0; rb < NRB; ++rb) {
    const int64_t row = r0 + 16 * rb + c;
    pa[rb] = A + (row < M ? row : M - 1) * lda + kb + 4 * g;
  }
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) {
    const int64_t col = j0 + 16 * cb + c;
    pb[cb] = B + (col < N ? col : N - 1) * ldb + kb + 4 * g;
  }
  f64x4 acc[NRB][4];
#pragma unroll
  for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) acc[rb][cb] = f64x4{0.0, 0.0, 0.0, 0.0};
  const int64_t nt = len / TH_KT;
  if (nt > 0) {
    // two operand sets, loaded one step ahead; the last step's loads are
    // repeated (in bounds, unused) so that every step issues the same loads
    const auto kof = [&](int64_t t) { return TH_KT * (t < nt ? t : nt - 1); };
    ThinOps<NRB> o0, o1;
    thin_load<NRB, NT>(o0, pa, pb, 0);
    int64_t t = 0;
    for (; t + 2 <= nt; t += 2) {
      thin_load<NRB, NT>(o1, pa, pb, kof(t + 1));
      thin_mfma<NRB>(o0, acc);
      thin_load<NRB, NT>(o0, pa, pb, kof(t + 2));
      thin_mfma<NRB>(o1, acc);
    }
    if (t < nt) thin_mfma<NRB>(o0, acc);
  }
  const int64_t tail = len - nt * TH_KT;
  if (tail > 0) {
    ThinOps<NRB> o;
    const int64_t k = nt * TH_KT;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const bool in0 = 4 * g + 2 * h < tail, in1 = 4 * g + 2 * h + 1 < tail;
#pragma unroll
      for (int rb = 0; rb < NRB; ++rb) {
        o.a[rb][h].x = in0 ? pa[rb][k + 2 * h] : 0.0;
        o.a[rb][h].y = in1 ? pa[rb][k + 2 * h + 1] : 0.0;
      }
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) {
        o.b[cb][h].x = in0 ? pb[cb][k + 2 * h] : 0.0;
        o.b[cb][h].y = in1 ? pb[cb][k + 2 * h + 1] : 0.0;
      }
    }
    thin_mfma<NRB>(o, acc);
  }
  // accumulator element r of block (rb, cb): row 16 rb + g + 4 r, column 16 cb + c
  double* o = out + split * split_stride;
#pragma unroll
  for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = r0 + 16 * rb + g + 4 * r, col = j0 + 16 * cb + c;
        if (row < rend && col < N) o[row * ld_out + col] = alpha * acc[rb][cb][r];
      }
}

// SIMDs of the current device (one thin-GEMM wave each).
// Per device, queried once (devices 0..63; others are queried each call).
static int64_t device_simds() {
  constexpr int kDevs = 64;
  static int simds[kDevs] = {};
  static std::mutex mu;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0) return 1024;
  std::lock_guard<std::mutex> lock(mu);
  if (dev < kDevs && simds[dev] > 0) return simds[dev];
  int cus = 0;
  const int v = (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
                 cus > 0) ? 4 * cus : 1024;
  if (dev < kDevs) simds[dev] = v;
  return v;
}

// Launch shape of the thin GEMM: row passes of at most 16 TH_MAXRB rows, and
// the split of the reduction that minimises (rounds of waves) x (steps per
// wave) plus the partials' write and read (estimated: 0.49 us per step and
// 16-row block, 5 TB/s for the partials).
struct ThinShape {
  int64_t passes, rows_per_pass, nrb, ncw, S, kchunk, waves;
};

static ThinShape thin_shape(int64_t M, int64_t N, int64_t Kr, int64_t work_elems) {
  ThinShape s;
  s.passes = (M + 16 * TH_MAXRB - 1) / (16 * TH_MAXRB);
  s.rows_per_pass = (M + s.passes - 1) / s.passes;
  s.nrb = (s.rows_per_pass + 15) / 16;
  s.ncw = (N + TH_COLS - 1) / TH_COLS;
  const int64_t simds = device_simds();
  double best = -1.0;
  s.S = 1;
  s.kchunk = Kr;
  for (int64_t S = 1; S <= 32; ++S) {
    const int64_t kc = ((Kr + S - 1) / S + TH_KT - 1) / TH_KT * TH_KT;
    const int64_t Sx = (Kr + kc - 1) / kc;
    if (Sx != S && S > 1) continue;   // the same split as a smaller S
    if (Sx > 1 && Sx * M * N > work_elems) break;
    const int64_t waves = s.ncw * s.passes * Sx;
    const double rounds = static_cast<double>((waves + simds - 1) / simds);
    const double us = rounds * static_cast<double>((kc + TH_KT - 1) / TH_KT) * 0.49 * s.nrb +
                      (Sx > 1 ? static_cast<double>(Sx * M * N) * 16.0 / 5e6 : 0.0);
    if (best < 0.0 || us < best) {
      best = us;
      s.S = Sx;
      s.kchunk = kc;
    }
  }
  s.waves = s.ncw * s.passes * s.S;
  return s;
}

// C[i][j] = sum_z P[z][i][j] over the split partials, in split order.
__global__ void thin_split_sum_kernel(int64_t M, int64_t N, int S, const double* __restrict__ P,
                                      double* __restrict__ C, int64_t ldc) {
  const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= M * N) return;
  const int64_t i = idx / N, j = idx - i * N;
  double acc = 0.0;
  for (int z = 0; z < S; ++z) acc += P[(int64_t)z * M * N + idx];
  C[i * ldc + j] = acc;
}

int gemm_thin64(hipStream_t st, int64_t M, int64_t N, int64_t Kr, double alpha, const double* A,
                int64_t lda, const double* B, int64_t ldb, double* C, int64_t ldc, double* work,
                int64_t work_elems) {
  SMX_REQUIRE(A && B && C && M > 0 && M <= GEMM_THIN_MAX_ROWS && N > 0 && Kr > 0 && lda >= Kr &&
                  ldb >= Kr && ldc >= N && lda % 2 == 0 && ldb % 2 == 0 &&
                  reinterpret_cast<uintptr_t>(A) % 16 == 0 && reinterpret_cast<uintptr_t>(B) % 16 == 0,
              "gemm_thin64: bad arguments");
  const ThinShape s = thin_shape(M, N, Kr, work ? work_elems : 0);
  double* out = s.S > 1 ? work : C;
  const int64_t ld_out = s.S > 1 ? N : ldc;
  const unsigned grid = static_cast<unsigned>((s.waves + 3) / 4);
  const bool nt = s.passes == 1;   // B streams once; with row passes its tiles are re-read from L2
#define SMX_THIN(R)                                                                               \
  case R:                                                                                         \
    if (nt)                                                                                       \
      hipLaunchKernelGGL((gemm_thin64_kernel<R, true>), dim3(grid), dim3(256), 0, st, M, N, Kr,    \
                         alpha, A, lda, B, ldb, out, ld_out, s.kchunk, M * N, s.rows_per_pass,     \
                         s.passes, s.ncw, s.waves);                                               \
    else                                                                                          \
      hipLaunchKernelGGL((gemm_thin64_kernel<R, false>), dim3(grid), dim3(256), 0, st, M, N, Kr,   \
                         alpha, A, lda, B, ldb, out, ld_out, s.kchunk, M * N, s.rows_per_pass,     \
                         s.passes, s.ncw, s.waves);                                               \
    break;
  switch (s.nrb) {
    SMX_THIN(1)
    SMX_THIN(2)
    SMX_THIN(3)
    SMX_THIN(4)
    SMX_THIN(5)
    SMX_THIN(6)
    SMX_THIN(7)
    default:
      SMX_REQUIRE(false, "gemm_thin64: %lld row blocks", (long long)s.nrb);
  }
#undef SMX_THIN
  if (s.S > 1)
    hipLaunchKernelGGL(thin_split_sum_kernel, dim3(blocks_for(M * N, 256)), dim3(256), 0, st, M, N,
                       (int)s.S, work, C, ldc);
  SMX_LAUNCHED();
  return 0;
}

}  // namespace smx
