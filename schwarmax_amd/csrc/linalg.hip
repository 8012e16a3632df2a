// NNLS by ADMM with a shared factor (nnls.py:6-166) and model maps
// (model.py:14-68).
//
// Device pipeline for K right-hand sides:
//   U^T (n, m)        FP32, from the FP64 library columns and per-row scales
//   Q = U^T U         FP32 MFMA GEMM (n, n), widened to FP64; + lambda/n I;
//                     rho = tr(Q)/n
//   M = (Q + rho I)^-1  FP64, in place, recursive Schur complements on FP64
//                     MFMA GEMMs
//                     (the reference factors Q + rho I once and solves with
//                     the factor every iteration; the explicit inverse makes
//                     each iteration one product)
//   C = -Y U          right-hand-side terms (K, n), FP64 MFMA GEMM
//   per iteration:    W = W_a + (R - R_a) M  (R = rho (z - u)) on fp16-split
//                     MFMA, then the relaxation / projection / dual update
//                     of nnls.py:152-157, which also writes the next R - R_a;
//   every 100:        the anchor R_a = R and W_a = M (R_a - c) in FP64 (FP64
//                     MFMA GEMM) instead of the split product.
// z and u are kept in FP64.  The reference's solve is (Q + rho I)^-1 (rho
// (z - u) - c) (nnls.py:152-154).  The right-hand side is dominated by c,
// and cond(Q + rho I) amplifies any rounding relative to it: taking the
// anchor's part in FP64 leaves the split products (and their FP32
// accumulation) only R's drift since the anchor to round.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include <vector>
#include <cstring>
#include <cstdlib>
#include <algorithm>
#include <type_traits>

#include <mutex>
#include <atomic>
#include <rocblas/rocblas.h>

#include "common.hpp"
#include "timing.hpp"
#include "tuning.hpp"
#include "gemm64.hpp"
#include "../../include/smx.h"

namespace smx {

// ---------------------------------------------------------------------------
// GEMM conventions:  C = alpha * op(A) op(B) + beta * C
//   op(A) is (M, K): A[i*lda + k] (TA = false) or A[k*lda + i] (TA = true)
//   op(B) is (K, N): B[k*ldb + j] (TB = false) or B[j*ldb + k] (TB = true)
// lower_only: skip tiles strictly above the diagonal.
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// NT GEMM on MFMA:  C[i][j] = alpha * sum_k A[i*lda + k] B[j*ldb + k] + beta * C[i][j]
// (both operands k-contiguous).  128 x 128 output tiles, 4 waves of 64 x 64,
// K in tiles of 32 staged through LDS k-major; the next tile is loaded into
// registers while the current one is multiplied.
//   FP64: v_mfma_f64_16x16x4f64, 4 x 4 blocks per wave; lane (g = l>>4, c = l&15),
//         accumulator element r -> row g + 4r, column c.
//   FP32: v_mfma_f32_32x32x2f32, 2 x 2 blocks per wave; lane (g = l>>5, c = l&31),
//         accumulator element r -> row 8(r/4) + 4g + r%4, column c.
// lower_only skips output tiles strictly above the diagonal and stores only j <= i.
// ---------------------------------------------------------------------------
constexpr int NT_TM = 128, NT_TN = 128, NT_TK = 32;
typedef double f64x4 __attribute__((ext_vector_type(4)));

typedef float f32x16m __attribute__((ext_vector_type(16)));

template <class T> struct NtMfma;
template <> struct NtMfma<double> {
  static constexpr int BLK = 16, KS = 4, NB = 4;       // block edge, k per MFMA, blocks per wave edge
  typedef f64x4 acc_t;
  static constexpr int ACC = 4;
  __device__ static acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int lane, int r) { return (lane >> 4) + 4 * r; }
  __device__ static int col(int lane) { return lane & 15; }
  __device__ static int kof(int lane) { return lane >> 4; }
};
template <> struct NtMfma<float> {
  static constexpr int BLK = 32, KS = 2, NB = 2;
  typedef f32x16m acc_t;
  static constexpr int ACC = 16;
  __device__ static acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int lane, int r) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }
  __device__ static int col(int lane) { return lane & 31; }
  __device__ static int kof(int lane) { return lane >> 5; }
};

// One operand tile (128 rows x 32 k) per thread: EPT elements as VEC-wide vectors.
template <class T, bool VEC>
struct NtLoader {
  static constexpr int V = VEC ? 16 / (int)sizeof(T) : 1;  // elements per load
  static constexpr int TPR = NT_TK / V;                     // threads per row
  static constexpr int RPP = 256 / TPR;                     // rows per pass
  static constexpr int PASSES = NT_TM / RPP;
  T v[PASSES][V];
  __device__ void load(const T* __restrict__ P, int64_t ld, int rows, int r0, int K, int k0) {
    const int t = threadIdx.x;
    const int kk = (t % TPR) * V;
#pragma unroll
    for (int p = 0; p < PASSES; ++p) {
      const int rr = t / TPR + p * RPP;
      const int gr = r0 + rr;
      const int gk = k0 + kk;
      if (VEC) {
        if (gr < rows && gk < K) {
          const T* src = P + (int64_t)gr * ld + gk;
          if (sizeof(T) == 8) {
            const double2 x = *reinterpret_cast<const double2*>(src);
            v[p][0] = (T)x.x;
            v[p][1] = (T)x.y;
          } else {
            const float4 x = *reinterpret_cast<const float4*>(src);
            v[p][0] = (T)x.x;
            v[p][1] = (T)x.y;
            if (V > 2) {
              v[p][V > 2 ? 2 : 0] = (T)x.z;
              v[p][V > 3 ? 3 : 0] = (T)x.w;
            }
          }
        } else {
#pragma unroll
          for (int e = 0; e < V; ++e) v[p][e] = (T)0;
        }
      } else {
        v[p][0] = (gr < rows && gk < K) ? P[(int64_t)gr * ld + gk] : (T)0;
      }
    }
  }
  template <int PITCH>
  __device__ void store(T (*S)[PITCH]) const {
    const int t = threadIdx.x;
    const int kk = (t % TPR) * V;
#pragma unroll
    for (int p = 0; p < PASSES; ++p) {
      const int rr = t / TPR + p * RPP;
#pragma unroll
      for (int e = 0; e < V; ++e) S[kk + e][rr] = v[p][e];
    }
  }
};

// kchunk > 0: split-K, block z reduces indices [z kchunk, (z + 1) kchunk) and
// writes its partial product to C + z split_stride (beta must be 0).
template <class T, bool VEC>
__global__ __launch_bounds__(256, 2) void gemm_nt_kernel(int M, int N, int K, T alpha,
                                                         const T* __restrict__ A, int64_t lda,
                                                         const T* __restrict__ B, int64_t ldb, T beta,
                                                         T* __restrict__ C, int64_t ldc, int lower_only,
                                                         int64_t kchunk, int64_t split_stride) {
  typedef NtMfma<T> MM;
  if (kchunk > 0) {
    const int64_t kz = (int64_t)blockIdx.z * kchunk;
    A += kz;
    B += kz;
    C += (int64_t)blockIdx.z * split_stride;
    K = static_cast<int>(K - kz < kchunk ? K - kz : kchunk);
  }
  constexpr int PITCH = NT_TM + (sizeof(T) == 8 ? 2 : 4);
  __shared__ T As[NT_TK][PITCH];
  __shared__ T Bs[NT_TK][PITCH];
  const int i0 = blockIdx.y * NT_TM, j0 = blockIdx.x * NT_TN;
  if (lower_only && j0 > i0 + NT_TM - 1) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wi = (wave >> 1) * 64, wj = (wave & 1) * 64;
  const int g = MM::kof(lane), c = MM::col(lane);
  typename MM::acc_t acc[MM::NB][MM::NB];
#pragma unroll
  for (int a = 0; a < MM::NB; ++a)
#pragma unroll
    for (int b = 0; b < MM::NB; ++b)
#pragma unroll
      for (int r = 0; r < MM::ACC; ++r) acc[a][b][r] = (T)0;
  NtLoader<T, VEC> la, lb;
  la.load(A, lda, M, i0, K, 0);
  lb.load(B, ldb, N, j0, K, 0);
  la.store(As);
  lb.store(Bs);
  __syncthreads();
  for (int k0 = 0; k0 < K; k0 += NT_TK) {
    const bool more = k0 + NT_TK < K;
    if (more) {
      la.load(A, lda, M, i0, K, k0 + NT_TK);
      lb.load(B, ldb, N, j0, K, k0 + NT_TK);
    }
#pragma unroll
    for (int kk = 0; kk < NT_TK; kk += MM::KS) {
      T av[MM::NB], bv[MM::NB];
#pragma unroll
      for (int r = 0; r < MM::NB; ++r) av[r] = As[kk + g][wi + r * MM::BLK + c];
#pragma unroll
      for (int r = 0; r < MM::NB; ++r) bv[r] = Bs[kk + g][wj + r * MM::BLK + c];
#pragma unroll
      for (int rb = 0; rb < MM::NB; ++rb)
#pragma unroll
        for (int cb = 0; cb < MM::NB; ++cb) acc[rb][cb] = MM::mma(av[rb], bv[cb], acc[rb][cb]);
    }
    __syncthreads();
    if (more) {
      la.store(As);
      lb.store(Bs);
    }
    __syncthreads();
  }
#pragma unroll
  for (int rb = 0; rb < MM::NB; ++rb)
#pragma unroll
    for (int cb = 0; cb < MM::NB; ++cb)
#pragma unroll
      for (int r = 0; r < MM::ACC; ++r) {
        const int gi = i0 + wi + rb * MM::BLK + MM::row(lane, r);
        const int gj = j0 + wj + cb * MM::BLK + c;
        if (gi >= M || gj >= N) continue;
        if (lower_only && gj > gi) continue;
        T* p = C + (int64_t)gi * ldc + gj;
        *p = (beta == (T)0) ? alpha * acc[rb][cb][r] : alpha * acc[rb][cb][r] + beta * (*p);
      }
}

// C[i][j] = sum_z P[z][i][j] (+ beta C[i][j]) over the split-K partials.
template <class T>
__global__ void split_sum_kernel(int64_t M, int64_t N, int S, const T* __restrict__ P, T beta,
                                 T* __restrict__ C, int64_t ldc) {
  const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= M * N) return;
  const int64_t i = idx / N, j = idx - i * N;
  T acc = (T)0;
  for (int z = 0; z < S; ++z) acc += P[(int64_t)z * M * N + idx];
  T* p = C + i * ldc + j;
  *p = (beta == (T)0) ? acc : acc + beta * (*p);
}

// work (optional, work_elems elements): when the output has fewer 128 x 128
// tiles than four per CU, the reduction is split over up to 8 blocks of
// workgroups (about four workgroups per CU, for latency hiding at one
// output tile per workgroup) whose partials go to work and are summed in T.
template <class T>
static int gemm_nt(hipStream_t st, int64_t M, int64_t N, int64_t K, T alpha, const T* A, int64_t lda,
                   const T* B, int64_t ldb, T beta, T* C, int64_t ldc, int lower_only = 0,
                   T* work = nullptr, int64_t work_elems = 0) {
  if (M <= 0 || N <= 0) return 0;
  if constexpr (std::is_same<T, double>::value) {
    // few rows, a workspace for split partials, 16-byte aligned rows: the thin kernel
    if (work && !lower_only && beta == 0.0 && K > 0 && M <= GEMM_THIN_MAX_ROWS && lda % 2 == 0 &&
        ldb % 2 == 0 && reinterpret_cast<uintptr_t>(A) % 16 == 0 &&
        reinterpret_cast<uintptr_t>(B) % 16 == 0 && knob(KN_GEMM64_KERNEL) == 1)
      return gemm_thin64(st, M, N, K, alpha, A, lda, B, ldb, C, ldc, work, work_elems);
  }
  constexpr int V = 16 / (int)sizeof(T);
  const bool vec = (lda % V == 0) && (ldb % V == 0) && (K % V == 0) &&
                   (reinterpret_cast<uintptr_t>(A) % 16 == 0) && (reinterpret_cast<uintptr_t>(B) % 16 == 0);
  dim3 grid((unsigned)((N + NT_TN - 1) / NT_TN), (unsigned)((M + NT_TM - 1) / NT_TM));
  int64_t S = 1, kchunk = 0;
  const int64_t tiles = (int64_t)grid.x * grid.y;
  const int64_t target = knob(KN_GEMM64_SPLIT);   // workgroups the splits should fill
  if (work && !lower_only && tiles < target) {
    S = (target + tiles - 1) / tiles;
    if (S > 16) S = 16;
    kchunk = ((K + S - 1) / S + NT_TK - 1) / NT_TK * NT_TK;
    S = (K + kchunk - 1) / kchunk;
    if (S < 2 || S * M * N > work_elems) S = 1;
  }
  T* out = C;
  int64_t ld_out = ldc;
  T b = beta;
  if (S > 1) {
    grid.z = (unsigned)S;
    out = work;
    ld_out = N;
    b = (T)0;
  } else {
    kchunk = 0;
  }
  if (vec)
    hipLaunchKernelGGL((gemm_nt_kernel<T, true>), grid, dim3(256), 0, st, (int)M, (int)N, (int)K, alpha,
                       A, lda, B, ldb, b, out, ld_out, lower_only, kchunk, M * N);
  else
    hipLaunchKernelGGL((gemm_nt_kernel<T, false>), grid, dim3(256), 0, st, (int)M, (int)N, (int)K, alpha,
                       A, lda, B, ldb, b, out, ld_out, lower_only, kchunk, M * N);
  if (S > 1)
    hipLaunchKernelGGL(split_sum_kernel<T>, dim3(blocks_for(M * N, 256)), dim3(256), 0, st, M, N,
                       (int)S, work, beta, C, ldc);
  SMX_LAUNCHED();
  return 0;
}

// ---------------------------------------------------------------------------
// design matrix, Gram regularisation, trace
// ---------------------------------------------------------------------------
// dst[i][j] = src[i][j] * scale[j] for j < cols, 0 up to ld_dst (FP64).
__global__ void scale_cols_kernel(const double* __restrict__ src, int64_t rows, int64_t cols,
                                  const double* __restrict__ scale, int64_t ld_dst,
                                  double* __restrict__ dst) {
  const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= rows * ld_dst) return;
  const int64_t i = idx / ld_dst, j = idx - i * ld_dst;
  dst[idx] = (j < cols) ? src[i * cols + j] * scale[j] : 0.0;
}

// ---------------------------------------------------------------------------
// ADMM helpers
// ---------------------------------------------------------------------------
// The ADMM iterate is kept as one FP64 array x = w_hat + u_old: then
// z = max(0, x) and u = min(0, x) exactly, as nnls.py:155-156 compute them
// (z = max(0, x); u = (u_old + w_hat) - z is x - x = 0 or x - 0 = x).
// x has rows of nx = n rounded up to 4 (the update's aligned vectors).
__global__ void zmax_kernel(int K, int64_t n, int64_t nx, const double* __restrict__ x,
                            unsigned* __restrict__ zmax) {
  const int k = blockIdx.y;
  float m = 0.f;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    m = fmaxf(m, static_cast<float>(fmax(0.0, x[(int64_t)k * nx + i])));
  if (m > 0.f) atomicMax(zmax + k, __float_as_uint(m));
}

// W (K, n) = z = max(0, x) after at least one iteration, z_init = w_init
// before any; x has rows of nx.
__global__ void copy_out_kernel(int64_t K, int64_t n, int64_t nx, const double* __restrict__ x,
                                int iterated, double w_init, double* __restrict__ W) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= K * n) return;
  const int64_t k = i / n;
  W[i] = iterated ? fmax(0.0, x[k * nx + (i - k * n)]) : w_init;
}

// ---------------------------------------------------------------------------
// ADMM product on MFMA.  Each iteration needs W^T = R M with R (K, n) and
// M = (Q + rho I)^-1 (n, n), symmetric.  Operands are split into fp16 pairs
// x = s (hi + lo) with power-of-two scales (M: one scale from rho, since
// |M_ij| <= 1/rho; R: one scale per row from its max), and the product is
// hi*hi + hi*lo + lo*hi on v_mfma_f32_32x32x16_f16 with FP32 accumulation.
// Every fp16 product is exact in FP32, so the result carries FP32-level
// error; the dropped lo*lo term is 2^-22 relative.
//
// Delta products.  R changes slowly from one iteration to the next (its rows
// move by 1e-4 .. 1e-5 of their size per iteration at the library's
// conditioning), so after a full product w_b = M R_b an iteration may take
//   w = w_b + M_hi (R - R_b)_hi,
// reading only the hi plane of M (half the bytes) and multiplying fp16 hi
// parts only (one MFMA instead of three).  Each hi part carries at most
// 2^-12 relative error, so the delta's error is ~ 2^-11 |M| |R - R_b|: at
// most twice the full product's 2^-22 |M| |R| while every row satisfies
// max|R_k - R_b,k| <= tau max|R_k| with tau = 2^-10.  An iteration where some
// row does not takes the full product and makes its R the new base.  The choice is made on the device, from the
// row maxima the previous update collected, identically in the product and
// the update kernel, so the loop needs no host synchronisation.
//
// Work split: a workgroup (4 waves, 2 per CU) owns 128 rows of R x 256
// columns of M and a contiguous range of the reduction index i; it writes
// its partial sum (no atomics).  The update kernel adds the partials and
// applies the ADMM step.  Symmetry of M lets the B operand be read along
// rows: B[i][j] = M[j][i].
//
// Data paths.  M streams once per launch from HBM, straight into the waves'
// registers: its planes are stored in the order the MFMA's B operand wants
// them (fragment-blocked, below), so each wave loads its own 64 columns with
// one contiguous 1 KB run per (column block, 16-index step, plane) and no
// staging through LDS.  R is shared by every wave: each chunk (32 reduction
// indices x 128 rows, FP32, from L2) is split into fp16 hi / lo once per
// workgroup into a double-buffered LDS stage.  Loads run two chunks ahead of
// the MFMAs (two register buffers of M fragments), with one barrier per chunk.
// ---------------------------------------------------------------------------
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int AG_KR = 128;   // rows of R per workgroup
constexpr int AG_BN = 256;   // columns of M per workgroup
constexpr int AG_OCC = 2;    // workgroups per CU (registers: 2 waves per SIMD)
constexpr int AG_KP = 32;    // reduction indices per chunk (npad is a multiple)
constexpr int AG_WAVES = 4;
constexpr int AG_NT = 64 * AG_WAVES;            // threads of a workgroup
constexpr int AG_WCB = AG_BN / (32 * AG_WAVES); // 32-column blocks per wave
static_assert(AG_WCB == 2, "a wave owns 64 columns: two 32-column MFMA blocks");
static_assert(AG_NT == 256 && AG_KR == 128, "the R chunk staging moves 4 float4 per thread");
constexpr float AG_TAU = 1.0f / 1024.0f;   // default delta-product bound max|R - R_b| / max|R| per row
// Anchor refresh period of the ADMM loop.  A split product's rounding (the
// fp16 hi / lo representation and the FP32 accumulation over n / S terms of
// each split) is relative to the size of its operand.  The operand is
// R - R_a, where the anchor R_a = rho (z - u) is taken every AG_ANCHOR
// iterations and its product M (R_a - c) is formed in FP64 (FP64 MFMA
// GEMM), so the rounding scales with R's drift since the anchor rather than
// with R.  In those iterations the FP64 product replaces the split one.
constexpr int AG_ANCHOR = 100;

// Full product this iteration?  force > 0: yes; force < 0: no (timing of the
// delta product only); otherwise yes when some row k < K has max|D_k| >
// tau max|R_k| (D = R - R_b).  Block-uniform; every thread must call it.
__device__ __forceinline__ bool admm_full_product(int force, int64_t K,
                                                  const unsigned* __restrict__ rmax,
                                                  const unsigned* __restrict__ dmax, float tau) {
  int need = force > 0 ? 1 : 0;
  if (force == 0)
    for (int64_t k = threadIdx.x; k < K; k += blockDim.x)
      if (__uint_as_float(dmax[k]) > tau * __uint_as_float(rmax[k])) need = 1;
  return __syncthreads_or(need) != 0;
}

__device__ __forceinline__ float pow2_scale_for(float amax, int target_exp) {
  // 2^(target - ceil(log2 amax)), or 1 for an all-zero row
  if (!(amax > 0.f)) return 1.f;
  int e;
  frexpf(amax, &e);            // amax = f * 2^e, f in [0.5, 1)
  return ldexpf(1.f, target_exp - e);
}

// Fragment-blocked layout of the M planes.  The B operand of
// v_mfma_f32_32x32x16_f16 gives lane l = 32 g + cc column cc of a 32-column
// block and reduction indices 8 g .. 8 g + 7 of a 16-index step, so
//   M[j][i] -> (((jb * nch + c) * 2 + t) * 64 + 32 g + cc) * 8 + e,
// jb = j / 32, cc = j % 32, c = i / 32, t = i % 32 / 16, g = i % 16 / 8,
// e = i % 8 (nch = npad / 32): one (column block, chunk, step) is a
// contiguous 1 KB run in lane order.

// M (n, n) FP64 -> scaled fp16 hi / lo (npad x npad, fragment-blocked), zero
// padded.  One thread per 16-byte piece of the planes, in their order: piece
// p = ((jb nch + c) 2 + t) 64 + 32 g + cc holds column j = 32 jb + cc at
// indices i0 = 32 c + 16 t + 8 g .. i0 + 7 (elements 8 p .. 8 p + 7).  A
// wave writes 1 KB runs of each plane and reads the 128-byte lines of 32
// rows (lanes cc and 32 + cc share a line); with one thread per 8
// consecutive i of a row instead, the reads were contiguous but each wave's
// writes were 64 scattered 16-byte pieces (2.2 ms at n = 20,000).
__global__ void split_m_kernel(const double* __restrict__ M, int64_t n, int64_t npad,
                               const float* __restrict__ rho_p, _Float16* __restrict__ Mh,
                               _Float16* __restrict__ Ml) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= npad * npad / 8) return;
  const int64_t nch = npad / AG_KP;
  const int64_t cc = p % 32, g = (p / 32) % 2, t = (p / 64) % 2, rest = p / 128;
  const int64_t j = 32 * (rest / nch) + cc, i0 = AG_KP * (rest % nch) + 16 * t + 8 * g;
  const double sc = pow2_scale_for(1.f / *rho_p, 13);   // power of two; |M_ij| <= 1/rho
  half8 h, l;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int64_t i = i0 + e;
    const double v = (i < n && j < n) ? M[j * n + i] * sc : 0.0;
    h[e] = static_cast<_Float16>(v);
    l[e] = static_cast<_Float16>(v - static_cast<double>(h[e]));
  }
  *reinterpret_cast<half8*>(Mh + 8 * p) = h;
  *reinterpret_cast<half8*>(Ml + 8 * p) = l;
}

// R operand staging.  Blocked FP32 layout of R (written by the update
// kernel): chunk c = R[k][32 c + q], k < Kp, at (c * Kp + k) * 32 + q, so the
// 128 rows of a chunk are 16 KB in a row.  Thread t moves float4 number
// 256 u + t of it (u < 4: one contiguous 1 KB run per wave and load), i.e.
// row 32 u + t / 8, indices 4 (t % 8) .. +3, and splits it into fp16 hi / lo
// with that row's power-of-two scale: max |R_k| -> 2^14, so hi + lo = R *
// scale to 2^-22 relative.  Rows k >= K (zero padding) read row K - 1
// instead, one cache line shared by the wave, and stage zeros (scale 0).
// Column (in halfs) of 16-byte piece q (0..3) of row `row` in an R stage of
// 64-byte rows: the pieces are XOR-swizzled by bits 2..3 of the row, so that
// the 16 lanes of a fragment read (rows c .. c + 15, one piece each) and the
// 32 lanes of a staging store (4 rows x 64 bytes) each cover all 64 banks.
__device__ __forceinline__ int stage_col(int row, int q) { return 8 * (q ^ ((row >> 2) & 3)); }

typedef _Float16 half4 __attribute__((ext_vector_type(4)));
struct AdmmA {
  float4 v[4];
  float sc[4];
  int off[4];   // float4 offsets of this thread's pieces in a chunk
  __device__ void scale(const unsigned* __restrict__ rowmax, int64_t K, int64_t k0) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t k = k0 + 32 * u + threadIdx.x / 8;
      sc[u] = k < K ? pow2_scale_for(__uint_as_float(rowmax[k]), 14) : 0.f;
      off[u] = k < K ? 256 * u + threadIdx.x : (int)(K - 1 - k0) * 8 + threadIdx.x % 8;
    }
  }
  __device__ void load(const float* __restrict__ A, int64_t Kp, int64_t k0, int64_t c) {
    const float4* p = reinterpret_cast<const float4*>(A + (c * Kp + k0) * AG_KP);
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = p[off[u]];
  }
  template <bool FULL, int P>
  __device__ void store(_Float16 (*Ah)[P], _Float16 (*Al)[P]) const {
    const int t = threadIdx.x;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const float f[4] = {v[u].x * sc[u], v[u].y * sc[u], v[u].z * sc[u], v[u].w * sc[u]};
      half4 h, l;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        h[e] = static_cast<_Float16>(f[e]);
        if (FULL) l[e] = static_cast<_Float16>(f[e] - static_cast<float>(h[e]));
      }
      const int row = 32 * u + t / 8;
      const int col = stage_col(row, (t % 8) >> 1) + 4 * (t & 1);
      *reinterpret_cast<half4*>(&Ah[row][col]) = h;
      if (FULL) *reinterpret_cast<half4*>(&Al[row][col]) = l;
    }
  }
};

// This wave's M fragments of one chunk: hi (and, for a full product, lo)
// planes of its AG_WCB column blocks, both 16-index steps.  With one 128-row
// block of R, M streams once per launch and non-temporal loads (NT) keep it
// from evicting the R chunks that the workgroups of one split share in L2.
// With two or more row blocks (K > 128) the workgroups of one column block,
// adjacent in dispatch order on one XCD, read the same fragments at nearly
// the same time: plain loads let the later ones hit L2 (n = 40,000, K = 200:
// 2.37 / 0.96 ms per full / delta launch against 2.58 / 1.24 ms).
template <bool NT>
__device__ __forceinline__ half8 load_frag(const _Float16* p) {
  if constexpr (NT) return __builtin_nontemporal_load(reinterpret_cast<const half8*>(p));
  else return *reinterpret_cast<const half8*>(p);
}
struct AdmmB {
  half8 h[AG_WCB][2], l[AG_WCB][2];
  template <bool FULL, bool NT>
  __device__ void load(const _Float16* __restrict__ Mh, const _Float16* __restrict__ Ml,
                       int64_t jb0, int64_t nch, int64_t c) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int cb = 0; cb < AG_WCB; ++cb)
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int64_t off = ((((jb0 + cb) * nch + c) * 2 + t) * 64 + lane) * 8;
        h[cb][t] = load_frag<NT>(Mh + off);
        if (FULL) l[cb][t] = load_frag<NT>(Ml + off);
      }
  }
};

// MFMAs of one chunk: this wave's 128 x 64 outputs (4 row blocks x AG_WCB
// column blocks of 32 x 32).  FULL: hi*hi + hi*lo + lo*hi of R and M;
// otherwise hi*hi of D = R - R_b and M's hi plane.
template <bool FULL, int P>
__device__ __forceinline__ void admm_mfma_chunk(const _Float16 (*Ah)[P], const _Float16 (*Al)[P],
                                                const AdmmB& b, f32x16 (&acc)[4][AG_WCB]) {
  const int lane = threadIdx.x & 63, g = lane >> 5, c = lane & 31;
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
      const int col = stage_col(rb * 32 + c, 2 * t + g);
      const half8 ah = *reinterpret_cast<const half8*>(&Ah[rb * 32 + c][col]);
      half8 al;
      if (FULL) al = *reinterpret_cast<const half8*>(&Al[rb * 32 + c][col]);
#pragma unroll
      for (int cb = 0; cb < AG_WCB; ++cb) {
        acc[rb][cb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, b.h[cb][t], acc[rb][cb], 0, 0, 0);
        if (FULL) {
          acc[rb][cb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, b.l[cb][t], acc[rb][cb], 0, 0, 0);
          acc[rb][cb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, b.h[cb][t], acc[rb][cb], 0, 0, 0);
        }
      }
    }
}

// One step of the chunk loop: MFMAs of the chunk in stage s (M fragments in
// b); stage the next chunk's R piece (loaded one step earlier into a) into
// the other stage; load the R piece and M fragments of chunk c2 (two steps
// ahead) into a and b; one barrier.  The R piece is loaded before the M
// fragments, so waiting for it never waits for M loads issued later: M
// fragments stay in flight for two steps.
template <bool FULL, bool NT, int P>
__device__ __forceinline__ void admm_step(int s, int64_t c2, const float* __restrict__ A, int64_t Kp,
                                          int64_t k0, const _Float16* __restrict__ Mh,
                                          const _Float16* __restrict__ Ml, int64_t jb0, int64_t nch,
                                          _Float16 (*Ah)[AG_KR][P], _Float16 (*Al)[AG_KR][P],
                                          AdmmA& a, AdmmB& b, f32x16 (&acc)[4][AG_WCB]) {
  admm_mfma_chunk<FULL, P>(Ah[s], Al[s], b, acc);
  a.template store<FULL, P>(Ah[s ^ 1], Al[s ^ 1]);
  a.load(A, Kp, k0, c2);
  __builtin_amdgcn_sched_barrier(0);   // keep the R loads ahead of the M loads
  b.template load<FULL, NT>(Mh, Ml, jb0, nch, c2);
  __syncthreads();
}

// Main loop over the split's chunks [ch0, ch1): pairs of steps (stage 0 /
// fragments b0, stage 1 / fragments b1), then at most one chunk.  M
// fragments are loaded two steps ahead, R pieces one step ahead.  Every step
// issues its loads, for the last chunk again once the range is exhausted (in
// bounds, unused): a loop without conditional loads lets the compiler count
// outstanding loads exactly, so waiting for one chunk's data does not wait
// for a later chunk's.  The prologue issues in the loop's order, so the loop
// entry sees the outstanding loads of the steady state.
template <bool FULL, bool NT, int P>
__device__ __forceinline__ void admm_gemm_loop(int64_t K, int64_t Kp, int64_t nch, int64_t k0, int64_t jb0,
                                               int64_t ch0, int64_t ch1, const float* __restrict__ A,
                                               const unsigned* __restrict__ rowmax,
                                               const _Float16* __restrict__ Mh,
                                               const _Float16* __restrict__ Ml,
                                               _Float16 (*Ah)[AG_KR][P], _Float16 (*Al)[AG_KR][P],
                                               f32x16 (&acc)[4][AG_WCB]) {
  const int64_t nq = ch1 - ch0;
  if (nq <= 0) return;   // block-uniform
  const auto chunk = [&](int64_t q) { return ch0 + (q < nq ? q : nq - 1); };
  AdmmA a;
  AdmmB b0, b1;
  a.scale(rowmax, K, k0);
  a.load(A, Kp, k0, ch0);
  __builtin_amdgcn_sched_barrier(0);
  b0.template load<FULL, NT>(Mh, Ml, jb0, nch, ch0);
  a.template store<FULL, P>(Ah[0], Al[0]);
  a.load(A, Kp, k0, chunk(1));
  __builtin_amdgcn_sched_barrier(0);
  b1.template load<FULL, NT>(Mh, Ml, jb0, nch, chunk(1));
  __syncthreads();
  int64_t q = 0;
  for (; q + 2 <= nq; q += 2) {
    admm_step<FULL, NT, P>(0, chunk(q + 2), A, Kp, k0, Mh, Ml, jb0, nch, Ah, Al, a, b0, acc);
    admm_step<FULL, NT, P>(1, chunk(q + 3), A, Kp, k0, Mh, Ml, jb0, nch, Ah, Al, a, b1, acc);
  }
  if (q < nq) admm_mfma_chunk<FULL, P>(Ah[0], Al[0], b0, acc);   // odd count: the last chunk
}

// One ADMM product: the full one on R (scales from rmax), or the delta one on
// D (scales from dmax), as admm_full_product decides.  D == nullptr: always
// full (force must then be 1).
// The product of one workgroup: block `blk` of the launch's `nblk` blocks
// (blockIdx.x and gridDim.x for one model's launch; a grouped launch passes
// the block's index within its model's run of blocks).
template <int KC, bool NT>
__device__ __forceinline__ void admm_gemm_body(
    int64_t blk, int64_t nblk, int64_t K, int64_t Kp, int64_t npad, int chunks_per_split,
    const float* __restrict__ R, const unsigned* __restrict__ rmax, const float* __restrict__ D,
    const unsigned* __restrict__ dmax, const unsigned* __restrict__ fmax_, int force, float tau,
    const _Float16* __restrict__ Mh, const _Float16* __restrict__ Ml, float* __restrict__ part) {
  static_assert(KC == AG_KP, "the blocked layouts use 32-index chunks");
  constexpr int P = KC;   // LDS row pitch (halfs): 64 bytes, 16-byte pieces swizzled (stage_col)
  __shared__ __align__(16) _Float16 Ah[2][AG_KR][P];
  __shared__ __align__(16) _Float16 Al[2][AG_KR][P];
  const bool full = admm_full_product(force, K, fmax_, dmax, tau);
  // 1-D grid, row block of R fastest (the Kp/128 row blocks that share a tile
  // of M run back to back and share it through the caches), then column
  // block, then split; with one row block this is column block, then split.
  // Workgroups are dispatched to the 8 XCDs round-robin; xcd_block gives each
  // XCD a contiguous run of this order, so an XCD's L2 holds the R (or D)
  // chunk ranges of one or two splits rather than all of R.
  const int64_t nrb = Kp / AG_KR, ncb = npad / AG_BN;
  const int64_t bid = xcd_block(blk, nblk);
  const int64_t k0 = (bid % nrb) * AG_KR;
  const int64_t j0 = ((bid / nrb) % ncb) * AG_BN;
  const int split = (int)(bid / (nrb * ncb));
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = lane >> 5, c = lane & 31;
  const int64_t nch = npad / KC;
  const int64_t ch0 = (int64_t)split * chunks_per_split;
  const int64_t ch1 = ch0 + chunks_per_split < nch ? ch0 + chunks_per_split : nch;
  const int64_t jb0 = j0 / 32 + AG_WCB * wave;   // this wave's first 32-column block
  f32x16 acc[4][AG_WCB];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < AG_WCB; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[a][b][r] = 0.f;
  if (full)
    admm_gemm_loop<true, NT, P>(K, Kp, nch, k0, jb0, ch0, ch1, R, rmax, Mh, Ml, Ah, Al, acc);
  else
    admm_gemm_loop<false, NT, P>(K, Kp, nch, k0, jb0, ch0, ch1, D, dmax, Mh, Ml, Ah, Al, acc);
  // 32x32 accumulator layout: lane (g, c), element r -> row 8(r/4) + 4g + r%4,
  // column c.  Rows K..Kp-1 (zero padding of R) are not stored: every reader
  // of the partials sums rows k < K only.
  float* out = part + ((int64_t)split * Kp + k0) * npad + j0 + wave * (32 * AG_WCB);
  const int64_t rows = K - k0;
#pragma unroll
  for (int rb = 0; rb < 4; ++rb)
#pragma unroll
    for (int cb = 0; cb < AG_WCB; ++cb)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = rb * 32 + 8 * (r >> 2) + 4 * g + (r & 3);
        if (row < rows) out[(int64_t)row * npad + cb * 32 + c] = acc[rb][cb][r];
      }
}

template <int KC, bool NT>
__global__ __launch_bounds__(AG_NT, AG_OCC) void admm_gemm_kernel(
    int64_t K, int64_t Kp, int64_t npad, int chunks_per_split, const float* __restrict__ R,
    const unsigned* __restrict__ rmax, const float* __restrict__ D,
    const unsigned* __restrict__ dmax, const unsigned* __restrict__ fmax_, int force, float tau,
    const _Float16* __restrict__ Mh, const _Float16* __restrict__ Ml, float* __restrict__ part) {
  admm_gemm_body<KC, NT>(blockIdx.x, gridDim.x, K, Kp, npad, chunks_per_split, R, rmax, D, dmax,
                         fmax_, force, tau, Mh, Ml, part);
}

__device__ __forceinline__ float block_max_256(float v, float* red);

// Row stride of the ADMM's K x n FP64 arrays: n rounded up to 4.
__host__ __device__ __forceinline__ int64_t align_up4(int64_t n) { return (n + 3) / 4 * 4; }

// Index of R[k][i] in the blocked FP32 layout the product reads.
__device__ __forceinline__ int64_t r_tile_index(int64_t k, int64_t i, int64_t Kp) {
  return ((i / AG_KP) * Kp + k) * AG_KP + i % AG_KP;
}

__device__ __forceinline__ float block_max_256(float v, float* red) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  float m = red[0];
  for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = fmaxf(m, red[w]);
  return m;
}

// ADMM step on the summed partial products (nnls.py:152-157), next R and its
// row maxima.  Grid (ceil(n / (256 V)), K): one row per blockIdx.y, V
// consecutive elements per thread (V = 4 needs n % 4 == 0, so that every
// vector access is aligned; 4 consecutive indices stay inside one 32-wide
// chunk of R's blocked layout), one atomic per block and maximum.
//   full product:  w = sum / (sM sR); w_b <- w, R_b <- this iteration's R;
//   delta product: w = w_b + sum / (sM sD).
// It writes the next R and D = R - R_b for the next product, whichever kind
// that is.  rmax/dmax: cur scales this iteration's product, next collects the
// next one, clear is reset for the iteration after.  first != 0: the iterate
// before this step is the initial (z, u) = (w_init, 0) rather than x.
struct AdmmStep {
  double rho, alpha;
  // x <- w_hat + u_old with w_hat = alpha w + (1 - alpha) z_old; returns
  // rho (z - u) of the new iterate and the step's |z - z_old|
  __device__ __forceinline__ double apply(double w, double zo, double uo, double& x,
                                          float& dz) const {
    const double wh = alpha * w + (1.0 - alpha) * zo;
    x = wh + uo;
    const double zn = fmax(0.0, x), un = fmin(0.0, x);
    dz = static_cast<float>(fabs(zn - zo));
    return rho * (zn - un);
  }
};

// Sum over the S split partials at one position, four loads in flight, added
// in split order.
template <int V>
__device__ __forceinline__ void load_vec(const float* __restrict__ p, float (&v)[V]) {
  if constexpr (V == 4) {
    const float4 t = *reinterpret_cast<const float4*>(p);
    v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
  } else {
#pragma unroll
    for (int e = 0; e < V; ++e) v[e] = p[e];
  }
}

template <int V>
__device__ __forceinline__ void store_vec(float* __restrict__ p, const float (&v)[V]) {
  if constexpr (V == 4) {
    *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
  } else {
#pragma unroll
    for (int e = 0; e < V; ++e) p[e] = v[e];
  }
}

template <int V>
__device__ __forceinline__ void load_vec(const double* __restrict__ p, double (&v)[V]) {
  if constexpr (V == 4) {
    const double2 a = *reinterpret_cast<const double2*>(p);
    const double2 b = *reinterpret_cast<const double2*>(p + 2);
    v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
  } else {
#pragma unroll
    for (int e = 0; e < V; ++e) v[e] = p[e];
  }
}

template <int V>
__device__ __forceinline__ void store_vec(double* __restrict__ p, const double (&v)[V]) {
  if constexpr (V == 4) {
    *reinterpret_cast<double2*>(p) = make_double2(v[0], v[1]);
    *reinterpret_cast<double2*>(p + 2) = make_double2(v[2], v[3]);
  } else {
#pragma unroll
    for (int e = 0; e < V; ++e) p[e] = v[e];
  }
}

template <int V>
__device__ __forceinline__ void sum_partials(const float* __restrict__ p, int64_t stride, int S,
                                             double (&sum)[V]) {
#pragma unroll
  for (int e = 0; e < V; ++e) sum[e] = 0.0;
  int s = 0;
  for (; s + 4 <= S; s += 4) {
    float a[4][V];
#pragma unroll
    for (int q = 0; q < 4; ++q) load_vec<V>(p + (int64_t)(s + q) * stride, a[q]);
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int e = 0; e < V; ++e) sum[e] += static_cast<double>(a[q][e]);
  }
  for (; s < S; ++s) {
    float a[V];
    load_vec<V>(p + (int64_t)s * stride, a);
#pragma unroll
    for (int e = 0; e < V; ++e) sum[e] += static_cast<double>(a[e]);
  }
}

// Per-iteration buffers of the ADMM loop (all device pointers).
struct AdmmIter {
  float* R;                 // Kp x npad blocked: this iteration's operand R - R_a
  float* Rb;                // Kp x npad blocked: the operand at the last full product
  float* D;                 // Kp x npad blocked: next operand - R_b
  double* wb;               // K x n: the product of the last full product
  const double* Ra;         // K x n: the anchor R_a
  const unsigned* rcur;     // max |R_k - R_a,k| of this iteration's operand (split scales)
  unsigned* rnext;
  unsigned* rclear;
  const unsigned* dcur;     // max |D_k| of this iteration's D
  unsigned* dnext;
  unsigned* dclear;
  const unsigned* fcur;     // max |R_k| = max rho |x_k| (the delta test's reference)
  unsigned* fnext;
  unsigned* fclear;
  unsigned* full_count;     // full products taken
};

// Iteration `it` of a solve whose nine row-maximum sets start at rmax: the
// product reads set it % 3 (of each kind), the update fills set (it + 1) % 3
// for the next product and clears set (it + 2) % 3 for the one after.
__host__ __device__ inline AdmmIter admm_iter(float* R, float* Rb, float* D, double* wb,
                                              const double* Ra, unsigned* rmax,
                                              unsigned* full_count, int64_t Kp, int it) {
  AdmmIter ib;
  ib.R = R;
  ib.Rb = Rb;
  ib.D = D;
  ib.wb = wb;
  ib.Ra = Ra;
  ib.rcur = rmax + (it % 3) * Kp;
  ib.rnext = rmax + ((it + 1) % 3) * Kp;
  ib.rclear = rmax + ((it + 2) % 3) * Kp;
  ib.dcur = rmax + (3 + it % 3) * Kp;
  ib.dnext = rmax + (3 + (it + 1) % 3) * Kp;
  ib.dclear = rmax + (3 + (it + 2) % 3) * Kp;
  ib.fcur = rmax + (6 + it % 3) * Kp;
  ib.fnext = rmax + (6 + (it + 1) % 3) * Kp;
  ib.fclear = rmax + (6 + (it + 2) % 3) * Kp;
  ib.full_count = full_count;
  return ib;
}

// anchor == 1: this iteration's product was not run (the anchor was just
// refreshed, so the operand R - R_a is zero): the partial sums count as 0 and
// the iteration counts as a full product (new base R_b = 0, w_b = 0).
// One block of the update: elements [256 V bx, 256 V (bx + 1)) of row k.
// The K x n FP64 arrays (x, c, W_c, R_a, w_b) have rows of nx = n rounded up
// to 4, so every 4-element vector is aligned; the elements past n are
// padding: masked to zero here (x, R, D) and left out of every maximum.
template <int V>
__device__ __forceinline__ void admm_update_body(
    int64_t bx, int64_t k, int64_t K, int64_t n, int64_t nx, int64_t Kp, int64_t npad, int S,
    const float* __restrict__ part, const double* __restrict__ rho_p, const double* __restrict__ wc,
    double* __restrict__ x, int first, int force, int anchor, float tau, double w_init,
    const AdmmIter& b, double alpha, const int* __restrict__ active, unsigned* __restrict__ delta) {
  __shared__ float red[4];
  const bool full = admm_full_product(force, K, b.fcur, b.dcur, tau);
  const int64_t i0 = (bx * (int64_t)blockDim.x + threadIdx.x) * V;
  if (bx == 0 && threadIdx.x == 0) {
    b.rclear[k] = 0u;
    b.dclear[k] = 0u;
    b.fclear[k] = 0u;
    if (k == K - 1)
      for (int64_t q = K; q < Kp; ++q) {
        b.rclear[q] = 0u;
        b.dclear[q] = 0u;
        b.fclear[q] = 0u;
      }
    if (k == 0 && full && !anchor) *b.full_count += 1u;
  }
  float rmax = 0.f, dmx = 0.f, fmx = 0.f, zmv = 0.f;
  if (i0 < n) {
    const int64_t ri = r_tile_index(k, i0, Kp);
    float rcur[V], rb[V], rn[V], dn[V];
    // this iteration's R: the new base after a full product, and kept by an
    // inactive row; an active row after a delta product does not need it
    if (full || !active[k]) load_vec<V>(b.R + ri, rcur);
    if (full) {
      store_vec<V>(b.Rb + ri, rcur);
#pragma unroll
      for (int e = 0; e < V; ++e) rb[e] = rcur[e];
    } else {
      load_vec<V>(b.Rb + ri, rb);
    }
    const int64_t idx = k * nx + i0;
    if (!active[k]) {
      double xo[V];
      load_vec<V>(x + idx, xo);
#pragma unroll
      for (int e = 0; e < V; ++e) {
        rn[e] = rcur[e];
        if (i0 + e < n) fmx = fmaxf(fmx, static_cast<float>(*rho_p * fabs(xo[e])));
      }
    } else {
      const float sM = pow2_scale_for(static_cast<float>(1.0 / *rho_p), 13);
      const float sA = pow2_scale_for(__uint_as_float(full ? b.rcur[k] : b.dcur[k]), 14);
      const double inv = 1.0 / (static_cast<double>(sM) * static_cast<double>(sA));
      const AdmmStep st{*rho_p, alpha};
      double sum[V];
      if (anchor) {
#pragma unroll
        for (int e = 0; e < V; ++e) sum[e] = 0.0;
      } else {
        sum_partials<V>(part + k * npad + i0, Kp * npad, S, sum);
      }
      double w[V], xo[V], xn[V];
      // the split partials are added in FP64; w_b keeps the full product
      if (full) {
#pragma unroll
        for (int e = 0; e < V; ++e) {
          const double p = sum[e] * inv;
          b.wb[idx + e] = p;
          w[e] = wc[idx + e] + p;
        }
      } else {
#pragma unroll
        for (int e = 0; e < V; ++e) w[e] = wc[idx + e] + (b.wb[idx + e] + sum[e] * inv);
      }
      if (first) {
#pragma unroll
        for (int e = 0; e < V; ++e) xo[e] = w_init;
      } else {
        load_vec<V>(x + idx, xo);
      }
#pragma unroll
      for (int e = 0; e < V; ++e) {
        const double zo = first ? w_init : fmax(0.0, xo[e]);
        const double uo = first ? 0.0 : fmin(0.0, xo[e]);
        float dz;
        const double rf = st.apply(w[e], zo, uo, xn[e], dz);
        if (i0 + e < n) {
          rn[e] = static_cast<float>(rf - b.Ra[idx + e]);
          fmx = fmaxf(fmx, static_cast<float>(fabs(rf)));
          zmv = fmaxf(zmv, dz);
        } else {
          xn[e] = 0.0;
          rn[e] = 0.f;
        }
      }
      store_vec<V>(x + idx, xn);
      store_vec<V>(b.R + ri, rn);
    }
#pragma unroll
    for (int e = 0; e < V; ++e) {
      dn[e] = rn[e] - rb[e];
      rmax = fmaxf(rmax, fabsf(rn[e]));
      dmx = fmaxf(dmx, fabsf(dn[e]));
    }
    store_vec<V>(b.D + ri, dn);
  }
  rmax = block_max_256(rmax, red);
  __syncthreads();
  dmx = block_max_256(dmx, red);
  __syncthreads();
  fmx = block_max_256(fmx, red);
  __syncthreads();
  zmv = block_max_256(zmv, red);
  if (threadIdx.x == 0) {
    if (rmax > 0.f) atomicMax(b.rnext + k, __float_as_uint(rmax));
    if (dmx > 0.f) atomicMax(b.dnext + k, __float_as_uint(dmx));
    if (fmx > 0.f) atomicMax(b.fnext + k, __float_as_uint(fmx));
    if (zmv > 0.f) atomicMax(delta + k, __float_as_uint(zmv));
  }
}

template <int V>
__global__ __launch_bounds__(256) void admm_update_kernel(
    int64_t K, int64_t n, int64_t Kp, int64_t npad, int S, const float* __restrict__ part,
    const double* __restrict__ rho_p, const double* __restrict__ wc, double* __restrict__ x,
    int first, int force, int anchor, float tau, double w_init, AdmmIter b, double alpha,
    const int* __restrict__ active, unsigned* __restrict__ delta) {
  admm_update_body<V>(blockIdx.x, blockIdx.y, K, n, align_up4(n), Kp, npad, S, part, rho_p, wc, x,
                      first, force, anchor, tau, w_init, b, alpha, active, delta);
}

// ---------------------------------------------------------------------------
// Grouped ADMM loop: the solves of several models (one K for all) advance one
// iteration per launch pair, one product launch and one update launch for
// the whole group.  Each model owns a contiguous run of a launch's blocks.
// A product run starts at a multiple of 8, so that a block's index within
// its run and its index in the launch agree modulo 8 and xcd_block gives
// every model the XCD-contiguous order of its own launch: each model's
// arithmetic is that of its single-model solve.
// ---------------------------------------------------------------------------
struct AdmmModel {
  int64_t n, npad;
  int64_t pblk0, pnblk;   // product blocks [pblk0, pblk0 + pnblk); the run is padded to 8
  int64_t ublk0, unblk;   // update blocks (x) [ublk0, ublk0 + unblk)
  int S, cps;
  float* Rp;
  float* Rb;
  float* D;
  double* wb;
  const double* Ra;
  unsigned* rmax;
  unsigned* full_count;
  const _Float16* Mh;
  const _Float16* Ml;
  float* part;
  const double* rho64;
  const double* Wc;
  double* x;
  const int* active;
  unsigned* delta;
  double w_init;
};

// The model whose run holds launch block b (runs are in increasing order).
__device__ __forceinline__ int admm_model_of(const AdmmModel* __restrict__ mods, int G, int64_t b,
                                             bool product) {
  int g = 0;
  for (int q = 1; q < G; ++q)
    if (b >= (product ? mods[q].pblk0 : mods[q].ublk0)) g = q;
  return g;
}

template <int KC, bool NT>
__global__ __launch_bounds__(AG_NT, AG_OCC) void admm_gemm_group_kernel(
    const AdmmModel* __restrict__ mods, int G, int64_t K, int64_t Kp, int it, int force, float tau) {
  const AdmmModel& d = mods[admm_model_of(mods, G, blockIdx.x, true)];
  const int64_t blk = (int64_t)blockIdx.x - d.pblk0;
  if (blk >= d.pnblk) return;   // padding block: the whole workgroup leaves
  const AdmmIter ib = admm_iter(d.Rp, d.Rb, d.D, d.wb, d.Ra, d.rmax, d.full_count, Kp, it);
  admm_gemm_body<KC, NT>(blk, d.pnblk, K, Kp, d.npad, d.cps, d.Rp, ib.rcur, d.D, ib.dcur, ib.fcur,
                         force, tau, d.Mh, d.Ml, d.part);
}

template <int V>
__global__ __launch_bounds__(256) void admm_update_group_kernel(
    const AdmmModel* __restrict__ mods, int G, int64_t K, int64_t Kp, int it, int first, int force,
    int anchor, float tau, double alpha) {
  const AdmmModel& d = mods[admm_model_of(mods, G, blockIdx.x, false)];
  const int64_t bx = (int64_t)blockIdx.x - d.ublk0;
  const AdmmIter ib = admm_iter(d.Rp, d.Rb, d.D, d.wb, d.Ra, d.rmax, d.full_count, Kp, it);
  admm_update_body<V>(bx, blockIdx.y, K, d.n, align_up4(d.n), Kp, d.npad, d.S, d.part, d.rho64, d.Wc,
                      d.x, first, force, anchor, tau, d.w_init, ib, alpha, d.active, d.delta);
}

// Anchor refresh (before the product of an iteration): R_a = rho (z - u) =
// rho |x| of the current iterate (rho w_init before the first), V = R_a - c
// for the FP64 product W_a = M V that follows, and the operand R - R_a = 0.
// Grid (ceil(n / 256), K).
__global__ __launch_bounds__(256) void admm_anchor_kernel(int64_t K, int64_t n, int64_t Kp,
                                                          int first, double w_init,
                                                          const double* __restrict__ rho_p,
                                                          const double* __restrict__ x,
                                                          const double* __restrict__ C,
                                                          double* __restrict__ Ra,
                                                          double* __restrict__ Va,
                                                          float* __restrict__ R) {
  const int64_t k = blockIdx.y;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t idx = k * align_up4(n) + i;
  const double rf = *rho_p * (first ? w_init : fabs(x[idx]));
  Ra[idx] = rf;
  Va[idx] = rf - C[idx];
  R[r_tile_index(k, i, Kp)] = 0.f;
}

static int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

struct NnlsWs {
  _Float16* Uh;   // gpad x mpad, fp16 hi plane of Ut (row i scaled by us[i]), Gram-blocked
  _Float16* Ul;   // gpad x mpad, fp16 lo plane
  float* us;      // gpad power-of-two row scales
  int64_t gpad, mpad;
  double* Q64;    // n x n FP64, inverted in place
  double* work64; // n*n + n work of the recursive inverse
  double* Ys;     // K x ldu: Y times the design rows' scales
  double* C;      // K x n: c = -Y U (nnls.py:143)
  double* Wc;     // K x n: W_a = M (R_a - c), the solve's FP64 part
  double* Ra;     // K x n: the anchor R_a = rho (z - u) at the last refresh
  double* Va;     // K x n: R_a - c
  double* x;      // K x n ADMM iterate, z = max(0, x), u = min(0, x)
  float* rho;     // 1
  double* rho64;  // 1
  unsigned* delta;  // K
  unsigned* zmax;   // K
  int* active;      // K
  _Float16* Mh;     // npad x npad
  _Float16* Ml;     // npad x npad
  float* Rp;        // Kp x npad, blocked (r_tile_index)
  float* Rb;        // Kp x npad, blocked: R at the last full product
  float* D;         // Kp x npad, blocked: R - Rb
  double* wb;       // K x n: the product of the last full product
  unsigned* full_count;  // full ADMM products taken by the last solve
  float* part;      // S x Kp x npad
  unsigned* rmax;   // 9 x Kp: three rotating row-maximum sets of the operand |R - R_a|,
                    // three of |D|, three of |R| = rho |x|
  int64_t npad, Kp;
  int64_t nx;     // row stride of C, W_c, R_a, V_a, x, w_b: n rounded up to 4
  int S;
  int64_t scratch_bytes, state_bytes, bytes;
};

// Splits of the ADMM product's reduction: enough workgroups for `occ` on each
// of the 256 CUs, and at least one workgroup per CU.  occ is the product's
// tuning knob (smx_admm_set_occupancy, set by the autotuner; default AG_OCC,
// the workgroups per CU that the kernel's LDS and registers allow at once;
// larger values queue further rounds of workgroups).  Workspaces size the
// partial buffer for occ = AG_OCC_MAX, and S grows with occ, so a later
// setting never overruns one.
constexpr int AG_OCC_MAX = 8;

static int admm_splits_for(int64_t npad, int64_t Kp, int occ) {
  int64_t tiles = (npad / AG_BN) * (Kp / AG_KR);   // workgroups of one split
  if (tiles < 1) tiles = 1;
  int64_t S = (256 * (int64_t)occ) / tiles;
  const int64_t fill = (256 + tiles - 1) / tiles;
  if (S < fill) S = fill;
  if (S > 16) S = 16;
  const int64_t nch = npad / AG_KP;
  if (S > nch) S = nch;
  return static_cast<int>(S);
}

// Non-temporal loads of M (knob KN_ADMM_LOADS).  Automatic: non-temporal
// when M streams once per launch through caches it would only evict, i.e.
// with one 128-row block of R and a hi plane larger than the memory-side
// cache can keep between iterations; cached when the row blocks' workgroups
// share M's tiles through L2 (K > 128), or when the hi plane stays resident
// (n = 8,000: delta products 30 us cached against 41 us non-temporal, full
// products 70 / 69 us; n = 20,000: 347 / 189 us cached against 329 / 169).
constexpr int64_t AG_RESIDENT_PLANE = 192ll << 20;
static bool admm_nt_loads(int64_t Kp, int64_t hi_plane_bytes) {
  const int64_t v = knob(KN_ADMM_LOADS);
  if (v != 0) return v == 1;
  return Kp == AG_KR && hi_plane_bytes > AG_RESIDENT_PLANE;
}

// One ADMM product launch (admm_gemm_kernel) over S splits of the reduction.
// rmax: the operand's row maxima (split scales of a full product); fmax_:
// max rho |x_k|, against which tau tests dmax (unused when force != 0).
static void launch_admm_product(hipStream_t st, int64_t K, int64_t Kp, int64_t npad, int S,
                                const float* R, const unsigned* rmax, const float* D,
                                const unsigned* dmax, const unsigned* fmax_, int force, float tau,
                                const _Float16* Mh, const _Float16* Ml, float* part) {
  const int64_t nch = npad / AG_KP;
  const int cps = static_cast<int>((nch + S - 1) / S);
  const dim3 grid((unsigned)((npad / AG_BN) * S * (Kp / AG_KR)));
  if (admm_nt_loads(Kp, npad * npad * 2))
    hipLaunchKernelGGL((admm_gemm_kernel<AG_KP, true>), grid, dim3(AG_NT), 0, st, K, Kp, npad, cps,
                       R, rmax, D, dmax, fmax_, force, tau, Mh, Ml, part);
  else
    hipLaunchKernelGGL((admm_gemm_kernel<AG_KP, false>), grid, dim3(AG_NT), 0, st, K, Kp, npad, cps,
                       R, rmax, D, dmax, fmax_, force, tau, Mh, Ml, part);
}

// Delta-product bound tau (the product's accuracy / speed knob, process-wide
// like the occupancy): an iteration takes the delta product while every row
// satisfies max|R_k - R_b,k| <= tau max|R_k|.  0 makes every product full.
static std::atomic<float> g_admm_tau{AG_TAU};

static int admm_splits(int64_t npad, int64_t Kp) {
  return admm_splits_for(npad, Kp, static_cast<int>(knob(KN_ADMM_OCC)));
}

// The workspace of a solve has two parts.  The factor scratch holds the
// Gram's fp16 planes of U and the recursive inverse's work area, which the
// FP64 GEMMs also use for their split-K partials; the solve state holds
// everything the ADMM loop keeps.  A single solve keeps both in one buffer
// (scratch first, `carve`); a grouped solve gives each model its own state
// and lets the models share one scratch, sized for the largest (the models'
// Gram, inverse and anchor products run one after another).
static NnlsWs carve_parts(void* scratch, void* state, int64_t n, int64_t m, int64_t K) {
  NnlsWs w{};
  const int64_t ldu = align_up(m, 16);
  char* p = static_cast<char*>(scratch);
  int64_t off = 0;
  auto take = [&](int64_t bytes) {
    char* q = p ? p + off : nullptr;
    off += align_up(bytes, 256);
    return q;
  };
  w.gpad = align_up(n, 256);
  w.mpad = align_up(m, 32);
  w.Uh = reinterpret_cast<_Float16*>(take(w.gpad * w.mpad * 2));
  w.Ul = reinterpret_cast<_Float16*>(take(w.gpad * w.mpad * 2));
  w.us = reinterpret_cast<float*>(take(w.gpad * 4));
  w.work64 = reinterpret_cast<double*>(take((n * n + n) * 8));
  w.scratch_bytes = off;
  p = static_cast<char*>(state);
  off = 0;
  w.Q64 = reinterpret_cast<double*>(take(n * n * 8));
  w.Ys = reinterpret_cast<double*>(take(K * ldu * 8));
  w.nx = align_up4(n);
  w.C = reinterpret_cast<double*>(take(K * w.nx * 8));
  w.Wc = reinterpret_cast<double*>(take(K * w.nx * 8));
  w.Ra = reinterpret_cast<double*>(take(K * w.nx * 8));
  w.Va = reinterpret_cast<double*>(take(K * w.nx * 8));
  w.x = reinterpret_cast<double*>(take(K * w.nx * 8));
  w.rho = reinterpret_cast<float*>(take(4));
  w.rho64 = reinterpret_cast<double*>(take(8));
  w.delta = reinterpret_cast<unsigned*>(take(K * 4));
  w.zmax = reinterpret_cast<unsigned*>(take(K * 4));
  w.active = reinterpret_cast<int*>(take(K * 4));
  w.npad = align_up(n, AG_BN);
  w.Kp = align_up(K, AG_KR);
  w.S = admm_splits(w.npad, w.Kp);
  w.Mh = reinterpret_cast<_Float16*>(take(w.npad * w.npad * 2));
  w.Ml = reinterpret_cast<_Float16*>(take(w.npad * w.npad * 2));
  w.Rp = reinterpret_cast<float*>(take(w.Kp * w.npad * 4));
  w.part = reinterpret_cast<float*>(
      take((int64_t)admm_splits_for(w.npad, w.Kp, AG_OCC_MAX) * w.Kp * w.npad * 4));
  w.rmax = reinterpret_cast<unsigned*>(take(9 * w.Kp * 4));
  w.Rb = reinterpret_cast<float*>(take(w.Kp * w.npad * 4));
  w.D = reinterpret_cast<float*>(take(w.Kp * w.npad * 4));
  w.wb = reinterpret_cast<double*>(take(K * w.nx * 8));
  w.full_count = reinterpret_cast<unsigned*>(take(4));
  w.state_bytes = off;
  w.bytes = w.scratch_bytes + w.state_bytes;
  return w;
}

static NnlsWs carve(void* base, int64_t n, int64_t m, int64_t K) {
  char* b = static_cast<char*>(base);
  const int64_t sb = carve_parts(nullptr, nullptr, n, m, K).scratch_bytes;
  return carve_parts(b, b ? b + sb : nullptr, n, m, K);
}

__global__ void fill_u32_kernel(unsigned* p, int64_t n, unsigned v) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// Gram Q = U^T U on fp16 MFMA with hi/lo splits (three products, FP32
// accumulation), written straight into the FP64 matrix that is inverted.
// Row i of Ut (orbit i, m data rows) is scaled by a power of two s_i taken
// from its maximum, so hi + lo holds it to 2^-22 relative and both stay in
// fp16's normal range; Q_ij = acc_ij / (s_i s_j).  At plain-20k this Gram
// moves logL by 3.8e-7 against an FP64 Gram (an FP32 Gram: 2.8e-7), far
// below the ADMM product's share of the error.
// Planes are blocked like the ADMM product's: tile (128-row block I, 32-index
// chunk c) = rows 128 I + r, indices 32 c + q at ((I nchk + c) 128 + r) 32 + q.
// ---------------------------------------------------------------------------
constexpr int GR_TM = 128;   // rows (A) per tile; the B operand is two such blocks
// Waves per Gram workgroup (the tuning knob KN_GRAM_WAVES: 4 or 8), two
// workgroups per CU either way (61 KB of LDS each).  8 waves of 128 x 32
// outputs (114 VGPRs) read all 128 A rows from LDS per wave, twice the LDS
// reads of 4 waves of 128 x 64 (242 VGPRs), but run four waves per SIMD
// instead of two: at plain-20k 25.8 against 29.8 ms, at n = 40,000,
// m = 35,600 174.5 against 192.0 ms.
template <int WAVES>
struct GramShape {
  static constexpr int NT = 64 * WAVES;
  static constexpr int WCB = 2 * GR_TM / (32 * WAVES);   // 32-column blocks per wave
  static constexpr int PASSES = GR_TM / (NT / 4);         // staging passes per 128-row tile
  static constexpr int OCC = 2;                            // workgroups per CU (registers, LDS)
  static_assert(WCB >= 1 && PASSES >= 1, "Gram wave layout");
};

__device__ __forceinline__ int64_t g_tile_index(int64_t i, int64_t k, int64_t nchk) {
  return ((i / GR_TM * nchk + k / AG_KP) * GR_TM + i % GR_TM) * AG_KP + k % AG_KP;
}

// One workgroup per (padded) row i: the scaled fp16 hi / lo planes of
// U^T[i, :] = lib[i, :m] * scale.
__global__ __launch_bounds__(256) void build_ut_split_kernel(
    const double* __restrict__ lib, int64_t ld_lib, int64_t n, int64_t m,
    int64_t mpad, const double* __restrict__ scale,
    _Float16* __restrict__ Uh, _Float16* __restrict__ Ul, float* __restrict__ us) {
  __shared__ float red[4];
  const int64_t i = blockIdx.x;
  const int64_t nchk = mpad / AG_KP;
  float mx = 0.f;
  if (i < n)
    for (int64_t j = threadIdx.x; j < m; j += blockDim.x)
      mx = fmaxf(mx, static_cast<float>(fabs(lib[i * ld_lib + j] * scale[j])));
  mx = block_max_256(mx, red);
  const float sc = pow2_scale_for(mx, 14);
  if (threadIdx.x == 0) us[i] = sc;
  for (int64_t j = threadIdx.x; j < mpad; j += blockDim.x) {
    const double u = (i < n && j < m) ? lib[i * ld_lib + j] * scale[j] : 0.0;
    const double v = u * static_cast<double>(sc);
    const _Float16 h = static_cast<_Float16>(v);
    const int64_t d = g_tile_index(i, j, nchk);
    Uh[d] = h;
    Ul[d] = static_cast<_Float16>(v - static_cast<double>(h));
  }
}

// Lower-triangle tiles: row block I (128 rows), column block J (256 columns)
// with 256 J <= 128 I, i.e. I >= 2 J, enumerated column block by column
// block: column block J holds the nrb - 2 J tiles I = 2 J .. nrb - 1,
// starting at C(J) = J (nrb + 1 - J).
__device__ __forceinline__ void gram_tile(int64_t q, int64_t nrb, int64_t& I, int64_t& J) {
  const double b = static_cast<double>(nrb + 1);
  int64_t j = static_cast<int64_t>((b - sqrt(b * b - 4.0 * static_cast<double>(q))) * 0.5);
  if (j < 0) j = 0;
  while (j > 0 && j * (nrb + 1 - j) > q) --j;
  while ((j + 1) * (nrb - j) <= q) ++j;
  J = j;
  I = 2 * j + (q - j * (nrb + 1 - j));
}

// Workgroups go to the 8 XCDs round-robin by id; tile q of XCD x's
// consecutive share runs on that XCD, so its co-resident workgroups work on
// neighbouring tiles of one column block and read that block's B operand
// through the same L2.  Returns false for the padding workgroups.
__device__ __forceinline__ bool gram_tile_xcd(int64_t w, int64_t tiles, int64_t nrb, int64_t& I,
                                              int64_t& J) {
  const int64_t per = (tiles + 7) / 8;
  const int64_t q = (w % 8) * per + w / 8;
  if (q >= tiles) return false;
  gram_tile(q, nrb, I, J);
  return true;
}

static int64_t gram_tiles(int64_t gpad) {
  const int64_t nrb = gpad / GR_TM;     // even: gpad is a multiple of 256
  const int64_t q = nrb / 2;
  return q * (q + 1);
}

static int64_t gram_grid(int64_t gpad) { return 8 * ((gram_tiles(gpad) + 7) / 8); }

// Segments of the Gram's reduction accumulated in FP32: each of GR_FLUSH
// segments covers 1/GR_FLUSH of the m products of every entry, and the
// workgroup adds its FP32 sums into the FP64 Q at the end of the segment
// (the first segment stores them).  One FP32 accumulation over all
// m = 17,288 design rows of plain-20k moves the NNLS weights by 1.7e-4 from
// an FP64 Gram (the library's conditioning amplifies it); 8 segments keep
// that near 1e-5 for about two extra reads and writes of the lower triangle
// of Q.  A workgroup runs all segments of its tile, so the read-modify-writes
// of Q overlap other workgroups' MFMAs (one launch per segment: 26.5 ms at
// plain-20k).
#ifndef GR_FLUSH
#define GR_FLUSH 8
#endif

// One (128 x 256) tile of Q per workgroup (WAVES waves of 128 x 256/WAVES),
// 32-index chunks staged through LDS, hi*hi + hi*lo + lo*hi per 16-deep
// step; the nchk_total chunks in segments of seg_chunks chunks, after each of which the
// segment's sums are stored (first segment) or added to Q[i][j] for
// j <= i < n.
template <int WAVES>
__global__ __launch_bounds__(GramShape<WAVES>::NT, GramShape<WAVES>::OCC) void gram_split_kernel(
    int64_t n, int64_t gpad, int64_t mpad, const _Float16* __restrict__ Uh,
    const _Float16* __restrict__ Ul, const float* __restrict__ us, double* __restrict__ Q,
    int64_t seg_chunks, int64_t nchk_total) {
  constexpr int GR_NT = GramShape<WAVES>::NT;
  constexpr int GR_WCB = GramShape<WAVES>::WCB;
  constexpr int GR_PASSES = GramShape<WAVES>::PASSES;
  constexpr int GR_WAVES = WAVES;
  constexpr int P = AG_KP + 8;
  __shared__ __align__(16) _Float16 Bh[2 * GR_TM][P];
  __shared__ __align__(16) _Float16 Bl[2 * GR_TM][P];
  __shared__ __align__(16) _Float16 Ah[GR_TM][P];
  __shared__ __align__(16) _Float16 Al[GR_TM][P];
  const int64_t nrb = gpad / GR_TM;
  const int64_t nq = (nrb / 2) * (nrb / 2 + 1);
  int64_t I, J;
  if (!gram_tile_xcd(blockIdx.x, nq, nrb, I, J)) return;   // whole workgroup
  const int64_t i0 = I * GR_TM, j0 = J * 2 * GR_TM;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int g = lane >> 5, c = lane & 31;
  const int64_t nchk = mpad / AG_KP;
  const int64_t tileA = I * nchk, tileB0 = 2 * J * nchk, tileB1 = (2 * J + 1) * nchk;
  const int r0 = tid >> 2, seg = 8 * (tid & 3);   // rows r0 + (GR_NT / 4) h of a 128-row tile
  f32x16 acc[4][GR_WCB];
  for (int64_t ch0 = 0; ch0 < nchk_total; ch0 += seg_chunks) {
  const int64_t ch1 = ch0 + seg_chunks < nchk_total ? ch0 + seg_chunks : nchk_total;
  const bool first = ch0 == 0;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < GR_WCB; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[a][b][r] = 0.f;
  for (int64_t ch = ch0; ch < ch1; ++ch) {
    // stage chunk ch: each plane tile is 512 16-byte runs, thread t moves runs t + GR_NT h
    const int64_t oa = ((tileA + ch) * GR_TM) * AG_KP;
    const int64_t ob0 = ((tileB0 + ch) * GR_TM) * AG_KP;
    const int64_t ob1 = ((tileB1 + ch) * GR_TM) * AG_KP;
#pragma unroll
    for (int h = 0; h < GR_PASSES; ++h) {
      const int rr = r0 + (GR_NT / 4) * h;
      const int64_t e = (int64_t)rr * AG_KP + seg;
      *reinterpret_cast<uint4*>(&Ah[rr][seg]) = *reinterpret_cast<const uint4*>(Uh + oa + e);
      *reinterpret_cast<uint4*>(&Al[rr][seg]) = *reinterpret_cast<const uint4*>(Ul + oa + e);
      *reinterpret_cast<uint4*>(&Bh[rr][seg]) = *reinterpret_cast<const uint4*>(Uh + ob0 + e);
      *reinterpret_cast<uint4*>(&Bl[rr][seg]) = *reinterpret_cast<const uint4*>(Ul + ob0 + e);
      *reinterpret_cast<uint4*>(&Bh[GR_TM + rr][seg]) = *reinterpret_cast<const uint4*>(Uh + ob1 + e);
      *reinterpret_cast<uint4*>(&Bl[GR_TM + rr][seg]) = *reinterpret_cast<const uint4*>(Ul + ob1 + e);
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < AG_KP / 16; ++t) {
      half8 bh[GR_WCB], bl[GR_WCB];
#pragma unroll
      for (int cb = 0; cb < GR_WCB; ++cb) {
        bh[cb] = *reinterpret_cast<const half8*>(&Bh[wave * (2 * GR_TM / GR_WAVES) + cb * 32 + c][16 * t + 8 * g]);
        bl[cb] = *reinterpret_cast<const half8*>(&Bl[wave * (2 * GR_TM / GR_WAVES) + cb * 32 + c][16 * t + 8 * g]);
      }
#pragma unroll
      for (int rb = 0; rb < 4; ++rb) {
        const half8 ah = *reinterpret_cast<const half8*>(&Ah[rb * 32 + c][16 * t + 8 * g]);
        const half8 al = *reinterpret_cast<const half8*>(&Al[rb * 32 + c][16 * t + 8 * g]);
#pragma unroll
        for (int cb = 0; cb < GR_WCB; ++cb) {
          acc[rb][cb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh[cb], acc[rb][cb], 0, 0, 0);
          acc[rb][cb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl[cb], acc[rb][cb], 0, 0, 0);
          acc[rb][cb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh[cb], acc[rb][cb], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
  // The scales and addresses of the flush are the same in every segment; the
  // compiler would hoist them out of the segment loop and keep them (64 FP64
  // scales, addresses and masks per lane) live across the chunk loop,
  // spilling.  Laundering their inputs once per segment keeps them local.
  const float* usg = us;
  double* Qg = Q;
  int64_t i0g = i0, j0g = j0;
  int gg = g, cg = c;
  asm volatile("" : "+s"(usg), "+s"(Qg), "+s"(i0g), "+s"(j0g), "+v"(gg), "+v"(cg));
  // 32x32 accumulator layout: lane (g, c), element r -> row 8(r/4) + 4g + r%4, column c
#pragma unroll
  for (int rb = 0; rb < 4; ++rb)
#pragma unroll
    for (int cb = 0; cb < GR_WCB; ++cb) {
      const int64_t j = j0g + wave * (2 * GR_TM / GR_WAVES) + cb * 32 + cg;
      const double sj = (j < n) ? static_cast<double>(usg[j]) : 1.0;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int64_t i = i0g + rb * 32 + 8 * (r >> 2) + 4 * gg + (r & 3);
        if (i < n && j <= i) {
          const double v = static_cast<double>(acc[rb][cb][r]) / (static_cast<double>(usg[i]) * sj);
          double* q = Qg + i * n + j;
          *q = first ? v : *q + v;
        }
      }
    }
  }   // segments
}


// ---------------------------------------------------------------------------
// SPD inverse in FP64 by recursive Schur complements, in place:
//   [A11 A21^T; A21 A22]^-1 with X = A11^-1, T = A21 X, S = A22 - T A21^T:
//   A22 <- S^-1,  A21 <- -S^-1 T,  A11 <- X - T^T A21.
// The recursion reads and writes the lower triangle: the two symmetric
// updates (S and the new A11) are formed on their lower triangles only, and
// X and S^-1 are mirrored before they serve as full GEMM operands; the caller
// mirrors the final lower triangle.  n^3 flops in FP64 GEMMs (4n^3/3 with
// full updates); 128 x 128 base blocks by the symmetric sweep in one
// workgroup (SPD: no pivoting).  The Schur complements must stay positive
// definite: in FP32 their error, ~ eps cond(A11) |A21|^2, destroys that at
// the library's conditioning, in FP64 it is ~1e-9 of the smallest eigenvalue.
// ---------------------------------------------------------------------------
constexpr int INV_BASE = 128;

// One workgroup of 1024 threads, symmetric sweep operator on the (nb <= 128)
// block, read from the lower triangle and padded with the identity.  Thread t
// keeps the 4 x 4 block (rows 4 (t / 32) + r, columns 4 (t % 32) + c) in
// registers.  Sweeping pivot k keeps the matrix symmetric, so pivot row and
// pivot column are the same numbers: the threads holding column k publish it
// to LDS, and every element updates from it:
//   a_kk <- -1/a_kk,  a_ik <- a_ik/a_kk,  a_kj <- a_kj/a_kk,
//   a_ij <- a_ij - a_ik a_kj / a_kk.
// Sweeping all pivots leaves -A^-1; the full block is written back negated.
// Every element takes the rank-1 update, then the threads holding the pivot
// row or column overwrite those.  The published column alternates between
// two LDS buffers: a thread writing pivot k's column has passed barrier k-1,
// which every reader of pivot k-2 (the same buffer) reached only after its
// reads, so one barrier per pivot suffices.  The sweep is bound by
// instruction issue (64 FP64 FMAs per SIMD and pivot plus the per-wave
// reads, division and patches): four waves per SIMD overlap one another's
// latencies, where one wave of 8 x 8 blocks per SIMD waited on each.
// Pivots run in groups of 4 with the position inside the group unrolled, so
// every register index is static.
__global__ __launch_bounds__(1024) void inv_base_kernel(double* __restrict__ A, int64_t lda, int nb) {
  __shared__ __align__(16) double col[2][INV_BASE];
  const int t = threadIdx.x;
  const int bi = t >> 5, bj = t & 31;
  double a[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int i = 4 * bi + r, j = 4 * bj + c;
      double x = (i == j) ? 1.0 : 0.0;
      if (i < nb && j < nb) x = (i >= j) ? A[(int64_t)i * lda + j] : A[(int64_t)j * lda + i];
      a[r][c] = x;
    }
  for (int kb = 0; kb < INV_BASE / 4; ++kb) {
#pragma unroll
    for (int kc = 0; kc < 4; ++kc) {
      const int k = 4 * kb + kc;
      double* cb = col[kc & 1];
      if (bj == kb) {
#pragma unroll
        for (int r = 0; r < 4; ++r) cb[4 * bi + r] = a[r][kc];
      }
      __syncthreads();
      const double f = 1.0 / cb[k];
      double gr[4], gc[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) gr[r] = cb[4 * bi + r] * f;
#pragma unroll
      for (int c = 0; c < 4; ++c) gc[c] = cb[4 * bj + c];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) a[r][c] -= gr[r] * gc[c];
      if (bi == kb) {
#pragma unroll
        for (int c = 0; c < 4; ++c) a[kc][c] = gc[c] * f;
      }
      if (bj == kb) {
#pragma unroll
        for (int r = 0; r < 4; ++r) a[r][kc] = gr[r];
        if (bi == kb) a[kc][kc] = -f;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int i = 4 * bi + r, j = 4 * bj + c;
      if (i < nb && j < nb) A[(int64_t)i * lda + j] = -a[r][c];
    }
}

// Upper triangle of the n x n matrix <- its lower triangle, through 32 x 32
// LDS tiles (coalesced on both sides).
template <class T>
__global__ __launch_bounds__(256) void mirror_lower_kernel(T* __restrict__ A, int64_t n,
                                                           int64_t lda, int nt) {
  __shared__ T tile[32][33];
  // block b -> destination tile (bi, bj), bj >= bi, row-major over the upper triangle
  const int64_t b = blockIdx.x;
  int bi = (int)(((2.0 * nt + 1.0) - sqrt((2.0 * nt + 1.0) * (2.0 * nt + 1.0) - 8.0 * b)) * 0.5);
  while (bi > 0 && (int64_t)bi * (2 * nt - bi + 1) / 2 > b) --bi;
  while ((int64_t)(bi + 1) * (2 * nt - bi) / 2 <= b) ++bi;
  const int bj = (int)(b - (int64_t)bi * (2 * nt - bi + 1) / 2 + bi);
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  // source: tile (bj, bi) of the lower triangle
  for (int r = ty; r < 32; r += 8) {
    const int64_t gi = (int64_t)bj * 32 + r, gj = (int64_t)bi * 32 + tx;
    if (gi < n && gj < n) tile[r][tx] = A[gi * lda + gj];
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int64_t gi = (int64_t)bi * 32 + r, gj = (int64_t)bj * 32 + tx;
    if (gi < n && gj < n && gj > gi) A[gi * lda + gj] = tile[tx][r];
  }
}

template <class T>
static void mirror_lower(hipStream_t st, T* A, int64_t n, int64_t lda) {
  const int nt = (int)((n + 31) / 32);
  const unsigned tiles = (unsigned)((int64_t)nt * (nt + 1) / 2);
  hipLaunchKernelGGL(mirror_lower_kernel<T>, dim3(tiles), dim3(256), 0, st, A, n, lda, nt);
}

// ---------------------------------------------------------------------------
// The inverse's products are plain FP64 GEMMs, run by rocBLAS (its library
// kernels reach 0.83-0.92 of the FP64 MFMA peak at n = 5k-10k here, against
// about 0.5 for gemm_nt_kernel<double>).  rocBLAS is column-major: a row-major
// r x c array X with leading dimension ld is the column-major c x r matrix
// X^T, so every product below is written for those transposed views.  A
// product known to be symmetric updates only the row-major lower triangle,
// i.e. the column-major upper one (dsyrkx).  One handle per device, created
// on first use.
// ---------------------------------------------------------------------------
static rocblas_handle blas_handle(hipStream_t st) {
  static std::mutex mu;
  static rocblas_handle handles[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
  std::lock_guard<std::mutex> lk(mu);
  if (!handles[dev] && rocblas_create_handle(&handles[dev]) != rocblas_status_success) {
    handles[dev] = nullptr;
    return nullptr;
  }
  if (rocblas_set_stream(handles[dev], st) != rocblas_status_success) return nullptr;
  return handles[dev];
}

// column-major C (m x n) = alpha op(P) op(Q) + beta C
static int blas_dgemm(hipStream_t st, bool tP, bool tQ, int64_t m, int64_t n, int64_t k,
                      double alpha, const double* P, int64_t ldp, const double* Q, int64_t ldq,
                      double beta, double* C, int64_t ldc) {
  if (m <= 0 || n <= 0) return 0;
  rocblas_handle h = blas_handle(st);
  SMX_REQUIRE(h, "rocBLAS handle could not be created");
  const rocblas_status r = rocblas_dgemm(
      h, tP ? rocblas_operation_transpose : rocblas_operation_none,
      tQ ? rocblas_operation_transpose : rocblas_operation_none, (rocblas_int)m, (rocblas_int)n,
      (rocblas_int)k, &alpha, P, (rocblas_int)ldp, Q, (rocblas_int)ldq, &beta, C, (rocblas_int)ldc);
  SMX_REQUIRE(r == rocblas_status_success, "rocblas_dgemm failed (status %d)", (int)r);
  return 0;
}

// column-major upper triangle of C (n x n) = alpha op(P) op(Q)^T + beta C,
// op(P) op(Q)^T known to be symmetric
static int blas_dsyrkx(hipStream_t st, bool trans, int64_t n, int64_t k, double alpha,
                       const double* P, int64_t ldp, const double* Q, int64_t ldq, double beta,
                       double* C, int64_t ldc) {
  if (n <= 0) return 0;
  rocblas_handle h = blas_handle(st);
  SMX_REQUIRE(h, "rocBLAS handle could not be created");
  const rocblas_status r = rocblas_dsyrkx(
      h, rocblas_fill_upper, trans ? rocblas_operation_transpose : rocblas_operation_none,
      (rocblas_int)n, (rocblas_int)k, &alpha, P, (rocblas_int)ldp, Q, (rocblas_int)ldq, &beta, C,
      (rocblas_int)ldc);
  SMX_REQUIRE(r == rocblas_status_success, "rocblas_dsyrkx failed (status %d)", (int)r);
  return 0;
}

// work holds T (n2 x n1, row-major, leading dimension n1) per level; n^2 in
// total suffices.  In the column-major views (A21^T is n1 x n2, T^T n1 x n2,
// the diagonal blocks are symmetric):
//   T^T = X A21^T                      dgemm(N, N, n1, n2, n1)
//   S  -= T A21^T   (lower)            dsyrkx(T): (A21^T)^T T^T ... = A21 T^T
//   A21^T = -T^T S^-1                  dgemm(N, N, n1, n2, n2)
//   X  -= T^T A21   (lower)            dsyrkx(N): A21^T (T^T)^T = A21^T T
// Below knob(KN_INV_SYRKX_MIN) rows (default 16,384) the two symmetric
// updates run as full dgemm
// products instead: twice the flops, but one launch where rocBLAS syrkx runs
// a diagonal-block kernel plus several off-diagonal products.  Measured on
// the 20,000-row inverse, the full products are as fast as syrkx at the
// 10,000-row nodes and much faster below (2,500-row inverse 5.7 -> 2.9 ms);
// syrkx keeps the 20,000-row node, where its halved flops win.  The full
// product also rewrites the block's upper triangle, which nothing reads
// before mirror_lower restores it.
static int inv_spd_rec(hipStream_t st, double* A, int64_t n, int64_t lda, double* work) {
  if (n <= INV_BASE) {
    hipLaunchKernelGGL(inv_base_kernel, dim3(1), dim3(1024), 0, st, A, lda, (int)n);
    SMX_LAUNCHED();
    return 0;
  }
  const int64_t n1 = ((n / 2 + INV_BASE - 1) / INV_BASE) * INV_BASE;
  const int64_t n2 = n - n1;
  double* A11 = A;
  double* A21 = A + n1 * lda;
  double* A22 = A + n1 * lda + n1;
  double* T = work;                       // n2 x n1
  double* deeper = work + n1 * n2;
  if (int e = inv_spd_rec(st, A11, n1, lda, deeper)) return e;
  mirror_lower(st, A11, n1, lda);
  const bool sym = n >= knob(KN_INV_SYRKX_MIN);
  if (int e = blas_dgemm(st, false, false, n1, n2, n1, 1.0, A11, lda, A21, lda, 0.0, T, n1)) return e;
  if (int e = sym ? blas_dsyrkx(st, true, n2, n1, -1.0, A21, lda, T, n1, 1.0, A22, lda)
                  : blas_dgemm(st, true, false, n2, n2, n1, -1.0, A21, lda, T, n1, 1.0, A22, lda))
    return e;
  if (int e = inv_spd_rec(st, A22, n2, lda, deeper)) return e;
  mirror_lower(st, A22, n2, lda);
  if (int e = blas_dgemm(st, false, false, n1, n2, n2, -1.0, T, n1, A22, lda, 0.0, A21, lda)) return e;
  return sym ? blas_dsyrkx(st, false, n1, n2, -1.0, A21, lda, T, n1, 1.0, A11, lda)
             : blas_dgemm(st, false, true, n1, n1, n2, -1.0, A21, lda, T, n1, 1.0, A11, lda);
}

// Full inverse of the SPD matrix whose lower triangle is given.
static int spd_inverse64(hipStream_t st, double* A, int64_t n, int64_t lda, double* work) {
  if (int e = inv_spd_rec(st, A, n, lda, work)) return e;
  mirror_lower(st, A, n, lda);
  SMX_LAUNCHED();
  return 0;
}

// Q64 (lower triangle): Q += reg I, rho = tr(Q)/n, Q += rho I (nnls.py:115-121)

__global__ __launch_bounds__(1024) void regularise64_kernel(double* __restrict__ Q, int64_t n,
                                                            double reg, double* __restrict__ rho_out,
                                                            float* __restrict__ rho_f) {
  __shared__ double red[16];
  __shared__ double rho_s;
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    const double d = Q[i * n + i] + reg;
    Q[i * n + i] = d;
    acc += d;
  }
  for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
    rho_s = t / static_cast<double>(n);
    *rho_out = rho_s;
    *rho_f = static_cast<float>(rho_s);
  }
  __syncthreads();
  const double rho = rho_s;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) Q[i * n + i] = Q[i * n + i] + rho;
}

// Q64 (lower triangle) = U^T U + (lambda/n) I + rho I with rho = tr(U^T U +
// (lambda/n) I)/n (nnls.py:115-121), U[j, i] = lib[i, j] row_scale[j]; also
// the FP32 Ut for the right-hand sides.
static int form_gram64(hipStream_t st, const NnlsWs& w, const double* lib, int64_t ld_lib,
                       int64_t n, int64_t m, const double* row_scale, double lambda_reg) {
  const int64_t ldu = align_up(m, 16);
  hipLaunchKernelGGL(build_ut_split_kernel, dim3((unsigned)w.gpad), dim3(256), 0, st, lib, ld_lib,
                     n, m, w.mpad, row_scale, w.Uh, w.Ul, w.us);
  SMX_LAUNCHED();
  const int64_t nchk = w.mpad / AG_KP;
  const int64_t seg = (nchk + GR_FLUSH - 1) / GR_FLUSH;
  if (knob(KN_GRAM_WAVES) == 8)
    hipLaunchKernelGGL(gram_split_kernel<8>, dim3((unsigned)gram_grid(w.gpad)), dim3(GramShape<8>::NT),
                       0, st, n, w.gpad, w.mpad, w.Uh, w.Ul, w.us, w.Q64, seg, nchk);
  else
    hipLaunchKernelGGL(gram_split_kernel<4>, dim3((unsigned)gram_grid(w.gpad)), dim3(GramShape<4>::NT),
                       0, st, n, w.gpad, w.mpad, w.Uh, w.Ul, w.us, w.Q64, seg, nchk);
  SMX_LAUNCHED();
  hipLaunchKernelGGL(regularise64_kernel, dim3(1), dim3(1024), 0, st, w.Q64, n,
                     lambda_reg / static_cast<double>(n), w.rho64, w.rho);
  SMX_LAUNCHED();
  return 0;
}

}  // namespace smx

using namespace smx;

extern "C" {

int64_t smx_nnls_workspace_bytes(int64_t n, int64_t m, int64_t K) {
  if (n <= 0 || m <= 0 || K <= 0) return -1;
  return carve(nullptr, n, m, K).bytes;
}

int smx_nnls_admm(const double* lib, int64_t ld_lib, int64_t n, int64_t m, const double* row_scale,
                  const double* Y, int64_t K, double lambda_reg, int32_t maxiter, double w_init,
                  double tol, void* workspace, double* W, int32_t* iters_done, void* stream) {
  SMX_REQUIRE(lib && row_scale && Y && workspace && W, "smx_nnls_admm: null pointer");
  SMX_REQUIRE(n > 0 && m > 0 && K > 0 && ld_lib >= m && maxiter >= 0, "smx_nnls_admm: bad sizes");
  SMX_REQUIRE(n < (int64_t)1 << 31 && m < (int64_t)1 << 31, "smx_nnls_admm: sizes exceed int32");
  hipStream_t st = as_stream(stream);
  NnlsWs w = carve(workspace, n, m, K);
  const int64_t ldu = align_up(m, 16);
  stage_begin(ST_GRAM, st);
  // Q = U^T U (lower triangle, FP64 out of fp16-split MFMA), + lambda/n I,
  // rho = tr/n, + rho I
  if (int e = form_gram64(st, w, lib, ld_lib, n, m, row_scale, lambda_reg)) return e;
  stage_end(ST_GRAM, st, 3);
  // Q + rho I <- (Q + rho I)^-1 in place, FP64
  stage_begin(ST_FACTOR, st);
  if (int e = spd_inverse64(st, w.Q64, n, n, w.work64)) return e;
  stage_end(ST_FACTOR, st, 1);
  stage_begin(ST_RHS, st);
  // C = -Y U in FP64 (FP64 MFMA): U[j][i] = lib[i][j] row_scale[j], so
  // C[k][i] = -sum_j (Y[k][j] row_scale[j]) lib[i][j] (nnls.py:143)
  hipLaunchKernelGGL(scale_cols_kernel, dim3(blocks_for(K * ldu, 256)), dim3(256), 0, st, Y, K, m,
                     row_scale, ldu, w.Ys);
  SMX_LAUNCHED();
  if (int e = gemm_nt<double>(st, K, n, m, -1.0, w.Ys, ldu, lib, ld_lib, 0.0, w.C, w.nx, 0, w.work64,
                              n * n + n))
    return e;
  const int64_t tot = K * n;
  const int64_t npad = w.npad, Kp = w.Kp;
  const int S = w.S;
  hipLaunchKernelGGL(split_m_kernel, dim3(blocks_for(npad * npad / 8, 256)), dim3(256), 0, st, w.Q64, n,
                     npad, w.rho, w.Mh, w.Ml);
  SMX_LAUNCHED();
  for (float* buf : {w.Rp, w.Rb, w.D})
    hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(Kp * npad, 256)), dim3(256), 0, st,
                       reinterpret_cast<unsigned*>(buf), Kp * npad, 0u);
  hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(9 * Kp, 256)), dim3(256), 0, st, w.rmax,
                     9 * Kp, 0u);
  hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(256), 0, st, w.full_count, 1, 0u);
  SMX_LAUNCHED();
  stage_end(ST_RHS, st, 4);
  std::vector<int> act(K, 1);
  SMX_HIP(hipMemcpyAsync(w.active, act.data(), K * sizeof(int), hipMemcpyHostToDevice, st));
  const int check = 25;
  std::vector<unsigned> dl(K), zm(K);
  int done = 0;
  stage_begin(ST_ADMM, st);
  for (int it = 0; it < maxiter; ++it) {
    const bool chk = tol > 0.0 && ((it + 1) % check == 0);
    if (chk)
      hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(K, 256)), dim3(256), 0, st, w.delta, K, 0u);
    const AdmmIter ib = admm_iter(w.Rp, w.Rb, w.D, w.wb, w.Ra, w.rmax, w.full_count, Kp, it);
    // the first product has no base; the last is full so that the result
    // carries no delta error.  An anchor iteration replaces its product by
    // the FP64 one: the operand is zero and w = W_a exactly.
    const int anchor = (it % AG_ANCHOR == 0) ? 1 : 0;
    const int force = (anchor || it == maxiter - 1) ? 1 : 0;
    const int first = it == 0 ? 1 : 0;
    const float tau = g_admm_tau.load(std::memory_order_relaxed);
    if (anchor) {
      hipLaunchKernelGGL(admm_anchor_kernel, dim3(blocks_for(n, 256), (unsigned)K), dim3(256), 0, st,
                         K, n, Kp, first, w_init, w.rho64, w.x, w.C, w.Ra, w.Va, w.Rp);
      SMX_LAUNCHED();
      if (int e = gemm_nt<double>(st, K, n, n, 1.0, w.Va, w.nx, w.Q64, n, 0.0, w.Wc, w.nx, 0, w.work64,
                                  n * n + n))
        return e;
    } else {
      stage_begin(ST_ADMM_PRODUCT, st);
      launch_admm_product(st, K, Kp, npad, S, w.Rp, ib.rcur, w.D, ib.dcur, ib.fcur, force, tau,
                          w.Mh, w.Ml, w.part);
      SMX_LAUNCHED();
      stage_end(ST_ADMM_PRODUCT, st, 1);
    }
    hipLaunchKernelGGL(admm_update_kernel<4>, dim3(blocks_for(w.nx / 4, 256), (unsigned)K), dim3(256),
                       0, st, K, n, Kp, npad, S, w.part, w.rho64, w.Wc, w.x, first, force, anchor,
                       tau, w_init, ib, 1.6, w.active, w.delta);
    SMX_LAUNCHED();
    done = it + 1;
    if (chk) {
      hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(K, 256)), dim3(256), 0, st, w.zmax, K, 0u);
      hipLaunchKernelGGL(zmax_kernel, dim3(64, (unsigned)K), dim3(256), 0, st, (int)K, n, w.nx, w.x,
                         w.zmax);
      SMX_HIP(hipMemcpyAsync(dl.data(), w.delta, K * 4, hipMemcpyDeviceToHost, st));
      SMX_HIP(hipMemcpyAsync(zm.data(), w.zmax, K * 4, hipMemcpyDeviceToHost, st));
      SMX_HIP(hipStreamSynchronize(st));
      bool any = false;
      for (int64_t k = 0; k < K; ++k) {
        if (!act[k]) continue;
        float d, zz;
        std::memcpy(&d, &dl[k], 4);
        std::memcpy(&zz, &zm[k], 4);
        if (d <= tol * (zz > 0.f ? zz : 1.f)) act[k] = 0; else any = true;
      }
      SMX_HIP(hipMemcpyAsync(w.active, act.data(), K * sizeof(int), hipMemcpyHostToDevice, st));
      if (!any) break;
    }
  }
  stage_end(ST_ADMM, st, done);
  hipLaunchKernelGGL(copy_out_kernel, dim3(blocks_for(tot, 256)), dim3(256), 0, st, K, n, w.nx, w.x,
                     done > 0 ? 1 : 0, w_init, W);
  SMX_LAUNCHED();
  if (iters_done) *iters_done = done;
  return 0;
}

int64_t smx_nnls_group_state_bytes(int64_t n, int64_t m, int64_t K) {
  if (n <= 0 || m <= 0 || K <= 0) return -1;
  return carve_parts(nullptr, nullptr, n, m, K).state_bytes;
}

// The shared scratch of a grouped solve: the factor scratch of the largest
// model, then the device copy of the models' descriptors.
static int64_t group_desc_offset(int64_t n_max, int64_t m, int64_t K) {
  return carve_parts(nullptr, nullptr, n_max, m, K).scratch_bytes;
}

int64_t smx_nnls_group_scratch_bytes(int64_t n_max, int64_t m, int64_t K) {
  if (n_max <= 0 || m <= 0 || K <= 0) return -1;
  return group_desc_offset(n_max, m, K) + align_up(SMX_NNLS_GROUP_MAX * (int64_t)sizeof(AdmmModel), 256);
}

int smx_nnls_admm_group(int32_t n_models, const smx_nnls_problem* probs, int64_t m, int64_t K,
                        double lambda_reg, int32_t maxiter, void* scratch, void* stream) {
  const int G = n_models;
  SMX_REQUIRE(G >= 1 && G <= SMX_NNLS_GROUP_MAX, "smx_nnls_admm_group: n_models must be 1 .. %d",
              SMX_NNLS_GROUP_MAX);
  SMX_REQUIRE(probs && scratch, "smx_nnls_admm_group: null pointer");
  SMX_REQUIRE(m > 0 && K > 0 && maxiter >= 0 && m < (int64_t)1 << 31,
              "smx_nnls_admm_group: bad sizes");
  int64_t n_max = 0;
  for (int g = 0; g < G; ++g) {
    const smx_nnls_problem& p = probs[g];
    SMX_REQUIRE(p.lib && p.row_scale && p.Y && p.state && p.W,
                "smx_nnls_admm_group: null pointer in problem %d", g);
    SMX_REQUIRE(p.n > 0 && p.n < (int64_t)1 << 31 && p.ld_lib >= m,
                "smx_nnls_admm_group: bad sizes in problem %d", g);
    if (p.n > n_max) n_max = p.n;
  }
  hipStream_t st = as_stream(stream);
  std::vector<NnlsWs> ws(G);
  for (int g = 0; g < G; ++g) ws[g] = carve_parts(scratch, probs[g].state, probs[g].n, m, K);
  const int64_t Kp = ws[0].Kp;
  const int64_t ldu = align_up(m, 16);
  // per model, one after another on the shared scratch: Gram, FP64 inverse,
  // c = -Y U, the split planes of M, the loop's initial state
  for (int g = 0; g < G; ++g) {
    const smx_nnls_problem& p = probs[g];
    const NnlsWs& w = ws[g];
    const int64_t n = p.n, npad = w.npad;
    stage_begin(ST_GRAM, st);
    if (int e = form_gram64(st, w, p.lib, p.ld_lib, n, m, p.row_scale, lambda_reg)) return e;
    stage_end(ST_GRAM, st, 3);
    stage_begin(ST_FACTOR, st);
    if (int e = spd_inverse64(st, w.Q64, n, n, w.work64)) return e;
    stage_end(ST_FACTOR, st, 1);
    stage_begin(ST_RHS, st);
    hipLaunchKernelGGL(scale_cols_kernel, dim3(blocks_for(K * ldu, 256)), dim3(256), 0, st, p.Y, K,
                       m, p.row_scale, ldu, w.Ys);
    SMX_LAUNCHED();
    if (int e = gemm_nt<double>(st, K, n, m, -1.0, w.Ys, ldu, p.lib, p.ld_lib, 0.0, w.C, w.nx, 0,
                                w.work64, n * n + n))
      return e;
    hipLaunchKernelGGL(split_m_kernel, dim3(blocks_for(npad * npad / 8, 256)), dim3(256), 0, st,
                       w.Q64, n, npad, w.rho, w.Mh, w.Ml);
    for (float* buf : {w.Rp, w.Rb, w.D})
      hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(Kp * npad, 256)), dim3(256), 0, st,
                         reinterpret_cast<unsigned*>(buf), Kp * npad, 0u);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(9 * Kp, 256)), dim3(256), 0, st, w.rmax,
                       9 * Kp, 0u);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(256), 0, st, w.full_count, 1, 0u);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(K, 256)), dim3(256), 0, st,
                       reinterpret_cast<unsigned*>(w.active), K, 1u);
    SMX_LAUNCHED();
    stage_end(ST_RHS, st, 4);
  }
  // the models' descriptors and block runs
  // Product runs longest workgroups first (chunks per split, descending):
  // the launch's last round then holds the short workgroups of small models
  // rather than a tail of long ones.  The descriptors are stored in run order.
  std::vector<int> order(G);
  for (int g = 0; g < G; ++g) order[g] = g;
  auto cps_of = [&](int g) { return (ws[g].npad / AG_KP + ws[g].S - 1) / ws[g].S; };
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cps_of(a) > cps_of(b); });
  std::vector<AdmmModel> md(G);
  int64_t pb = 0, ub = 0;
  for (int q = 0; q < G; ++q) {
    const int g = order[q];
    const NnlsWs& w = ws[g];
    AdmmModel& d = md[q];
    d.n = probs[g].n;
    d.npad = w.npad;
    d.S = w.S;
    d.cps = static_cast<int>((w.npad / AG_KP + w.S - 1) / w.S);
    d.pblk0 = pb;
    d.pnblk = (w.npad / AG_BN) * w.S * (Kp / AG_KR);
    d.ublk0 = ub;
    d.unblk = blocks_for(w.nx / 4, 256);
    pb += align_up(d.pnblk, 8);
    ub += d.unblk;
    d.Rp = w.Rp;
    d.Rb = w.Rb;
    d.D = w.D;
    d.wb = w.wb;
    d.Ra = w.Ra;
    d.rmax = w.rmax;
    d.full_count = w.full_count;
    d.Mh = w.Mh;
    d.Ml = w.Ml;
    d.part = w.part;
    d.rho64 = w.rho64;
    d.Wc = w.Wc;
    d.x = w.x;
    d.active = w.active;
    d.delta = w.delta;
    d.w_init = probs[g].w_init;
  }
  SMX_REQUIRE(pb < ((int64_t)1 << 31) && ub < ((int64_t)1 << 31),
              "smx_nnls_admm_group: too many blocks for one launch");
  int64_t hi_bytes = 0;   // the hi planes one product launch streams
  for (int g = 0; g < G; ++g) hi_bytes += ws[g].npad * ws[g].npad * 2;
  AdmmModel* dmods =
      reinterpret_cast<AdmmModel*>(static_cast<char*>(scratch) + group_desc_offset(n_max, m, K));
  SMX_HIP(hipMemcpyAsync(dmods, md.data(), G * sizeof(AdmmModel), hipMemcpyHostToDevice, st));
  SMX_HIP(hipStreamSynchronize(st));   // md is pageable host memory
  stage_begin(ST_ADMM, st);
  for (int it = 0; it < maxiter; ++it) {
    const int anchor = (it % AG_ANCHOR == 0) ? 1 : 0;
    const int force = (anchor || it == maxiter - 1) ? 1 : 0;
    const int first = it == 0 ? 1 : 0;
    const float tau = g_admm_tau.load(std::memory_order_relaxed);
    if (anchor) {
      for (int g = 0; g < G; ++g) {
        const NnlsWs& w = ws[g];
        const int64_t n = probs[g].n;
        hipLaunchKernelGGL(admm_anchor_kernel, dim3(blocks_for(n, 256), (unsigned)K), dim3(256), 0,
                           st, K, n, Kp, first, probs[g].w_init, w.rho64, w.x, w.C, w.Ra, w.Va,
                           w.Rp);
        SMX_LAUNCHED();
        if (int e = gemm_nt<double>(st, K, n, n, 1.0, w.Va, w.nx, w.Q64, n, 0.0, w.Wc, w.nx, 0,
                                    w.work64, n * n + n))
          return e;
      }
    } else {
      stage_begin(ST_ADMM_PRODUCT, st);
      if (admm_nt_loads(Kp, hi_bytes))
        hipLaunchKernelGGL((admm_gemm_group_kernel<AG_KP, true>), dim3((unsigned)pb), dim3(AG_NT), 0,
                           st, dmods, G, K, Kp, it, force, tau);
      else
        hipLaunchKernelGGL((admm_gemm_group_kernel<AG_KP, false>), dim3((unsigned)pb), dim3(AG_NT),
                           0, st, dmods, G, K, Kp, it, force, tau);
      SMX_LAUNCHED();
      stage_end(ST_ADMM_PRODUCT, st, 1);
    }
    hipLaunchKernelGGL(admm_update_group_kernel<4>, dim3((unsigned)ub, (unsigned)K), dim3(256), 0, st,
                       dmods, G, K, Kp, it, first, force, anchor, tau, 1.6);
    SMX_LAUNCHED();
  }
  stage_end(ST_ADMM, st, maxiter);
  for (int g = 0; g < G; ++g) {
    const int64_t tot = K * probs[g].n;
    hipLaunchKernelGGL(copy_out_kernel, dim3(blocks_for(tot, 256)), dim3(256), 0, st, K, probs[g].n,
                       ws[g].nx, ws[g].x, maxiter > 0 ? 1 : 0, probs[g].w_init, probs[g].W);
  }
  SMX_LAUNCHED();
  return 0;
}

int smx_nnls_group_full_products(int64_t n, int64_t m, int64_t K, const void* state,
                                 int32_t* full_products, void* stream) {
  SMX_REQUIRE(state && full_products && n > 0 && m > 0 && K > 0,
              "smx_nnls_group_full_products: bad arguments");
  hipStream_t st = as_stream(stream);
  NnlsWs w = carve_parts(nullptr, const_cast<void*>(state), n, m, K);
  unsigned v = 0;
  SMX_HIP(hipMemcpyAsync(&v, w.full_count, 4, hipMemcpyDeviceToHost, st));
  SMX_HIP(hipStreamSynchronize(st));
  *full_products = static_cast<int32_t>(v);
  return 0;
}

}  // extern "C"
namespace smx {
// ---------------------------------------------------------------------------
// The ADMM operator alone: M = (U^T U + (lambda/n) I + rho I)^-1 built once,
// then applied to any K right-hand sides (K <= the K the workspace was sized
// for).  Used where many solves share one Gram up to low-rank changes.
// ---------------------------------------------------------------------------
// R (K, n) -> zero-padded (Kp, npad) in the product's blocked layout, and
// the row maxima of |R|.  Grid (npad / 256, Kp).
__global__ __launch_bounds__(256) void pad_rows_kernel(const float* __restrict__ R, int64_t K,
                                                       int64_t n, int64_t Kp,
                                                       float* __restrict__ Rp,
                                                       unsigned* __restrict__ rmax) {
  __shared__ float red[4];
  const int64_t k = blockIdx.y;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  float v = 0.f;
  if (k < K && i < n) v = R[k * n + i];
  Rp[r_tile_index(k, i, Kp)] = v;
  const float m = block_max_256(fabsf(v), red);
  if (threadIdx.x == 0 && m > 0.f) atomicMax(rmax + k, __float_as_uint(m));
}

__global__ void apply_out_kernel(const float* __restrict__ part, int S, int64_t Kp, int64_t npad,
                                 int64_t K, int64_t n, const float* __restrict__ rho_p,
                                 const unsigned* __restrict__ rmax, float* __restrict__ out) {
  const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= K * n) return;
  const int64_t k = idx / n, i = idx - k * n;
  float sum = 0.f;
  for (int s = 0; s < S; ++s) sum += part[((int64_t)s * Kp + k) * npad + i];
  const float sM = pow2_scale_for(1.f / *rho_p, 13);
  const float sR = pow2_scale_for(__uint_as_float(rmax[k]), 14);
  out[idx] = static_cast<float>(static_cast<double>(sum) /
                                (static_cast<double>(sM) * static_cast<double>(sR)));
}

static int nnls_build_operator(hipStream_t st, const NnlsWs& w, const double* lib, int64_t ld_lib,
                               int64_t n, int64_t m, const double* row_scale, double lambda_reg) {
  const int64_t ldu = align_up(m, 16);
  stage_begin(ST_GRAM, st);
  if (int e = form_gram64(st, w, lib, ld_lib, n, m, row_scale, lambda_reg)) return e;
  stage_end(ST_GRAM, st, 3);
  stage_begin(ST_FACTOR, st);
  if (int e = spd_inverse64(st, w.Q64, n, n, w.work64)) return e;
  hipLaunchKernelGGL(split_m_kernel, dim3(blocks_for(w.npad * w.npad / 8, 256)), dim3(256), 0, st,
                     w.Q64, n, w.npad, w.rho, w.Mh, w.Ml);
  SMX_LAUNCHED();
  stage_end(ST_FACTOR, st, 1);
  return 0;
}
}  // namespace smx
using namespace smx;
extern "C" {

int smx_nnls_admm_full_products(int64_t n, int64_t m, int64_t K, const void* workspace,
                                int32_t* full_products, void* stream) {
  SMX_REQUIRE(workspace && full_products && n > 0 && m > 0 && K > 0,
              "smx_nnls_admm_full_products: bad arguments");
  hipStream_t st = as_stream(stream);
  NnlsWs w = carve(const_cast<void*>(workspace), n, m, K);
  unsigned v = 0;
  SMX_HIP(hipMemcpyAsync(&v, w.full_count, 4, hipMemcpyDeviceToHost, st));
  SMX_HIP(hipStreamSynchronize(st));
  *full_products = static_cast<int32_t>(v);
  return 0;
}

int smx_admm_set_occupancy(int32_t workgroups_per_cu) {
  SMX_REQUIRE(workgroups_per_cu >= 0 && workgroups_per_cu <= AG_OCC_MAX,
              "smx_admm_set_occupancy: workgroups_per_cu must be 0 (default) .. %d", AG_OCC_MAX);
  return smx_tune_set(SMX_KNOB_ADMM_OCCUPANCY, workgroups_per_cu);
}

int smx_admm_set_delta_threshold(double tau) {
  SMX_REQUIRE(tau >= 0.0 && tau <= 0.25, "smx_admm_set_delta_threshold: tau must be in [0, 0.25]");
  g_admm_tau.store(static_cast<float>(tau), std::memory_order_relaxed);
  return 0;
}

int smx_admm_get_delta_threshold(double* tau) {
  SMX_REQUIRE(tau, "smx_admm_get_delta_threshold: null pointer");
  *tau = static_cast<double>(g_admm_tau.load(std::memory_order_relaxed));
  return 0;
}

int smx_admm_get_occupancy(int32_t* workgroups_per_cu) {
  SMX_REQUIRE(workgroups_per_cu, "smx_admm_get_occupancy: null pointer");
  *workgroups_per_cu = static_cast<int32_t>(knob(KN_ADMM_OCC));
  return 0;
}

int smx_admm_product_time(int64_t n, int64_t m, int64_t K, void* workspace, int32_t full,
                          int32_t reps, float* ms_per_launch, void* stream) {
  SMX_REQUIRE(workspace && ms_per_launch && n > 0 && m > 0 && K > 0 && reps > 0,
              "smx_admm_product_time: bad arguments");
  hipStream_t st = as_stream(stream);
  NnlsWs w = carve(workspace, n, m, K);
  const int64_t npad = w.npad, Kp = w.Kp;
  hipEvent_t e0, e1;
  SMX_HIP(hipEventCreate(&e0));
  SMX_HIP(hipEventCreate(&e1));
  SMX_HIP(hipEventRecord(e0, st));
  for (int r = 0; r < reps; ++r)
    launch_admm_product(st, K, Kp, npad, w.S, w.Rp, w.rmax, w.D, w.rmax + 3 * Kp, w.rmax,
                        full ? 1 : -1, AG_TAU,
                        w.Mh, w.Ml, w.part);
  SMX_LAUNCHED();
  SMX_HIP(hipEventRecord(e1, st));
  SMX_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  SMX_HIP(hipEventElapsedTime(&ms, e0, e1));
  SMX_HIP(hipEventDestroy(e0));
  SMX_HIP(hipEventDestroy(e1));
  *ms_per_launch = ms / static_cast<float>(reps);
  return 0;
}

int64_t smx_nnls_operator_bytes(int64_t n, int64_t m, int64_t K_max) {
  if (n <= 0 || m <= 0 || K_max <= 0) return -1;
  return carve(nullptr, n, m, K_max).bytes;
}

int smx_nnls_operator(const double* lib, int64_t ld_lib, int64_t n, int64_t m,
                      const double* row_scale, double lambda_reg, int64_t K_max, void* workspace,
                      double* rho_out, void* stream) {
  SMX_REQUIRE(lib && row_scale && workspace && rho_out, "smx_nnls_operator: null pointer");
  SMX_REQUIRE(n > 0 && m > 0 && K_max > 0 && ld_lib >= m && n < (int64_t)1 << 31 &&
                  m < (int64_t)1 << 31,
              "smx_nnls_operator: bad sizes");
  hipStream_t st = as_stream(stream);
  NnlsWs w = carve(workspace, n, m, K_max);
  if (int e = nnls_build_operator(st, w, lib, ld_lib, n, m, row_scale, lambda_reg)) return e;
  SMX_HIP(hipMemcpyAsync(rho_out, w.rho64, sizeof(double), hipMemcpyDeviceToDevice, st));
  return 0;
}

// out = R M (hi_only = 0: the full split product) or R_hi M_hi (hi_only = 1:
// the delta product's arithmetic, for R a small change of an earlier input).
static int operator_apply(int64_t n, int64_t m, int64_t K_max, void* workspace, const float* R,
                          int64_t K, float* out, int hi_only, void* stream) {
  SMX_REQUIRE(workspace && R && out, "smx_nnls_operator_apply: null pointer");
  SMX_REQUIRE(n > 0 && m > 0 && K > 0 && K <= K_max, "smx_nnls_operator_apply: bad sizes");
  hipStream_t st = as_stream(stream);
  NnlsWs w = carve(workspace, n, m, K_max);
  const int64_t npad = w.npad, Kp = align_up(K, AG_KR);
  const int S = w.S;
  hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(Kp, 256)), dim3(256), 0, st, w.rmax, Kp, 0u);
  hipLaunchKernelGGL(pad_rows_kernel, dim3((unsigned)(npad / 256), (unsigned)Kp), dim3(256), 0, st,
                     R, K, n, Kp, w.Rp, w.rmax);
  SMX_LAUNCHED();
  stage_begin(ST_ADMM_PRODUCT, st);
  // the padded rows serve as R (full) or as D (delta), with the same row maxima
  launch_admm_product(st, K, Kp, npad, S, w.Rp, w.rmax, w.Rp, w.rmax, w.rmax, hi_only ? -1 : 1,
                      AG_TAU,
                      w.Mh, w.Ml,
                      w.part);
  SMX_LAUNCHED();
  stage_end(ST_ADMM_PRODUCT, st, 1);
  hipLaunchKernelGGL(apply_out_kernel, dim3(blocks_for(K * n, 256)), dim3(256), 0, st, w.part, S,
                     Kp, npad, K, n, w.rho, w.rmax, out);
  SMX_LAUNCHED();
  return 0;
}

int smx_nnls_operator_apply(int64_t n, int64_t m, int64_t K_max, void* workspace, const float* R,
                            int64_t K, float* out, void* stream) {
  return operator_apply(n, m, K_max, workspace, R, K, out, 0, stream);
}

int smx_nnls_operator_apply_hi(int64_t n, int64_t m, int64_t K_max, void* workspace,
                               const float* D, int64_t K, float* out, void* stream) {
  return operator_apply(n, m, K_max, workspace, D, K, out, 1, stream);
}

int smx_gemm_nt_f32(int64_t M, int64_t N, int64_t K, float alpha, const float* A, int64_t lda,
                    const float* B, int64_t ldb, float beta, float* C, int64_t ldc, void* stream) {
  SMX_REQUIRE(A && B && C && M >= 0 && N >= 0 && K >= 0 && M < (1LL << 31) && N < (1LL << 31) &&
                  K < (1LL << 31),
              "smx_gemm_nt_f32: bad arguments");
  return gemm_nt<float>(as_stream(stream), M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
}

int smx_gemm_nt_f64(int64_t M, int64_t N, int64_t K, double alpha, const double* A, int64_t lda,
                    const double* B, int64_t ldb, double beta, double* C, int64_t ldc, void* stream) {
  SMX_REQUIRE(A && B && C && M >= 0 && N >= 0 && K >= 0 && M < (1LL << 31) && N < (1LL << 31) &&
                  K < (1LL << 31),
              "smx_gemm_nt_f64: bad arguments");
  return gemm_nt<double>(as_stream(stream), M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
}

int smx_gemm_nt_f64_split(int64_t M, int64_t N, int64_t K, double alpha, const double* A,
                          int64_t lda, const double* B, int64_t ldb, double* C, int64_t ldc,
                          double* work, int64_t work_elems, void* stream) {
  SMX_REQUIRE(A && B && C && work && M >= 0 && N >= 0 && K >= 0 && M < (1LL << 31) &&
                  N < (1LL << 31) && K < (1LL << 31) && lda >= K && ldb >= K && ldc >= N &&
                  work_elems >= 0,
              "smx_gemm_nt_f64_split: bad arguments");
  return gemm_nt<double>(as_stream(stream), M, N, K, alpha, A, lda, B, ldb, 0.0, C, ldc, 0, work,
                         work_elems);
}

int smx_spd_inverse_f64(double* A, int64_t n, double* work, void* stream) {
  SMX_REQUIRE(A && work && n > 0 && n < ((int64_t)1 << 31), "smx_spd_inverse_f64: bad arguments");  // work: n*n + n doubles
  return spd_inverse64(as_stream(stream), A, n, n, work);
}

int smx_model_maps(const double* lib, int64_t ld_lib, int64_t n, int64_t m, const double* W,
                   int64_t K, double* maps, void* stream) {
  SMX_REQUIRE(lib && W && maps && n > 0 && m > 0 && K > 0 && ld_lib >= m, "smx_model_maps: bad arguments");
  // row-major out (K, m) = W (K, n) lib[:, :m] is the column-major product
  // out^T (m, K) = lib[:, :m]^T (m, n, leading dimension ld_lib) W^T (n, K):
  // one FP64 GEMM on the matrix cores
  hipStream_t st = as_stream(stream);
  stage_begin(ST_MAPS, st);
  if (int e = blas_dgemm(st, false, false, m, K, n, 1.0, lib, ld_lib, W, n, 0.0, maps, m)) return e;
  stage_end(ST_MAPS, st, 1);
  return 0;
}

}  // extern "C"
