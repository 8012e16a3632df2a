// NNLS by ADMM with a shared factor (nnls.py:6-166) and model maps
// (model.py:14-68).
//
// Device pipeline for K right-hand sides:
//   U^T (n, m)        FP32, from the FP64 library columns and per-row scales
//   Q = U^T U         FP32 GEMM (n, n); + lambda/n I; rho = tr(Q)/n
//   L L^T = Q + rho I blocked right-looking Cholesky (FP32)
//   Z = L^-1          blocked triangular inverse
//   M = Z^T Z         = (Q + rho I)^-1, so one ADMM iteration is ONE GEMM
//   C = -Y U          right-hand-side terms (K, n)
//   per iteration:    W = R M  (R = rho (z - u) - c), epilogue does the
//                     relaxation / projection / dual update of nnls.py:152-157
//                     and writes the next R.
// z and u are kept in FP64; the products run in FP32 with FP32 accumulation.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include <vector>
#include <cstring>
#include <algorithm>

#include "common.hpp"
#include "../../include/smx.h"

namespace smx {

// ---------------------------------------------------------------------------
// generic tiled SGEMM:  C = alpha * op(A) op(B) + beta * C
//   op(A) is (M, K): A[i*lda + k] (TA = false) or A[k*lda + i] (TA = true)
//   op(B) is (K, N): B[k*ldb + j] (TB = false) or B[j*ldb + k] (TB = true)
// kmode: 0 full k range; 1 lower-triangular A (k <= i); 2 k >= max(i, j)
//        (product of two lower-triangular operands Z^T Z).
// lower_only: skip tiles strictly above the diagonal.
// ---------------------------------------------------------------------------
constexpr int BM = 64, BN = 64, BK = 16;

template <bool TA, bool TB>
__global__ __launch_bounds__(256) void sgemm_kernel(int M, int N, int K, float alpha,
                                                    const float* __restrict__ A, int64_t lda,
                                                    const float* __restrict__ B, int64_t ldb,
                                                    float beta, float* __restrict__ C, int64_t ldc,
                                                    int kmode, int lower_only) {
  const int i0 = blockIdx.y * BM, j0 = blockIdx.x * BN;
  if (lower_only && j0 > i0 + BM - 1) return;
  __shared__ float As[BK][BM + 4];
  __shared__ float Bs[BK][BN + 4];
  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = tid >> 4;
  float acc[4][4] = {};
  int kbeg = 0, kend = K;
  if (kmode == 1) kend = min(K, i0 + BM);
  if (kmode == 2) kbeg = (max(i0, j0) / BK) * BK;
  for (int k0 = kbeg; k0 < kend; k0 += BK) {
    // A tile: BM x BK
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      const int e = tid + 256 * l;
      int ii, kk;
      if (TA) { ii = e & (BM - 1); kk = e >> 6; } else { kk = e & (BK - 1); ii = e >> 4; }
      const int gi = i0 + ii, gk = k0 + kk;
      float v = 0.f;
      if (gi < M && gk < kend) v = TA ? A[(int64_t)gk * lda + gi] : A[(int64_t)gi * lda + gk];
      As[kk][ii] = v;
    }
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      const int e = tid + 256 * l;
      int jj, kk;
      if (TB) { kk = e & (BK - 1); jj = e >> 4; } else { jj = e & (BN - 1); kk = e >> 6; }
      const int gj = j0 + jj, gk = k0 + kk;
      float v = 0.f;
      if (gj < N && gk < kend) v = TB ? B[(int64_t)gj * ldb + gk] : B[(int64_t)gk * ldb + gj];
      Bs[kk][jj] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; ++kk) {
      float a[4], b[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) a[r] = As[kk][ty * 4 + r];
#pragma unroll
      for (int c = 0; c < 4; ++c) b[c] = Bs[kk][tx * 4 + c];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] += a[r] * b[c];
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int gi = i0 + ty * 4 + r;
    if (gi >= M) continue;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int gj = j0 + tx * 4 + c;
      if (gj >= N) continue;
      if (lower_only && gj > gi) continue;
      float* p = C + (int64_t)gi * ldc + gj;
      *p = (beta == 0.f) ? alpha * acc[r][c] : alpha * acc[r][c] + beta * (*p);
    }
  }
}

template <bool TA, bool TB>
static int sgemm(hipStream_t st, int M, int N, int K, float alpha, const float* A, int64_t lda,
                 const float* B, int64_t ldb, float beta, float* C, int64_t ldc, int kmode = 0,
                 int lower_only = 0) {
  if (M <= 0 || N <= 0) return 0;
  dim3 grid((N + BN - 1) / BN, (M + BM - 1) / BM);
  hipLaunchKernelGGL((sgemm_kernel<TA, TB>), grid, dim3(256), 0, st, M, N, K, alpha, A, lda, B,
                     ldb, beta, C, ldc, kmode, lower_only);
  SMX_LAUNCHED();
  return 0;
}

// ---------------------------------------------------------------------------
// design matrix, Gram regularisation, trace
// ---------------------------------------------------------------------------
// Ut[i, j] = lib[i, j] * scale[j] for j < m, 0 for m <= j < ldu
__global__ void build_ut_kernel(const double* __restrict__ lib, int64_t ld_lib, int64_t n,
                                int64_t m, int64_t ldu, const double* __restrict__ scale,
                                float* __restrict__ Ut) {
  const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= n * ldu) return;
  const int64_t i = idx / ldu, j = idx - i * ldu;
  Ut[idx] = (j < m) ? static_cast<float>(lib[i * ld_lib + j] * scale[j]) : 0.f;
}

__global__ void to_float_kernel(const double* __restrict__ src, int64_t rows, int64_t cols,
                                int64_t ld_dst, float* __restrict__ dst) {
  const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= rows * ld_dst) return;
  const int64_t i = idx / ld_dst, j = idx - i * ld_dst;
  dst[idx] = (j < cols) ? static_cast<float>(src[i * cols + j]) : 0.f;
}

// Q += reg I; rho = tr(Q)/n; Q += rho I.  One block.
__global__ __launch_bounds__(1024) void regularise_kernel(float* __restrict__ Q, int64_t n,
                                                          float reg, float* __restrict__ rho_out) {
  __shared__ double red[16];
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    const float d = Q[i * n + i] + reg;
    Q[i * n + i] = d;
    acc += d;
  }
  for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  __shared__ float rho_s;
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
    rho_s = static_cast<float>(t) / static_cast<float>(n);
    *rho_out = rho_s;
  }
  __syncthreads();
  const float rho = rho_s;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) Q[i * n + i] = Q[i * n + i] + rho;
}

// ---------------------------------------------------------------------------
// blocked Cholesky (lower, in place, row-major), block NB
// ---------------------------------------------------------------------------
constexpr int NB = 64;

// factor the jb x jb diagonal block at (k, k); one block of NB*4 threads
__global__ __launch_bounds__(256) void potrf_diag_kernel(float* __restrict__ A, int64_t lda, int64_t k,
                                                         int jb) {
  __shared__ float T[NB][NB + 1];
  for (int e = threadIdx.x; e < jb * jb; e += blockDim.x) {
    const int r = e / jb, c = e - r * jb;
    T[r][c] = A[(k + r) * lda + k + c];
  }
  __syncthreads();
  for (int j = 0; j < jb; ++j) {
    if (threadIdx.x == 0) T[j][j] = sqrtf(T[j][j]);
    __syncthreads();
    const float d = T[j][j];
    for (int r = j + 1 + threadIdx.x; r < jb; r += blockDim.x) T[r][j] = T[r][j] / d;
    __syncthreads();
    // trailing update of the lower triangle
    for (int e = threadIdx.x; e < (jb - j - 1) * (jb - j - 1); e += blockDim.x) {
      const int r = j + 1 + e / (jb - j - 1), c = j + 1 + e % (jb - j - 1);
      if (c <= r) T[r][c] -= T[r][j] * T[c][j];
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < jb * jb; e += blockDim.x) {
    const int r = e / jb, c = e - r * jb;
    A[(k + r) * lda + k + c] = (c <= r) ? T[r][c] : 0.f;
  }
}

// rows i in [k+jb, n): L21[i,:] = A21[i,:] L11^-T (forward substitution)
__global__ __launch_bounds__(256) void trsm_panel_kernel(float* __restrict__ A, int64_t lda, int64_t n,
                                                         int64_t k, int jb) {
  __shared__ float L[NB][NB + 1];
  for (int e = threadIdx.x; e < jb * jb; e += blockDim.x) {
    const int r = e / jb, c = e - r * jb;
    L[r][c] = A[(k + r) * lda + k + c];
  }
  __syncthreads();
  const int64_t i = k + jb + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  float x[NB];
  float* row = A + i * lda + k;
  for (int j = 0; j < jb; ++j) {
    float s = row[j];
    for (int l = 0; l < j; ++l) s -= x[l] * L[j][l];
    x[j] = s / L[j][j];
  }
  for (int j = 0; j < jb; ++j) row[j] = x[j];
}

// zero the strict upper triangle of an n x n row-major matrix
__global__ void zero_upper_kernel(float* __restrict__ A, int64_t n) {
  const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= n * n) return;
  const int64_t i = idx / n, j = idx - i * n;
  if (j > i) A[idx] = 0.f;
}

// invert the lower-triangular jb x jb diagonal block at (j, j) in place
__global__ __launch_bounds__(64) void trti_diag_kernel(float* __restrict__ A, int64_t lda, int64_t j0,
                                                       int jb) {
  __shared__ float L[NB][NB + 1];
  __shared__ float X[NB][NB + 1];
  for (int e = threadIdx.x; e < jb * jb; e += blockDim.x) {
    const int r = e / jb, c = e - r * jb;
    L[r][c] = A[(j0 + r) * lda + j0 + c];
  }
  __syncthreads();
  // column c of X solves L x = e_c; thread c owns column c
  const int c = threadIdx.x;
  if (c < jb) {
    for (int r = 0; r < jb; ++r) {
      float s = (r == c) ? 1.f : 0.f;
      for (int l = c; l < r; ++l) s -= L[r][l] * X[l][c];
      X[r][c] = (r < c) ? 0.f : s / L[r][r];
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < jb * jb; e += blockDim.x) {
    const int r = e / jb, cc = e - r * jb;
    A[(j0 + r) * lda + j0 + cc] = X[r][cc];
  }
}

// ---------------------------------------------------------------------------
// ADMM
// ---------------------------------------------------------------------------
// state init: z = w_init, u = 0, R = rho (z - u) - c
__global__ void admm_init_kernel(int64_t total, double w_init, const float* __restrict__ rho_p,
                                 const float* __restrict__ c, double* __restrict__ z,
                                 double* __restrict__ u, float* __restrict__ R) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= total) return;
  const double rho = *rho_p;
  z[i] = w_init;
  u[i] = 0.0;
  R[i] = static_cast<float>(rho * (w_init - 0.0) - static_cast<double>(c[i]));
}

// Wt = R M (R: (K, n), M symmetric (n, n)) with the ADMM update fused in the
// epilogue (nnls.py:152-157).  delta[k] receives max |z_new - z| bits.
__global__ __launch_bounds__(256) void admm_step_kernel(int K, int n, const float* __restrict__ R,
                                                        const float* __restrict__ M,
                                                        const float* __restrict__ rho_p,
                                                        const float* __restrict__ c,
                                                        double* __restrict__ z,
                                                        double* __restrict__ u,
                                                        float* __restrict__ Rn, double alpha,
                                                        const int* __restrict__ active,
                                                        unsigned* __restrict__ delta) {
  const int i0 = blockIdx.y * BM, j0 = blockIdx.x * BN;
  __shared__ float As[BK][BM + 4];
  __shared__ float Bs[BK][BN + 4];
  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = tid >> 4;
  float acc[4][4] = {};
  for (int k0 = 0; k0 < n; k0 += BK) {
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      const int e = tid + 256 * l;
      const int kk = e & (BK - 1), ii = e >> 4;
      const int gi = i0 + ii, gk = k0 + kk;
      As[kk][ii] = (gi < K && gk < n) ? R[(int64_t)gi * n + gk] : 0.f;
    }
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      const int e = tid + 256 * l;
      const int kk = e & (BK - 1), jj = e >> 4;
      const int gj = j0 + jj, gk = k0 + kk;
      Bs[kk][jj] = (gj < n && gk < n) ? M[(int64_t)gj * n + gk] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; ++kk) {
      float a[4], b[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) a[r] = As[kk][ty * 4 + r];
#pragma unroll
      for (int cc = 0; cc < 4; ++cc) b[cc] = Bs[kk][tx * 4 + cc];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int cc = 0; cc < 4; ++cc) acc[r][cc] += a[r] * b[cc];
    }
    __syncthreads();
  }
  const double rho = *rho_p;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int gi = i0 + ty * 4 + r;
    if (gi >= K) continue;
    const bool act = active[gi] != 0;
    float dmax = 0.f;
#pragma unroll
    for (int cc = 0; cc < 4; ++cc) {
      const int gj = j0 + tx * 4 + cc;
      if (gj >= n) continue;
      const int64_t o = (int64_t)gi * n + gj;
      if (act) {
        const double w = acc[r][cc];
        const double zo = z[o], uo = u[o];
        const double wh = alpha * w + (1.0 - alpha) * zo;
        const double zn = fmax(0.0, wh + uo);
        const double un = (uo + wh) - zn;
        z[o] = zn;
        u[o] = un;
        Rn[o] = static_cast<float>(rho * (zn - un) - static_cast<double>(c[o]));
        dmax = fmaxf(dmax, static_cast<float>(fabs(zn - zo)));
      } else {
        Rn[o] = R[o];
      }
    }
    if (act && dmax > 0.f) atomicMax(delta + gi, __float_as_uint(dmax));
  }
}

__global__ void zmax_kernel(int K, int64_t n, const double* __restrict__ z, unsigned* __restrict__ zmax) {
  const int k = blockIdx.y;
  float m = 0.f;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    m = fmaxf(m, static_cast<float>(fabs(z[(int64_t)k * n + i])));
  if (m > 0.f) atomicMax(zmax + k, __float_as_uint(m));
}

__global__ void copy_out_kernel(int64_t total, const double* __restrict__ z, double* __restrict__ W) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < total) W[i] = z[i];
}

// ---------------------------------------------------------------------------
// FP64 maps: out (K, m) = W (K, n) . lib[:, :m]
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void maps_kernel(int K, int64_t n, int64_t m, const double* __restrict__ W,
                                                   const double* __restrict__ lib, int64_t ld,
                                                   double* __restrict__ out) {
  constexpr int TM = 16, TN = 64, TK = 32;
  __shared__ double Ws[TK][TM + 1];
  __shared__ double Ls[TK][TN + 1];
  const int k0 = blockIdx.y * TM;
  const int64_t j0 = blockIdx.x * (int64_t)TN;
  const int tid = threadIdx.x;
  const int tj = tid & 63, tk = tid >> 6;   // 4 row groups of 4
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t i0 = 0; i0 < n; i0 += TK) {
    for (int e = tid; e < TK * TM; e += 256) {
      const int ii = e % TK, kk = e / TK;
      const int64_t gi = i0 + ii;
      Ws[ii][kk] = (gi < n && k0 + kk < K) ? W[(int64_t)(k0 + kk) * n + gi] : 0.0;
    }
    for (int e = tid; e < TK * TN; e += 256) {
      const int jj = e % TN, ii = e / TN;
      const int64_t gi = i0 + ii, gj = j0 + jj;
      Ls[ii][jj] = (gi < n && gj < m) ? lib[gi * ld + gj] : 0.0;
    }
    __syncthreads();
#pragma unroll 8
    for (int ii = 0; ii < TK; ++ii) {
      const double l = Ls[ii][tj];
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[r] += Ws[ii][tk * 4 + r] * l;
    }
    __syncthreads();
  }
  const int64_t gj = j0 + tj;
  if (gj >= m) return;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int k = k0 + tk * 4 + r;
    if (k < K) out[(int64_t)k * m + gj] = acc[r];
  }
}

static int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

struct NnlsWs {
  float* Ut;      // n x ldu
  float* Q;       // n x n (factor / inverse work)
  float* Mi;      // n x n inverse
  float* tmp;     // n x NB
  float* Yf;      // K x ldu
  float* C;       // K x n
  float* R0;      // K x n
  float* R1;      // K x n
  double* z;      // K x n
  double* u;      // K x n
  float* rho;     // 1
  unsigned* delta;  // K
  unsigned* zmax;   // K
  int* active;      // K
  int64_t bytes;
};

static NnlsWs carve(void* base, int64_t n, int64_t m, int64_t K) {
  NnlsWs w{};
  const int64_t ldu = align_up(m, 16);
  char* p = static_cast<char*>(base);
  int64_t off = 0;
  auto take = [&](int64_t bytes) {
    char* q = p ? p + off : nullptr;
    off += align_up(bytes, 256);
    return q;
  };
  w.Ut = reinterpret_cast<float*>(take(n * ldu * 4));
  w.Q = reinterpret_cast<float*>(take(n * n * 4));
  w.Mi = reinterpret_cast<float*>(take(n * n * 4));
  w.tmp = reinterpret_cast<float*>(take(n * NB * 4));
  w.Yf = reinterpret_cast<float*>(take(K * ldu * 4));
  w.C = reinterpret_cast<float*>(take(K * n * 4));
  w.R0 = reinterpret_cast<float*>(take(K * n * 4));
  w.R1 = reinterpret_cast<float*>(take(K * n * 4));
  w.z = reinterpret_cast<double*>(take(K * n * 8));
  w.u = reinterpret_cast<double*>(take(K * n * 8));
  w.rho = reinterpret_cast<float*>(take(4));
  w.delta = reinterpret_cast<unsigned*>(take(K * 4));
  w.zmax = reinterpret_cast<unsigned*>(take(K * 4));
  w.active = reinterpret_cast<int*>(take(K * 4));
  w.bytes = off;
  return w;
}

__global__ void fill_u32_kernel(unsigned* p, int64_t n, unsigned v) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// in-place Cholesky of the n x n matrix A, then A <- L^-1 (lower), then
// Mi <- (L L^T)^-1
static int factor_and_invert(hipStream_t st, float* A, float* Mi, float* tmp, int64_t n) {
  for (int64_t k = 0; k < n; k += NB) {
    const int jb = static_cast<int>(std::min<int64_t>(NB, n - k));
    hipLaunchKernelGGL(potrf_diag_kernel, dim3(1), dim3(256), 0, st, A, n, k, jb);
    SMX_LAUNCHED();
    const int64_t rest = n - k - jb;
    if (rest > 0) {
      hipLaunchKernelGGL(trsm_panel_kernel, dim3(blocks_for(rest, 256)), dim3(256), 0, st, A, n, n,
                         k, jb);
      SMX_LAUNCHED();
      // A22 -= L21 L21^T (lower part)
      if (int e = sgemm<false, true>(st, (int)rest, (int)rest, jb, -1.f, A + (k + jb) * n + k, n,
                                     A + (k + jb) * n + k, n, 1.f, A + (k + jb) * n + k + jb, n, 0, 1))
        return e;
    }
  }
  hipLaunchKernelGGL(zero_upper_kernel, dim3(blocks_for(n * n, 256)), dim3(256), 0, st, A, n);
  SMX_LAUNCHED();
  // triangular inverse, last block first (LAPACK trtri, lower)
  const int64_t nblk = (n + NB - 1) / NB;
  for (int64_t b = nblk - 1; b >= 0; --b) {
    const int64_t j = b * NB;
    const int jb = static_cast<int>(std::min<int64_t>(NB, n - j));
    hipLaunchKernelGGL(trti_diag_kernel, dim3(1), dim3(64), 0, st, A, n, j, jb);
    SMX_LAUNCHED();
    const int64_t rest = n - j - jb;
    if (rest > 0) {
      // tmp = Z22 * L21   (Z22 lower-triangular, already inverted)
      if (int e = sgemm<false, false>(st, (int)rest, jb, (int)rest, 1.f, A + (j + jb) * n + j + jb, n,
                                      A + (j + jb) * n + j, n, 0.f, tmp, NB, 1, 0))
        return e;
      // Z21 = -tmp * Z11
      if (int e = sgemm<false, false>(st, (int)rest, jb, jb, -1.f, tmp, NB, A + j * n + j, n, 0.f,
                                      A + (j + jb) * n + j, n, 0, 0))
        return e;
    }
  }
  // Mi = Z^T Z, using k >= max(i, j)
  return sgemm<true, false>(st, (int)n, (int)n, (int)n, 1.f, A, n, A, n, 0.f, Mi, n, 2, 0);
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
  hipLaunchKernelGGL(build_ut_kernel, dim3(blocks_for(n * ldu, 256)), dim3(256), 0, st, lib, ld_lib,
                     n, m, ldu, row_scale, w.Ut);
  SMX_LAUNCHED();
  // Q = U^T U (lower only), mirrored later by Z^T Z
  if (int e = sgemm<false, true>(st, (int)n, (int)n, (int)ldu, 1.f, w.Ut, ldu, w.Ut, ldu, 0.f, w.Q, n,
                                 0, 1))
    return e;
  hipLaunchKernelGGL(regularise_kernel, dim3(1), dim3(1024), 0, st, w.Q, n,
                     static_cast<float>(lambda_reg / static_cast<double>(n)), w.rho);
  SMX_LAUNCHED();
  if (int e = factor_and_invert(st, w.Q, w.Mi, w.tmp, n)) return e;
  // C = -Y U
  hipLaunchKernelGGL(to_float_kernel, dim3(blocks_for(K * ldu, 256)), dim3(256), 0, st, Y, K, m, ldu,
                     w.Yf);
  SMX_LAUNCHED();
  if (int e = sgemm<false, true>(st, (int)K, (int)n, (int)ldu, -1.f, w.Yf, ldu, w.Ut, ldu, 0.f, w.C,
                                 n, 0, 0))
    return e;
  const int64_t tot = K * n;
  hipLaunchKernelGGL(admm_init_kernel, dim3(blocks_for(tot, 256)), dim3(256), 0, st, tot, w_init,
                     w.rho, w.C, w.z, w.u, w.R0);
  SMX_LAUNCHED();
  std::vector<int> act(K, 1);
  SMX_HIP(hipMemcpyAsync(w.active, act.data(), K * sizeof(int), hipMemcpyHostToDevice, st));
  const int check = 25;
  std::vector<unsigned> dl(K), zm(K);
  float* Rc = w.R0;
  float* Rn = w.R1;
  int done = 0;
  dim3 grid((unsigned)((n + BN - 1) / BN), (unsigned)((K + BM - 1) / BM));
  for (int it = 0; it < maxiter; ++it) {
    const bool chk = tol > 0.0 && ((it + 1) % check == 0);
    if (chk) {
      hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(K, 256)), dim3(256), 0, st, w.delta, K, 0u);
    }
    hipLaunchKernelGGL(admm_step_kernel, grid, dim3(256), 0, st, (int)K, (int)n, Rc, w.Mi, w.rho,
                       w.C, w.z, w.u, Rn, 1.6, w.active, w.delta);
    SMX_LAUNCHED();
    std::swap(Rc, Rn);
    done = it + 1;
    if (chk) {
      hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(K, 256)), dim3(256), 0, st, w.zmax, K, 0u);
      hipLaunchKernelGGL(zmax_kernel, dim3(64, (unsigned)K), dim3(256), 0, st, (int)K, n, w.z, w.zmax);
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
  hipLaunchKernelGGL(copy_out_kernel, dim3(blocks_for(tot, 256)), dim3(256), 0, st, tot, w.z, W);
  SMX_LAUNCHED();
  if (iters_done) *iters_done = done;
  return 0;
}

int smx_model_maps(const double* lib, int64_t ld_lib, int64_t n, int64_t m, const double* W,
                   int64_t K, double* maps, void* stream) {
  SMX_REQUIRE(lib && W && maps && n > 0 && m > 0 && K > 0 && ld_lib >= m, "smx_model_maps: bad arguments");
  dim3 grid((unsigned)((m + 63) / 64), (unsigned)((K + 15) / 16));
  hipLaunchKernelGGL(maps_kernel, grid, dim3(256), 0, as_stream(stream), (int)K, n, m, W, lib, ld_lib,
                     maps);
  SMX_LAUNCHED();
  return 0;
}

}  // extern "C"
