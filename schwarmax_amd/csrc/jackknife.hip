// The jackknife's ADMM step after the operator products (model.py:490-577,
// whose per-replicate re-solve runs nnls.py:51-66): all replicates at once,
// one workgroup per replicate g.  The caller's products give the T terms
// S_k = r_g M^(k+1) of the Neumann series for the shifted operator, so
// x_g = sum_k delta_g^k S_k = (Q + rho I - delta_g I)^-1 r_g; then the rank-d
// Woodbury correction for the d design rows V_g that replicate g drops and
// the ADMM update of nnls.py:155-162:
//   s = V_g x_g,  t = C_g^-1 s,  w = x_g + P_g^T t,
//   w_hat = alpha w + (1 - alpha) z,  z' = max(w_hat + u, 0),
//   u' = u + w_hat - z',  r' = rho_g (z' - u') - c_g,
// r' being the next iteration's right-hand side (nnls.py:152-154), stored in
// FP32, the products' input precision.
// x is formed in FP64 where it is used (twice: the dot products and the
// update), not stored.  The d dot products over n run in passes of JK_DC, each
// one pass over the row (partial sums per thread, then across the wave by
// shuffles and across the four waves through LDS, in a fixed order); t is a
// d x d product in LDS; the update is one more pass over the row.
#include <hip/hip_runtime.h>
#include <cstdint>

#include "common.hpp"
#include "../../include/smx.h"

namespace smx {

constexpr int JK_T = 256;         // threads per workgroup (4 waves)
constexpr int JK_DC = 8;          // dot products per pass over x_g
constexpr int JK_DMAX = 2048;     // dropped rows per replicate at most (LDS: 2 d doubles)

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// x_g[i] = sum_k delta^k S_k[g][i], in FP64, the powers by repeated products.
__device__ __forceinline__ double jk_x(const float* __restrict__ Sg, int64_t stride, int T,
                                       double delta, int64_t i) {
  double x = static_cast<double>(Sg[i]);
  double p = 1.0;
  for (int k = 1; k < T; ++k) {
    p *= delta;
    x += p * static_cast<double>(Sg[k * stride + i]);
  }
  return x;
}

__global__ __launch_bounds__(JK_T) void jk_admm_step_kernel(
    int64_t G, int64_t n, int d, int T, const float* __restrict__ S, const double* __restrict__ delta,
    const double* __restrict__ V, const double* __restrict__ P, const double* __restrict__ Cinv,
    const double* __restrict__ rho_g, const double* __restrict__ c, double alpha,
    double* __restrict__ z, double* __restrict__ u, float* __restrict__ r) {
  extern __shared__ double sh[];   // s[d], t[d], then per-wave partials [4][JK_DC]
  double* s = sh;
  double* t = sh + d;
  double* part = sh + 2 * d;
  const int64_t g = blockIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const float* Sg = S + g * n;       // term k of replicate g at Sg + k G n
  const int64_t stride = G * n;
  const double dg = delta[g];
  const double* Vg = V + g * (int64_t)d * n;
  const double* Pg = P + g * (int64_t)d * n;
  for (int d0 = 0; d0 < d; d0 += JK_DC) {
    const int nd = d - d0 < JK_DC ? d - d0 : JK_DC;
    double acc[JK_DC];
#pragma unroll
    for (int e = 0; e < JK_DC; ++e) acc[e] = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += JK_T) {
      const double xi = jk_x(Sg, stride, T, dg, i);
#pragma unroll
      for (int e = 0; e < JK_DC; ++e)
        if (e < nd) acc[e] += Vg[(int64_t)(d0 + e) * n + i] * xi;
    }
#pragma unroll
    for (int e = 0; e < JK_DC; ++e) {
      const double v = wave_sum(acc[e]);
      if (lane == 0) part[wave * JK_DC + e] = v;
    }
    __syncthreads();
    if (threadIdx.x < nd)
      s[d0 + threadIdx.x] = ((part[threadIdx.x] + part[JK_DC + threadIdx.x]) +
                             part[2 * JK_DC + threadIdx.x]) + part[3 * JK_DC + threadIdx.x];
    __syncthreads();
  }
  const double* Cg = Cinv + g * (int64_t)d * d;
  for (int e = threadIdx.x; e < d; e += JK_T) {
    double a = 0.0;
    for (int f = 0; f < d; ++f) a += Cg[(int64_t)e * d + f] * s[f];
    t[e] = a;
  }
  __syncthreads();
  const double rho = rho_g[g];
  double* zg = z + g * n;
  double* ug = u + g * n;
  float* rg = r + g * n;
  const double* cg = c + g * n;
  for (int64_t i = threadIdx.x; i < n; i += JK_T) {
    double w = jk_x(Sg, stride, T, dg, i);
    for (int e = 0; e < d; ++e) w += Pg[(int64_t)e * n + i] * t[e];
    const double zo = zg[i], uo = ug[i];
    const double w_hat = alpha * w + (1.0 - alpha) * zo;
    const double zn = fmax(w_hat + uo, 0.0);
    const double un = uo + w_hat - zn;
    zg[i] = zn;
    ug[i] = un;
    rg[i] = static_cast<float>(rho * (zn - un) - cg[i]);
  }
}

}  // namespace smx

using namespace smx;

extern "C" {

int smx_jackknife_admm_step(int64_t G, int64_t n, int64_t d, int64_t T, const float* S,
                            const double* delta, const double* V, const double* P,
                            const double* Cinv, const double* rho_g, const double* c, double alpha,
                            double* z, double* u, float* r, void* stream) {
  SMX_REQUIRE(S && delta && V && P && Cinv && rho_g && c && z && u && r && G > 0 &&
                  G < (1LL << 31) && n > 0 && d > 0 && d <= JK_DMAX && T > 0 && T <= 64,
              "smx_jackknife_admm_step: bad arguments (G %lld, n %lld, d %lld, T %lld; d <= %d, "
              "T <= 64)",
              (long long)G, (long long)n, (long long)d, (long long)T, JK_DMAX);
  const size_t lds = (2 * (size_t)d + 4 * JK_DC) * sizeof(double);
  hipLaunchKernelGGL(jk_admm_step_kernel, dim3((unsigned)G), dim3(JK_T), lds, as_stream(stream), G,
                     n, (int)d, (int)T, S, delta, V, P, Cinv, rho_g, c, alpha, z, u, r);
  SMX_LAUNCHED();
  return 0;
}

}  // extern "C"
