// The jackknife's ADMM step after the operator products (model.py:490-577,
// whose per-replicate re-solve runs nnls.py:51-66): all replicates at once,
// one workgroup per replicate g.  With x_g = (Q_g + rho_g I)^-1 applied to
// the unmasked problem's operator (the Neumann-series products, done by the
// caller), the rank-d Woodbury correction for the d design rows V_g that
// replicate g drops and the ADMM update of nnls.py:155-162:
//   s = V_g x_g,  t = C_g^-1 s,  w = x_g + P_g^T t,
//   w_hat = alpha w + (1 - alpha) z,  z' = max(w_hat + u, 0),
//   u' = u + w_hat - z',  r' = rho_g (z' - u') - c_g,
// r' being the next iteration's right-hand side (nnls.py:152-154).
// The d dot products over n run in passes of JK_DC, each one pass over x_g
// (partial sums per thread, then across the wave by shuffles and across the
// four waves through LDS, in a fixed order); t is a d x d product in LDS; the
// update is one more pass over the row.  All arithmetic in FP64.
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

__global__ __launch_bounds__(JK_T) void jk_admm_step_kernel(
    int64_t n, int d, const double* __restrict__ x, const double* __restrict__ V,
    const double* __restrict__ P, const double* __restrict__ Cinv, const double* __restrict__ rho_g,
    const double* __restrict__ c, double alpha, double* __restrict__ z, double* __restrict__ u,
    double* __restrict__ r) {
  extern __shared__ double sh[];   // s[d], t[d], then per-wave partials [4][JK_DC]
  double* s = sh;
  double* t = sh + d;
  double* part = sh + 2 * d;
  const int64_t g = blockIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const double* xg = x + g * n;
  const double* Vg = V + g * (int64_t)d * n;
  const double* Pg = P + g * (int64_t)d * n;
  for (int d0 = 0; d0 < d; d0 += JK_DC) {
    const int nd = d - d0 < JK_DC ? d - d0 : JK_DC;
    double acc[JK_DC];
#pragma unroll
    for (int e = 0; e < JK_DC; ++e) acc[e] = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += JK_T) {
      const double xi = xg[i];
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
  double* rg = r + g * n;
  const double* cg = c + g * n;
  for (int64_t i = threadIdx.x; i < n; i += JK_T) {
    double w = xg[i];
    for (int e = 0; e < d; ++e) w += Pg[(int64_t)e * n + i] * t[e];
    const double zo = zg[i], uo = ug[i];
    const double w_hat = alpha * w + (1.0 - alpha) * zo;
    const double zn = fmax(w_hat + uo, 0.0);
    const double un = uo + w_hat - zn;
    zg[i] = zn;
    ug[i] = un;
    rg[i] = rho * (zn - un) - cg[i];
  }
}

}  // namespace smx

using namespace smx;

extern "C" {

int smx_jackknife_admm_step(int64_t G, int64_t n, int64_t d, const double* x, const double* V,
                            const double* P, const double* Cinv, const double* rho_g,
                            const double* c, double alpha, double* z, double* u, double* r,
                            void* stream) {
  SMX_REQUIRE(x && V && P && Cinv && rho_g && c && z && u && r && G > 0 && G < (1LL << 31) &&
                  n > 0 && d > 0 && d <= JK_DMAX,
              "smx_jackknife_admm_step: bad arguments (G %lld, n %lld, d %lld; d <= %d)",
              (long long)G, (long long)n, (long long)d, JK_DMAX);
  const size_t lds = (2 * (size_t)d + 4 * JK_DC) * sizeof(double);
  hipLaunchKernelGGL(jk_admm_step_kernel, dim3((unsigned)G), dim3(JK_T), lds, as_stream(stream), n,
                     (int)d, x, V, P, Cinv, rho_g, c, alpha, z, u, r);
  SMX_LAUNCHED();
  return 0;
}

}  // extern "C"
