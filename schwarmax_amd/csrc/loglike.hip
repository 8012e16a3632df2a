// Model maps to Gauss-Hermite moments, and the bootstrap chi^2 with its
// log-mean-exp, for the forward evaluation (no autograd).
//
//  * model_epilogue_kernel: from the maps W lib[:, :m] (model.py:14-66) the
//    surface density (map_xy * mmpo) * lm (model.py:260-286), the clamped
//    h1..h4 = clip(map_h / gamma, -10, 10) and the LOSVD mean and dispersion
//    (ghmoments.py:53-79).
//  * bootstrap_rows_kernel + bootstrap_total_kernel: the rows of
//    likelihood.py:184-244.  Every row k adds the same density term, the sum
//    over all rows and cells of ((model - y_R) / (0.1 y_R amp))^2 (the
//    reference broadcasts its res_Rzphi), to its own surface-density and
//    h1..h4 terms; NaN terms are skipped (nansum); logL = lmax +
//    log(mean(exp(logl - lmax))).  One workgroup per row, fixed reduction
//    order, so a value does not depend on the launch.
#include <hip/hip_runtime.h>
#include <cmath>

#include "common.hpp"
#include "../../include/smx.h"

namespace smx {

constexpr double kGhEps = 1e-12;   // constants.py:11 (EPSILON)

__device__ __forceinline__ double clip10(double x) {
  return x < -10.0 ? -10.0 : (x > 10.0 ? 10.0 : x);   // NaN passes, as torch.clamp
}

__global__ __launch_bounds__(256) void model_epilogue_kernel(
    const double* __restrict__ maps, int64_t ldm, int64_t K, int P, int B,
    const double* __restrict__ gamma, const double* __restrict__ v0, const double* __restrict__ s,
    double mmpo, double lm, double* __restrict__ d2, double* __restrict__ h, double* __restrict__ V,
    double* __restrict__ sig) {
#pragma clang fp contract(off)
  const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= K * B) return;
  const int64_t k = idx / B, a = idx - k * B;
  const double gk = gamma[a];
  const double g = fabs(gk) > 1e-8 ? gk : 1e-8;
  d2[idx] = (maps[k * ldm + P + a] * mmpo) * lm;
  const double* row = maps + k * ldm + P + B;
  double hv[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    hv[j] = clip10(row[(int64_t)j * B + a] / g);
    h[((int64_t)j * K + k) * B + a] = hv[j];
  }
  // ghmoments.py:53-79
  const double sq2 = 1.4142135623730951, sq3 = 1.7320508075688772, sq6 = 2.449489742783178;
  const double dv = sq2 * hv[0] + sq3 * hv[2];
  V[idx] = v0[a] + s[a] * dv;
  const double s2 = ((1.0 + (2.0 * sq2) * hv[1]) + (2.0 * sq6) * hv[3]) - dv * dv;
  sig[idx] = s[a] * sqrt(s2 < 1e-10 ? 1e-10 : s2);
}

__device__ __forceinline__ double sq_term(double m, double y, double sg) {
#pragma clang fp contract(off)
  const double q = (m - y) / sg;
  const double t = q * q;
  return t != t ? 0.0 : t;   // nansum
}

__device__ __forceinline__ double block_sum_256(double v, double* red) {
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  const double t = ((red[0] + red[1]) + red[2]) + red[3];
  __syncthreads();
  return t;
}

// Row k: out_R[k] = its density term, out_S[k] = its surface-density and
// h1..h4 terms, as (((h1 + h2) + h3) + h4) + xy.
__global__ __launch_bounds__(256) void bootstrap_rows_kernel(smx_bootstrap_model m, int64_t K,
                                                             int P, int B, smx_bootstrap_data d,
                                                             double amp, double* __restrict__ out_R,
                                                             double* __restrict__ out_S) {
#pragma clang fp contract(off)
  __shared__ double red[4];
  const int64_t k = blockIdx.x;
  const double* d3 = m.d3 + k * m.ld3;
  double r = 0.0;
  for (int64_t p = threadIdx.x; p < P; p += 256)
    r += sq_term(d3[p], m.y_R[p], (0.1 * m.y_R[p]) * amp);
  double t[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int64_t a = threadIdx.x; a < B; a += 256) {
    const double sxr = d.xy_err[a] + kGhEps;
    const double yb = d.xy[a] + d.xy_normal[k * B + a] * sxr;
    t[4] += sq_term(m.surface_density[k * B + a], yb, (0.1 * d.xy[a]) * amp);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const double sr = d.h_err[j][a] + kGhEps;
      const double yh = d.h[j][a] + d.h_normal[j][k * B + a] * sr;
      t[j] += sq_term(m.h[j][k * B + a], yh, sr * amp);
    }
  }
  const double R = block_sum_256(r, red);
  double S[5];
#pragma unroll
  for (int j = 0; j < 5; ++j) S[j] = block_sum_256(t[j], red);
  if (threadIdx.x == 0) {
    out_R[k] = R;
    out_S[k] = (((S[0] + S[1]) + S[2]) + S[3]) + S[4];
  }
}

// logl_k = -0.5 (S_k + sum_k' R_k'), then lmax + log(mean(exp(logl - lmax))).
__global__ __launch_bounds__(256) void bootstrap_total_kernel(int64_t K, const double* __restrict__ R,
                                                              const double* __restrict__ S,
                                                              double* __restrict__ logl_rows,
                                                              double* __restrict__ logl) {
#pragma clang fp contract(off)
  __shared__ double red[4];
  __shared__ double res_R_s, lmax_s;
  if (threadIdx.x == 0) {
    double acc = 0.0;
    for (int64_t k = 0; k < K; ++k) acc += R[k];
    res_R_s = acc;
  }
  __syncthreads();
  double lm = -INFINITY;
  for (int64_t k = threadIdx.x; k < K; k += 256) {
    const double v = -0.5 * (S[k] + res_R_s);
    logl_rows[k] = v;
    lm = fmax(lm, v);
  }
  for (int o = 32; o >= 1; o >>= 1) lm = fmax(lm, __shfl_xor(lm, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = lm;
  __syncthreads();
  if (threadIdx.x == 0) lmax_s = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  __syncthreads();
  double e = 0.0;
  for (int64_t k = threadIdx.x; k < K; k += 256) e += exp(logl_rows[k] - lmax_s);
  const double tot = block_sum_256(e, red);
  if (threadIdx.x == 0) *logl = lmax_s + log(tot / static_cast<double>(K));
}

}  // namespace smx

using namespace smx;

extern "C" {

int smx_model_epilogue(const double* maps, int64_t ldm, int64_t K, int32_t P, int32_t B,
                       const double* gamma_kin, const double* v0, const double* s, double mmpo,
                       double lm, double* surface_density, double* h, double* V, double* sigma,
                       void* stream) {
  SMX_REQUIRE(maps && gamma_kin && v0 && s && surface_density && h && V && sigma,
              "smx_model_epilogue: null pointer");
  SMX_REQUIRE(K >= 0 && P >= 0 && B >= 0 && ldm >= (int64_t)P + 5 * (int64_t)B,
              "smx_model_epilogue: bad sizes");
  if (K * B == 0) return 0;
  hipLaunchKernelGGL(model_epilogue_kernel, dim3(blocks_for(K * B, 256)), dim3(256), 0,
                     as_stream(stream), maps, ldm, K, (int)P, (int)B, gamma_kin, v0, s, mmpo, lm,
                     surface_density, h, V, sigma);
  SMX_LAUNCHED();
  return 0;
}

int smx_bootstrap_loglike(const smx_bootstrap_model* model, int64_t K, int32_t P, int32_t B,
                          const smx_bootstrap_data* data, double sigma_amplify, double* work,
                          double* logl_rows, double* logl, void* stream) {
  SMX_REQUIRE(model && data && work && logl_rows && logl, "smx_bootstrap_loglike: null pointer");
  SMX_REQUIRE(model->d3 && model->y_R && model->surface_density,
              "smx_bootstrap_loglike: null model map");
  SMX_REQUIRE(data->xy && data->xy_err && data->xy_normal, "smx_bootstrap_loglike: null data");
  for (int j = 0; j < 4; ++j)
    SMX_REQUIRE(model->h[j] && data->h[j] && data->h_err[j] && data->h_normal[j],
                "smx_bootstrap_loglike: null h%d array", j + 1);
  SMX_REQUIRE(K >= 1 && P >= 0 && B >= 0 && model->ld3 >= P, "smx_bootstrap_loglike: bad sizes");
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(bootstrap_rows_kernel, dim3((unsigned)K), dim3(256), 0, st, *model, K, (int)P,
                     (int)B, *data, sigma_amplify, work, work + K);
  SMX_LAUNCHED();
  hipLaunchKernelGGL(bootstrap_total_kernel, dim3(1), dim3(256), 0, st, K, work, work + K, logl_rows,
                     logl);
  SMX_LAUNCHED();
  return 0;
}

}  // extern "C"
