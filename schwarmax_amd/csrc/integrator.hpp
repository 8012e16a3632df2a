// BS3 attempt in the rotating bar frame, the Jacobi energy, and the
// per-integration accumulator slab with the binning of one sample.
#pragma once
#include "common.hpp"
#include "potential.hpp"
#include "binning.hpp"

namespace smx {

// ---------------------------------------------------------------------------
// BS3 integrator (integrates.py:8-112)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void deriv(const smx_potential& p, double Om, const double* y, double* k) {
  double ax, ay, az;
  accel_fast(p, y[0], y[1], y[2], ax, ay, az);
  k[0] = y[3] + Om * y[1];
  k[1] = y[4] - Om * y[0];
  k[2] = y[5];
  k[3] = ax + Om * y[4];
  k[4] = ay - Om * y[3];
  k[5] = az;
}

__device__ __forceinline__ double jacobi(const smx_potential& p, double Om, const double* y) {
  const double ek = 0.5 * ((y[3] * y[3] + y[4] * y[4]) + y[5] * y[5]);
  return (ek + phi_total(p, y[0], y[1], y[2])) - Om * (y[0] * y[4] - y[1] * y[3]);
}

// One attempt.  On return y holds y_out, k1 the FSAL slope, dt the next step;
// returns the step length used when accepted, 0 otherwise.  The stage
// vectors are folded as soon as they are complete (the error sum's first
// three terms before k4 is formed), keeping the reference's summation order.
__device__ __forceinline__ double bs3_attempt(const smx_potential& p, const smx_schedule& s,
                                              double& t, double* y, double* k1, double& dt) {
  const double Om = s.Omega;
  double yt[6], k2[6], kk[6], yn[6], e3[6];
#pragma unroll
  for (int j = 0; j < 6; ++j) yt[j] = y[j] + (0.5 * dt) * k1[j];
  deriv(p, Om, yt, k2);
#pragma unroll
  for (int j = 0; j < 6; ++j) yt[j] = y[j] + (0.75 * dt) * k2[j];
  deriv(p, Om, yt, kk);                                  // k3
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    yn[j] = y[j] + dt * ((2.0 / 9.0 * k1[j] + 1.0 / 3.0 * k2[j]) + 4.0 / 9.0 * kk[j]);
    e3[j] = (-5.0 / 72.0 * k1[j] + 1.0 / 12.0 * k2[j]) + 1.0 / 9.0 * kk[j];
  }
  deriv(p, Om, yn, kk);                                  // k4
  double es = 0.0;
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    const double e = dt * (e3[j] + -1.0 / 8.0 * kk[j]);
    const double sc = s.atol + s.rtol * nmax(fabs(y[j]), fabs(yn[j]));
    const double q = e / sc;
    es += q * q;
  }
  const double err = sqrt(es / 6.0);
  const bool accept = err <= 1.0;
  const double used = accept ? dt : 0.0;
  if (accept) {
    t = t + dt;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      y[j] = yn[j];
      k1[j] = kk[j];
    }
  }
  double sf = 0.9 * pow(nmax(err, 1e-10), -1.0 / 3.0);
  sf = nmin(nmax(sf, 0.2), 5.0);
  dt = nmin(nmax(dt * sf, s.dt_min), s.dt_max);
  return used;
}

// The same attempt with the stage slopes k2, k3, k4 returned (k4 = f(y_new)
// whether or not the attempt is accepted); used to replay a segment for the
// adjoint.
__device__ __forceinline__ double bs3_attempt_stages(const smx_potential& p, const smx_schedule& s,
                                                     double& t, double* y, double* k1, double& dt,
                                                     double* k2, double* k3, double* k4) {
  const double Om = s.Omega;
  double yt[6], yn[6], e3[6];
#pragma unroll
  for (int j = 0; j < 6; ++j) yt[j] = y[j] + (0.5 * dt) * k1[j];
  deriv(p, Om, yt, k2);
#pragma unroll
  for (int j = 0; j < 6; ++j) yt[j] = y[j] + (0.75 * dt) * k2[j];
  deriv(p, Om, yt, k3);
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    yn[j] = y[j] + dt * ((2.0 / 9.0 * k1[j] + 1.0 / 3.0 * k2[j]) + 4.0 / 9.0 * k3[j]);
    e3[j] = (-5.0 / 72.0 * k1[j] + 1.0 / 12.0 * k2[j]) + 1.0 / 9.0 * k3[j];
  }
  deriv(p, Om, yn, k4);
  double es = 0.0;
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    const double e = dt * (e3[j] + -1.0 / 8.0 * k4[j]);
    const double sc = s.atol + s.rtol * nmax(fabs(y[j]), fabs(yn[j]));
    const double q = e / sc;
    es += q * q;
  }
  const double err = sqrt(es / 6.0);
  const bool accept = err <= 1.0;
  const double used = accept ? dt : 0.0;
  if (accept) {
    t = t + dt;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      y[j] = yn[j];
      k1[j] = k4[j];
    }
  }
  double sf = 0.9 * pow(nmax(err, 1e-10), -1.0 / 3.0);
  sf = nmin(nmax(sf, 0.2), 5.0);
  dt = nmin(nmax(dt * sf, s.dt_min), s.dt_max);
  return used;
}

// per-integration accumulator slab: P cells (padded to 8) then B records of
// 8 doubles [dt, g dt H0, ..., g dt H4, -, -]
__host__ __device__ __forceinline__ int64_t slab_rz_len(int P) { return (P + 7) & ~7; }
__host__ __device__ __forceinline__ int64_t slab_stride(int P, int B) {
  return slab_rz_len(P) + 8 * (int64_t)B;
}

__device__ __forceinline__ void bin_sample(const smx_grids& g, const int32_t* __restrict__ bm,
                                           const double* __restrict__ v0,
                                           const double* __restrict__ sd, const double* y,
                                           double w, int nsym, double* __restrict__ rz,
                                           double* __restrict__ ap, int P) {
  SampleProj sp;
  project(g, y, nsym > 2, sp);
  for (int c = 0; c < nsym; ++c) {
    double zs, phi, xr, yr, vlos;
    copy_coords(sp, c, zs, phi, xr, yr, vlos);
    const int cell = rzphi_cell(g, sp.R, zs * y[2], phi);
    if (cell >= 0) rz[cell] += w;
    const int a = aperture(g, bm, xr, yr);
    if (a >= 0) {
      double gw, H[5];
      gh_terms(vlos, v0[a], sd[a], gw, H);
      double* rec = ap + 8 * (int64_t)a;
      const double gd = gw * w;
      rec[0] += w;
#pragma unroll
      for (int k = 0; k < 5; ++k) rec[1 + k] += gd * H[k];
    }
  }
}

}  // namespace smx
