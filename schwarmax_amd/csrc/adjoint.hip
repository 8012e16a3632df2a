// Reverse-mode derivative of the model evaluation up to the orbit library,
// the derivative jax.grad takes of the reference (model.py:70-258 through
// integrates.py:37-300): bin membership, accept / reject decisions and
// validity flags held fixed, everything else differentiated, including the
// step-size controller (integrates.py:100-111), the Gauss-Hermite weights
// (ghmoments.py:6-36), the T normalisation and the realization average.
//
//  * library_vjp_kernel: one integration per lane.  The forward pass stored
//    the state (y, dt, t) before every ckpt_every-th attempt; each segment is
//    replayed from its checkpoint, storing the attempts' stage slopes, and
//    then swept backwards.  Output: cotangents of the integration's initial
//    state and dt0, of Omega, of the sky rotation's third row and of the
//    potential fields.
//  * The force's vector-Jacobian product uses the symmetry of the Hessian of
//    Phi (J^T w = J w, one directional derivative) and forward-mode tangents
//    for the component parameters, each component with its own dual size.
//  * initial_conditions_jac / density_target_jac: forward-mode Jacobians of
//    the Jeans ICs, time scales and density target with respect to the
//    potential fields (dual numbers seeded a few fields per pass).
#include <hip/hip_runtime.h>
#include <cmath>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "potential.hpp"
#include "binning.hpp"
#include "integrator.hpp"
#include "ics.hpp"
#include "timing.hpp"
#include "../../include/smx.h"

namespace smx {

static_assert(PF_COUNT == SMX_POT_FIELDS, "field count mismatch");

// ---------------------------------------------------------------------------
// Force VJP: k = deriv(y) (integrates.py:8-27).  For cotangent w of k, add
// (dk/dy)^T w to yb, dk/dOmega . w to omb and dk/dfield . w to fb.
// ---------------------------------------------------------------------------
template <int N>
__device__ __forceinline__ void seed_xyz(const double* y, const double* dir, Dual<N>& X, Dual<N>& Y,
                                         Dual<N>& Z) {
  X = dconst<N>(y[0]);
  Y = dconst<N>(y[1]);
  Z = dconst<N>(y[2]);
  X.d[0] = dir[0];
  Y.d[0] = dir[1];
  Z.d[0] = dir[2];
}

template <int N>
__device__ __forceinline__ void take_component(double sgn, const double* wv, const Dual<N>& gx,
                                               const Dual<N>& gy, const Dual<N>& gz, double* yb,
                                               double* fb, const int* slot_field) {
  // acceleration of this component = sgn * grad; Hessian symmetric
  yb[0] += sgn * gx.d[0];
  yb[1] += sgn * gy.d[0];
  yb[2] += sgn * gz.d[0];
#pragma unroll
  for (int k = 1; k < N; ++k)
    fb[slot_field[k]] += sgn * ((wv[0] * gx.d[k] + wv[1] * gy.d[k]) + wv[2] * gz.d[k]);
}

__device__ __noinline__ void deriv_vjp(const smx_potential& p, double Om, const double* y,
                                       const double* w, double* yb, double& omb, double* fb) {
  // rotating-frame terms: k0 = y3 + Om y1, k1 = y4 - Om y0, k2 = y5,
  // k3 = ax + Om y4, k4 = ay - Om y3, k5 = az
  yb[0] += -Om * w[1];
  yb[1] += Om * w[0];
  yb[3] += w[0] - Om * w[4];
  yb[4] += w[1] + Om * w[3];
  yb[5] += w[2];
  omb += (w[0] * y[1] - w[1] * y[0]) + (w[3] * y[4] - w[4] * y[3]);
  const double wv[3] = {w[3], w[4], w[5]};
  {  // NFW: slots [direction, mGM, Rs]; a = -grad
    typedef Dual<3> D;
    D X, Y, Z, gx, gy, gz;
    seed_xyz(y, wv, X, Y, Z);
    const D m = dseed<3>(p.nfw_mGM, 1), rs = dseed<3>(p.nfw_Rs, 2);
    nfw_grad_t(p, m, rs, X, Y, Z, gx, gy, gz);
    const int sf[3] = {0, PF_NFW_MGM, PF_NFW_RS};
    take_component(-1.0, wv, gx, gy, gz, yb, fb, sf);
  }
  {  // Miyamoto-Nagai: [direction, mGM, Rs, Hs]; a = -grad
    typedef Dual<4> D;
    D X, Y, Z, gx, gy, gz;
    seed_xyz(y, wv, X, Y, Z);
    const D m = dseed<4>(p.mn_mGM, 1), rs = dseed<4>(p.mn_Rs, 2), hs = dseed<4>(p.mn_Hs, 3);
    mn_grad_t(p, m, rs, hs, X, Y, Z, gx, gy, gz);
    const int sf[4] = {0, PF_MN_MGM, PF_MN_RS, PF_MN_HS};
    take_component(-1.0, wv, gx, gy, gz, yb, fb, sf);
  }
  {  // T3 bar: [direction, M, L, a, b]; a = +G grad(ph)
    typedef Dual<5> D;
    D X, Y, Z, gx, gy, gz;
    seed_xyz(y, wv, X, Y, Z);
    const D m = dseed<5>(p.t3_M, 1), L = dseed<5>(p.t3_L, 2), a = dseed<5>(p.t3_a, 3),
            b = dseed<5>(p.t3_b, 4);
    t3_grad_t(p, m, L, a, b, X, Y, Z, gx, gy, gz);
    const int sf[5] = {0, PF_T3_M, PF_T3_L, PF_T3_A, PF_T3_B};
    take_component(p.G, wv, gx, gy, gz, yb, fb, sf);
  }
  {  // V4 bulge: [direction, M]; a = +G grad(ph)
    typedef Dual<2> D;
    D X, Y, Z, gx, gy, gz;
    seed_xyz(y, wv, X, Y, Z);
    const D m = dseed<2>(p.v4_M, 1);
    v4_grad_t(p, m, X, Y, Z, gx, gy, gz);
    const int sf[2] = {0, PF_V4_M};
    take_component(p.G, wv, gx, gy, gz, yb, fb, sf);
  }
}

// ---------------------------------------------------------------------------
// Binning VJP of one accepted sample y with weight w (bin_sample's adjoint).
// gr is the library cotangent row of the sample's column.  Returns dL/dw
// (unscaled); adds sc * dL/dy to ynb (velocities only: positions enter
// through bin membership alone) and sc * dL/d(rot[6..8]) to rotb.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double bin_vjp(const smx_grids& g, const int32_t* __restrict__ bm,
                          const double* __restrict__ v0, const double* __restrict__ sd,
                          const double* __restrict__ npb, const double* __restrict__ gr,
                          const double* y, double w, double sc, int nsym, int P, double* ynb,
                          double* rotb) {
  const int B = g.num_vbin;
  SampleProj sp;
  project(g, y, nsym > 2, sp);
  double dw = 0.0;
  for (int c = 0; c < nsym; ++c) {
    double zs, phi, xr, yr, vlos;
    copy_coords(sp, c, zs, phi, xr, yr, vlos);
    const int cell = rzphi_cell(g, sp.R, zs * y[2], phi);
    if (cell >= 0) dw += gr[cell];
    const int a = aperture(g, bm, xr, yr);
    if (a >= 0) {
      double gw, H[5];
      gh_terms(vlos, v0[a], sd[a], gw, H);
      // library rows [Rzphi | xy | h1 | h2 | h3 | h4 | h0]
      const double gh[5] = {gr[P + 5 * B + a], gr[P + B + a], gr[P + 2 * B + a],
                            gr[P + 3 * B + a], gr[P + 4 * B + a]};
      double sH = 0.0, sdH = 0.0;
      const double u = (vlos - v0[a]) / sd[a];
      // d/du (g H_k) = g (H_k' - u H_k), H_k' = sqrt(2k) H_{k-1}
      const double dH[5] = {0.0, kSq2 * H[0], 2.0 * H[1], 2.4494897427831781 * H[2],
                            2.8284271247461901 * H[3]};
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        sH += gh[k] * H[k];
        sdH += gh[k] * (dH[k] - u * H[k]);
      }
      dw += gr[P + a] / (npb[a] * g.area_pixel + kEps) + gw * sH;
      // v_los = (sxy (M6 vx + M7 vy) + sz M8 vz) KPCGYR_TO_KMS
      const double dv = sc * w * gw * sdH / sd[a] * kKpcGyrToKms;
      const double sxy = (c >= 2) ? -1.0 : 1.0;
      const double szz = (c & 1) ? -1.0 : 1.0;
      ynb[3] += dv * sxy * g.rot[6];
      ynb[4] += dv * sxy * g.rot[7];
      ynb[5] += dv * szz * g.rot[8];
      rotb[0] += dv * sxy * y[3];
      rotb[1] += dv * sxy * y[4];
      rotb[2] += dv * szz * y[5];
    }
  }
  return dw;
}

// ---------------------------------------------------------------------------
// Per-integration scale and T cotangent.  lib[c] = wt_c sum_r valid_r X_r,
// X_r = slab_r * c_e / T_r (c_e = 1, or 1/(num_per_bin area + eps) for the
// surface-density rows), wt_c = 1/(sum_r valid_r + 0.1).  With gbar the
// library cotangent: scale_r = wt_c valid_r / T_r multiplies every binned
// contribution, and dL/dT_r = -scale_r / T_r * sum_e gbar[c][e] c_e slab_r[e].
// One wave per integration.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void tdot_kernel(const long long* __restrict__ ws, int64_t n_int,
                                                   int r, int P, int B,
                                                   const double* __restrict__ valid,
                                                   const double* __restrict__ T_int,
                                                   const double* __restrict__ npb, double area,
                                                   const double* __restrict__ dt0, smx_schedule s,
                                                   int NC, const double* __restrict__ sd,
                                                   const double* __restrict__ gbar,
                                                   double* __restrict__ scale,
                                                   double* __restrict__ tbar) {
  const int lane = threadIdx.x & 63;
  const int64_t i = blockIdx.x * (int64_t)(blockDim.x >> 6) + (threadIdx.x >> 6);
  if (i >= n_int) return;
  const int64_t c = i / r;
  double nv = 0.0;
  for (int k = 0; k < r; ++k) nv += valid[c * r + k];
  const double sc = valid[i] > 0.5 ? (1.0 / (nv + 0.1)) / T_int[i] : 0.0;
  const int64_t rows = (int64_t)P + 6 * (int64_t)B;
  const double* gr = gbar + c * rows;
  const long long* rz = ws + i * slab_stride(P, B);
  const long long* ap = rz + slab_rz_len(P);
  double dot = 0.0;
  if (sc != 0.0) {
    // the slab's fixed-point sums in the forward's units (fix_exp_time)
    const int e_t = fix_exp_time(dt0[i], s, NC);
    const double un_t = fix_unscale(e_t);
    for (int64_t e = lane; e < rows; e += 64) {
      double v;
      if (e < P) {
        v = static_cast<double>(rz[e]) * un_t;
      } else {
        const int64_t q = e - P;
        const int64_t blk = q / B, a = q - blk * B;
        const int comp = (blk == 0) ? 0 : (blk == 5 ? 1 : (int)blk + 1);
        v = static_cast<double>(ap[kApRec * a + comp]) *
            (comp == 0 ? un_t : fix_unscale(e_t + fix_exp_gh(sd[a])));
        if (blk == 0) v = v / aperture_area(npb[a], area);
      }
      dot += gr[e] * v;
    }
  }
  dot = wave_sum(dot);
  if (lane == 0) {
    scale[i] = sc;
    tbar[i] = -sc * dot / T_int[i];
  }
}

// per-lane replay buffer: attempt a, component j of [y(6) dt k1(6) k2(6) k3(6) k4(6)]
constexpr int SEG_W = 31;

__device__ __forceinline__ double& segv(double* seg, int64_t n_int, int64_t i, int a, int j) {
  return seg[((int64_t)a * SEG_W + j) * n_int + i];
}

// Adjoint state of one integration between units: the cotangents of the
// state y and FSAL slope k1 before the unit's first attempt, of dt, and the
// accumulated cotangents of Omega, the potential fields and the rotation.
struct VjpState {
  double yb[6], k1b[6], dtb, omb, fb[PF_COUNT], rotb[3];
};
constexpr int kVjpHand = 12 + 2 + PF_COUNT + 3;   // doubles handed from a head to its tail

__device__ __forceinline__ void vjp_zero(VjpState& v) {
#pragma unroll
  for (int j = 0; j < 6; ++j) v.yb[j] = v.k1b[j] = 0.0;
  v.dtb = v.omb = 0.0;
#pragma unroll
  for (int f = 0; f < PF_COUNT; ++f) v.fb[f] = 0.0;
  v.rotb[0] = v.rotb[1] = v.rotb[2] = 0.0;
}

// One unit of the adjoint: segment sg of integration i (every attempts from
// its checkpoint), replayed into i's buffer and swept backwards.  own: this
// lane owns (i, sg) at this step, replays it and applies the sweep to v.
// Otherwise the lane shadows (i, sg), which another lane of the wave owns at
// the same step: it reads the buffer that lane writes, with zero cotangents
// and scale, so every cotangent it forms is exactly zero and v is untouched.
// Every lane runs all `every` steps of the sweep (attempts past a short last
// segment re-read its last attempt with zero cotangents), so the three
// deriv_vjp calls of each step are made with all lanes of the wave active.
__device__ __forceinline__ void vjp_unit(const smx_potential& p, const smx_grids& g,
                                         const smx_schedule& s, int64_t n_int, int r,
                                         const int32_t* __restrict__ bm,
                                         const double* __restrict__ npb,
                                         const double* __restrict__ v0,
                                         const double* __restrict__ sd,
                                         const double* __restrict__ gbar,
                                         const double* __restrict__ scale,
                                         const double* __restrict__ tbar,
                                         const double* __restrict__ ckpt, int every,
                                         double* __restrict__ seg, int64_t i, int sg, bool own,
                                         VjpState& v) {
  const int P = g.nR * g.nz * g.nphi;
  const int B = g.num_vbin;
  const int nsym = g.symmetry ? 4 : 1;
  const double Om = s.Omega;
  const double sc = own ? scale[i] : 0.0;
  const double Tb = own ? tbar[i] : 0.0;
  const double* gr = gbar + (i / r) * ((int64_t)P + 6 * (int64_t)B);
  const int a0 = sg * every;
  const int na = (s.N_max - a0 < every) ? s.N_max - a0 : every;
  // replay the segment from its checkpoint (owner only: no out-of-line calls)
  if (own) {
    const double* c = ckpt + ((int64_t)sg * n_int + i) * 8;
    double y[6], k1[6], dt = c[6], t = c[7];
#pragma unroll
    for (int j = 0; j < 6; ++j) y[j] = c[j];
    deriv(p, Om, y, k1);
    for (int a = 0; a < na; ++a) {
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        segv(seg, n_int, i, a, j) = y[j];
        segv(seg, n_int, i, a, 7 + j) = k1[j];
      }
      segv(seg, n_int, i, a, 6) = dt;
      double k2[6], k3[6], k4[6];
      bs3_attempt_stages(p, s, t, y, k1, dt, k2, k3, k4);
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        segv(seg, n_int, i, a, 13 + j) = k2[j];
        segv(seg, n_int, i, a, 19 + j) = k3[j];
        segv(seg, n_int, i, a, 25 + j) = k4[j];
      }
    }
  }
  // the owner's buffer writes above are read below by lanes shadowing i
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  // reverse sweep
  for (int a_step = every - 1; a_step >= 0; --a_step) {
    const bool act = own && a_step < na;
    const int a = a_step < na ? a_step : na - 1;
    double yb[6], k1b[6], dtb;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      yb[j] = act ? v.yb[j] : 0.0;
      k1b[j] = act ? v.k1b[j] : 0.0;
    }
    dtb = act ? v.dtb : 0.0;
    double& omb = v.omb;
    double* fb = v.fb;
    double* rotb = v.rotb;
    {
        double y[6], k1[6], k2[6], k3[6], k4[6];
#pragma unroll
        for (int j = 0; j < 6; ++j) {
          y[j] = segv(seg, n_int, i, a, j);
          k1[j] = segv(seg, n_int, i, a, 7 + j);
          k2[j] = segv(seg, n_int, i, a, 13 + j);
          k3[j] = segv(seg, n_int, i, a, 19 + j);
          k4[j] = segv(seg, n_int, i, a, 25 + j);
        }
        const double dt = segv(seg, n_int, i, a, 6);
        double yn[6], e3[6], q[6], scl[6];
        double es = 0.0;
#pragma unroll
        for (int j = 0; j < 6; ++j) {
          yn[j] = y[j] + dt * ((2.0 / 9.0 * k1[j] + 1.0 / 3.0 * k2[j]) + 4.0 / 9.0 * k3[j]);
          e3[j] = (-5.0 / 72.0 * k1[j] + 1.0 / 12.0 * k2[j]) + 1.0 / 9.0 * k3[j];
        }
#pragma unroll
        for (int j = 0; j < 6; ++j) {
          const double e = dt * (e3[j] + -1.0 / 8.0 * k4[j]);
          scl[j] = s.atol + s.rtol * nmax(fabs(y[j]), fabs(yn[j]));
          q[j] = e / scl[j];
          es += q[j] * q[j];
        }
        const double err = sqrt(es / 6.0);
        const bool accept = err <= 1.0;
        const double sf_raw = 0.9 * pow(nmax(err, 1e-10), -1.0 / 3.0);
        const double sf = nmin(nmax(sf_raw, 0.2), 5.0);
        const double dtn_raw = dt * sf;

        double ybA[6] = {0, 0, 0, 0, 0, 0}, ynb[6] = {0, 0, 0, 0, 0, 0};
        double k2b[6], k3b[6], k4b[6], k1l[6];
        double dtA = 0.0;
        if (accept) {
#pragma unroll
          for (int j = 0; j < 6; ++j) ynb[j] = yb[j];
          // the sample y_{a+1} = yn binned with weight dt; T += nsym dt
          const double dw = act ? bin_vjp(g, bm, v0, sd, npb, gr, yn, dt, sc, nsym, P, ynb, rotb) : 0.0;
          dtA += sc * dw + nsym * Tb;
        } else {
#pragma unroll
          for (int j = 0; j < 6; ++j) ybA[j] = yb[j];
        }
        // step-size controller: dt_{a+1} = clip(dt clip(0.9 max(err,1e-10)^(-1/3), 0.2, 5), dt_min, dt_max)
        const double prodb = (dtn_raw > s.dt_min && dtn_raw < s.dt_max) ? dtb : 0.0;
        dtA += prodb * sf;
        const double sfb = (sf_raw > 0.2 && sf_raw < 5.0) ? prodb * dt : 0.0;
        const double eb = (err > 1e-10) ? sfb * (-1.0 / 3.0) * sf_raw / err : 0.0;
#pragma unroll
        for (int j = 0; j < 6; ++j) {
          double evb = 0.0;
          if (err > 0.0 && eb != 0.0) {
            const double qb = eb * q[j] / (6.0 * err);
            evb = qb / scl[j];
            const double scb = -qb * q[j] / scl[j];
            const double ay = fabs(y[j]), an = fabs(yn[j]);
            const double sy = y[j] > 0.0 ? 1.0 : (y[j] < 0.0 ? -1.0 : 0.0);
            const double sn = yn[j] > 0.0 ? 1.0 : (yn[j] < 0.0 ? -1.0 : 0.0);
            if (ay > an) {
              ybA[j] += s.rtol * scb * sy;
            } else if (an > ay) {
              ynb[j] += s.rtol * scb * sn;
            } else {
              ybA[j] += 0.5 * s.rtol * scb * sy;
              ynb[j] += 0.5 * s.rtol * scb * sn;
            }
          }
          // err_j = dt (e3_j - k4_j / 8)
          dtA += evb * (e3[j] + -1.0 / 8.0 * k4[j]);
          k1l[j] = dt * (-5.0 / 72.0) * evb;
          k2b[j] = dt * (1.0 / 12.0) * evb;
          k3b[j] = dt * (1.0 / 9.0) * evb;
          k4b[j] = dt * (-1.0 / 8.0) * evb;
          // FSAL: k1_{a+1} = accept ? k4 : k1
          if (accept) k4b[j] += k1b[j];
          else k1l[j] += k1b[j];
        }
        // k4 = f(yn)
        deriv_vjp(p, Om, yn, k4b, ynb, omb, fb);
        // yn = y + dt ((2/9 k1 + 1/3 k2) + 4/9 k3)
#pragma unroll
        for (int j = 0; j < 6; ++j) {
          ybA[j] += ynb[j];
          k1l[j] += dt * (2.0 / 9.0) * ynb[j];
          k2b[j] += dt * (1.0 / 3.0) * ynb[j];
          k3b[j] += dt * (4.0 / 9.0) * ynb[j];
          dtA += ((2.0 / 9.0 * k1[j] + 1.0 / 3.0 * k2[j]) + 4.0 / 9.0 * k3[j]) * ynb[j];
        }
        // k3 = f(y + 0.75 dt k2)
        {
          double ys[6], ysb[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
          for (int j = 0; j < 6; ++j) ys[j] = y[j] + (0.75 * dt) * k2[j];
          deriv_vjp(p, Om, ys, k3b, ysb, omb, fb);
#pragma unroll
          for (int j = 0; j < 6; ++j) {
            ybA[j] += ysb[j];
            k2b[j] += (0.75 * dt) * ysb[j];
            dtA += 0.75 * k2[j] * ysb[j];
          }
        }
        // k2 = f(y + 0.5 dt k1)
        {
          double ys[6], ysb[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
          for (int j = 0; j < 6; ++j) ys[j] = y[j] + (0.5 * dt) * k1[j];
          deriv_vjp(p, Om, ys, k2b, ysb, omb, fb);
#pragma unroll
          for (int j = 0; j < 6; ++j) {
            ybA[j] += ysb[j];
            k1l[j] += (0.5 * dt) * ysb[j];
            dtA += 0.5 * k1[j] * ysb[j];
          }
        }
#pragma unroll
        for (int j = 0; j < 6; ++j) {
          yb[j] = ybA[j];
          k1b[j] = k1l[j];
        }
        dtb = dtA;
    }
    if (act) {
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        v.yb[j] = yb[j];
        v.k1b[j] = k1b[j];
      }
      v.dtb = dtb;
    }
  }
}

// Balanced over the resident waves like the library integrator: each wave
// takes the contiguous share of m integrations [w n / W, (w + 1) n / W) and
// spreads their m U units (U = ceil(N_max / every) segments each, swept from
// the last segment to the first) over its lanes by LaneSchedule; a split
// integration's adjoint state passes from the lane that sweeps its last
// segments to the lane that sweeps the rest through LDS.  A lane without a
// unit at a step shadows lane 0's unit (lane 0 is busy at every step).  After
// each step every lane makes one more deriv_vjp call: k1_0 = f(y_0) of an
// integration finished at that step, with zero cotangent elsewhere.
__global__ __launch_bounds__(64) void library_vjp_kernel(
    smx_potential p, smx_grids g, smx_schedule s, const double* __restrict__ wb,
    const double* __restrict__ dt0, int64_t n_int, int r, const int32_t* __restrict__ bm,
    const double* __restrict__ npb, const double* __restrict__ v0, const double* __restrict__ sd,
    const double* __restrict__ gbar, const double* __restrict__ scale,
    const double* __restrict__ tbar, const double* __restrict__ ckpt, int every,
    double* __restrict__ seg, double* __restrict__ ybar0, double* __restrict__ dtbar0,
    double* __restrict__ fbar, double* __restrict__ ombar, double* __restrict__ rotbar) {
  __shared__ double handoff[kVjpHand * 64];
  const int lane = threadIdx.x;
  const int64_t W = gridDim.x, w = blockIdx.x;
  const int64_t j0 = w * n_int / W, m = (w + 1) * n_int / W - j0;
  const int64_t U = (s.N_max + every - 1) / every;
  if (m <= 0) return;
  const double Om = s.Omega;
  if (U <= 0) {
    // no attempts: the cotangents are those of k1_0 = f(y_0) with k1b = 0
    for (int64_t q = lane; q < m; q += 64) {
      const int64_t i = j0 + q;
#pragma unroll
      for (int j = 0; j < 6; ++j) ybar0[6 * i + j] = 0.0;
      dtbar0[i] = 0.0;
      for (int f = 0; f < PF_COUNT; ++f) fbar[i * PF_COUNT + f] = 0.0;
      ombar[i] = 0.0;
      for (int j = 0; j < 3; ++j) rotbar[3 * i + j] = 0.0;
    }
    return;
  }
  LaneSchedule ls, l0;
  ls.init(lane, m, U);
  l0.init(0, m, U);
  const int nseg = ls.nseg;
  int sidx = 0, kind = 0, k0 = 0;
  int64_t job = 0, first = 0, t0 = 0, t1 = 0;
  if (nseg > 0) ls.segment(0, kind, job, first, t0, t1);
  int s0 = 0, kind0 = 0;
  int64_t job0 = 0, first0 = 0, a00 = 0, a01 = 0;
  l0.segment(0, kind0, job0, first0, a00, a01);
  (void)k0;
  VjpState v;
  vjp_zero(v);
  int64_t cur = -1;    // integration this lane holds, or -1
  int64_t unit = 0;    // its next unit (0 = its last segment)
  for (int64_t step = 0; step < ls.C; ++step) {
    // lane 0's unit at this step (every lane follows lane 0's schedule)
    while (step >= a01) {
      ++s0;
      l0.segment(s0, kind0, job0, first0, a00, a01);
    }
    const int64_t u0 = first0 + (step - a00);
    // start of this lane's next segment
    if (sidx < nseg && cur < 0 && step == t0) {
      cur = job;
      unit = first;
      if (kind == 2) {
        const double* h = handoff + (lane - 1);
#pragma unroll
        for (int j = 0; j < 6; ++j) {
          v.yb[j] = h[j * 64];
          v.k1b[j] = h[(6 + j) * 64];
        }
        v.dtb = h[12 * 64];
        v.omb = h[13 * 64];
#pragma unroll
        for (int f = 0; f < PF_COUNT; ++f) v.fb[f] = h[(14 + f) * 64];
#pragma unroll
        for (int j = 0; j < 3; ++j) v.rotb[j] = h[(14 + PF_COUNT + j) * 64];
      } else {
        vjp_zero(v);
      }
    }
    const bool own = cur >= 0;
    const int64_t i = j0 + (own ? cur : job0);
    const int sgi = (int)(U - 1 - (own ? unit : u0));
    vjp_unit(p, g, s, n_int, r, bm, npb, v0, sd, gbar, scale, tbar, ckpt, every, seg, i, sgi, own,
             v);
    bool finished = false;
    if (own) {
      ++unit;
      if (step + 1 == t1) {   // end of this lane's segment
        if (kind == 0) {
          double* h = handoff + lane;
#pragma unroll
          for (int j = 0; j < 6; ++j) {
            h[j * 64] = v.yb[j];
            h[(6 + j) * 64] = v.k1b[j];
          }
          h[12 * 64] = v.dtb;
          h[13 * 64] = v.omb;
#pragma unroll
          for (int f = 0; f < PF_COUNT; ++f) h[(14 + f) * 64] = v.fb[f];
#pragma unroll
          for (int j = 0; j < 3; ++j) h[(14 + PF_COUNT + j) * 64] = v.rotb[j];
          cur = -1;
        } else {
          finished = true;
        }
        ++sidx;
        if (sidx < nseg) ls.segment(sidx, kind, job, first, t0, t1);
      }
    }
    // k1_0 = f(y_0) for an integration finished at this step; zero cotangent
    // (at a valid state) elsewhere
    {
      const int64_t ie = finished ? i : j0 + job0;
      double y0[6], k1z[6];
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        y0[j] = wb[6 * ie + j];
        k1z[j] = finished ? v.k1b[j] : 0.0;
      }
      deriv_vjp(p, Om, y0, k1z, v.yb, v.omb, v.fb);
    }
    if (finished) {
#pragma unroll
      for (int j = 0; j < 6; ++j) ybar0[6 * i + j] = v.yb[j];
      dtbar0[i] = v.dtb;
#pragma unroll
      for (int f = 0; f < PF_COUNT; ++f) fbar[i * PF_COUNT + f] = v.fb[f];
      ombar[i] = v.omb;
      for (int j = 0; j < 3; ++j) rotbar[3 * i + j] = v.rotb[j];
      cur = -1;
      vjp_zero(v);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");   // hand-overs before later reads
  }
}

// Waves of the balanced adjoint: the waves resident at once when there are
// more integrations than their lanes, else one lane per integration.
static int64_t vjp_waves(int64_t n_int) {
  static int resident = 0;
  static std::mutex mu;
  {
    std::lock_guard<std::mutex> lock(mu);
    if (resident == 0) {
      int dev = 0, cus = 0, per_cu = 0;
      if (hipGetDevice(&dev) == hipSuccess &&
          hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
          hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, library_vjp_kernel, 64, 0) ==
              hipSuccess &&
          cus > 0 && per_cu > 0)
        resident = cus * per_cu;
      else
        resident = 1024;
    }
  }
  const int64_t lanes = (n_int + 63) / 64;
  return lanes < resident ? lanes : resident;
}

// ---------------------------------------------------------------------------
// Field Jacobians of the initial conditions and the density target.
// ---------------------------------------------------------------------------
// One group of NJ fields (F0 .. F0 + NJ - 1) of the initial conditions'
// Jacobian, from the wave's three pressure integrals in gw (value, then the
// PF_COUNT field derivatives, each).  F0 is a constant, so the compiler drops
// the tangent arithmetic of the unseeded fields; out of line, so that the
// passes do not share (and keep live) their common primal values.
constexpr int IC_GW = PF_COUNT + 1;

template <int NJ, int F0>
__device__ __noinline__ void ic_jac_pass(const smx_potential& p, const double* __restrict__ w0,
                                         const double* __restrict__ normals,
                                         const double* __restrict__ uniforms, int64_t i, int lane,
                                         int r, int r_total, double N_step, const double* gw,
                                         double* __restrict__ dwb, double* __restrict__ ddt0) {
  Dual<NJ> I[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    I[q] = dconst<NJ>(gw[q * IC_GW]);
#pragma unroll
    for (int k = 0; k < NJ; ++k) I[q].d[k] = (F0 + k < PF_COUNT) ? gw[q * IC_GW + 1 + F0 + k] : 0.0;
  }
  StarIC<Dual<NJ>> o;
  star_ic_finish<Dual<NJ>>(p, fields_seeded<NJ>(p, F0), w0, normals, i, I[0], I[1], I[2], o);
  if (lane != 0) return;
  for (int k = 0; k < r; ++k) {
    const int64_t q = i * r + k;
    for (int j = 0; j < 6; ++j) {
      const double u = uniforms[(i * r_total + k) * 6 + j] - 0.5;
#pragma unroll
      for (int t = 0; t < NJ; ++t) {
        if (F0 + t >= PF_COUNT) continue;
        const double ds = (j < 3) ? 0.0 : o.vsc.d[t];
        dwb[(6 * q + j) * PF_COUNT + F0 + t] = o.wn[j].d[t] + u * ds;
      }
    }
#pragma unroll
    for (int t = 0; t < NJ; ++t)
      if (F0 + t < PF_COUNT) ddt0[q * PF_COUNT + F0 + t] = o.T_orb.d[t] / N_step;
  }
}

// The pressure integrals and their field derivatives come from one pass
// (pressure_part_grad); the rest of the initial conditions runs once per
// group of NJ fields on Dual<NJ> (ic_jac_pass).  The wave's three integrals
// wait in LDS (lane 0 writes them, the wave's lanes read them back in
// order), so that they hold no registers during those passes.
template <int NJ>
__global__ __launch_bounds__(256) void ic_jac_kernel(smx_potential p, const double* __restrict__ w0,
                                                     const double* __restrict__ normals,
                                                     const double* __restrict__ uniforms, int64_t n,
                                                     int r, int r_total, double N_step,
                                                     double* __restrict__ dwb,
                                                     double* __restrict__ ddt0) {
  static_assert(NJ == 4 && PF_COUNT <= 12, "three passes of four fields");
  __shared__ double gsh[4][3 * IC_GW];
  const int lane = threadIdx.x & 63;
  const int64_t i = blockIdx.x * (int64_t)(blockDim.x >> 6) + (threadIdx.x >> 6);
  if (i >= n) return;
  double* gw = gsh[threadIdx.x >> 6];
  {
    const double xs = w0[3 * i], ys = w0[3 * i + 1], zs = w0[3 * i + 2];
    const double Rs = sqrt(xs * xs + ys * ys);
    const double zlo = fabs(zs);
    const FieldGrad I[3] = {
        wave_sum(pressure_part_grad(p, xs, ys, zlo, JEANS_NPTS, lane)),
        wave_sum(pressure_part_grad(p, Rs + JEANS_DR, 0.0, zlo, JEANS_NPTS, lane)),
        wave_sum(pressure_part_grad(p, Rs - JEANS_DR, 0.0, zlo, JEANS_NPTS, lane))};
    if (lane == 0)
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        gw[q * IC_GW] = I[q].v;
#pragma unroll
        for (int k = 0; k < PF_COUNT; ++k) gw[q * IC_GW + 1 + k] = I[q].d[k];
      }
  }
  ic_jac_pass<NJ, 0>(p, w0, normals, uniforms, i, lane, r, r_total, N_step, gw, dwb, ddt0);
  ic_jac_pass<NJ, NJ>(p, w0, normals, uniforms, i, lane, r, r_total, N_step, gw, dwb, ddt0);
  ic_jac_pass<NJ, 2 * NJ>(p, w0, normals, uniforms, i, lane, r, r_total, N_step, gw, dwb, ddt0);
}

template <int NJ>
__global__ __launch_bounds__(256) void density_jac_kernel(
    smx_potential p, const double* __restrict__ Rg, const double* __restrict__ zg,
    const double* __restrict__ pg, int64_t P, double dR, double dz, double dphi,
    const double* __restrict__ smp, int64_t S, int f0, double* __restrict__ jac) {
  __shared__ double red[4][NJ];
  const int64_t c = blockIdx.x;
  if (c >= P) return;
  Dual<NJ> acc = cell_mass_part<Dual<NJ>>(p, fields_seeded<NJ>(p, f0), Rg[c], zg[c], pg[c], dR, dz,
                                          dphi, smp, S, threadIdx.x, blockDim.x);
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0)
    for (int t = 0; t < NJ; ++t) red[threadIdx.x >> 6][t] = acc.d[t];
  __syncthreads();
  if (threadIdx.x < NJ && f0 + (int)threadIdx.x < PF_COUNT) {
    double tot = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) tot += red[w][threadIdx.x];
    jac[c * PF_COUNT + f0 + threadIdx.x] = ((tot / static_cast<double>(S)) * dR * dz) * dphi;
  }
}

constexpr int JAC_SLOTS = 4;

}  // namespace smx

using namespace smx;

extern "C" {

// Host replay of library_vjp_kernel's step loop (segment starts and ends,
// hand-overs, shadows), with the kernel's own schedule: every integration
// must have its units 0 .. U-1 swept once and in order, a tail must start
// from the state its head handed over, a lane without a unit must find lane
// 0 on the unit it shadows, and every integration is finished once.
// Returns 0 when all hold; *steps_out = the loop length of wave 0.
int smx_adjoint_schedule_check(int64_t n_int, int64_t waves, int64_t U, int64_t* steps_out) {
  SMX_REQUIRE(n_int >= 0 && waves >= 1 && U >= 1 && steps_out,
              "smx_adjoint_schedule_check: bad arguments");
  std::vector<int64_t> next(n_int, 0), done(n_int, 0);
  *steps_out = 0;
  for (int64_t w = 0; w < waves; ++w) {
    const int64_t j0 = w * n_int / waves, m = (w + 1) * n_int / waves - j0;
    if (m <= 0) continue;
    LaneSchedule ls[64], l0;
    int sidx[64], kind[64];
    int64_t job[64], first[64], t0[64], t1[64], cur[64], unit[64], hand_job[64], hand_unit[64];
    for (int l = 0; l < 64; ++l) {
      ls[l].init(l, m, U);
      sidx[l] = 0;
      cur[l] = -1;
      hand_job[l] = -1;
      if (ls[l].nseg > 0) ls[l].segment(0, kind[l], job[l], first[l], t0[l], t1[l]);
    }
    l0.init(0, m, U);
    int s0 = 0, kind0 = 0;
    int64_t job0 = 0, first0 = 0, a00 = 0, a01 = 0;
    l0.segment(0, kind0, job0, first0, a00, a01);
    const int64_t C = ls[0].C;
    if (w == 0) *steps_out = C;
    for (int64_t step = 0; step < C; ++step) {
      while (step >= a01) {
        if (++s0 >= l0.nseg) return 5;
        l0.segment(s0, kind0, job0, first0, a00, a01);
      }
      const int64_t u0 = first0 + (step - a00);
      int64_t lane0_job = -2, lane0_unit = -2;
      for (int l = 0; l < 64; ++l) {
        if (sidx[l] < ls[l].nseg && cur[l] < 0 && step == t0[l]) {
          cur[l] = job[l];
          unit[l] = first[l];
          if (kind[l] == 2 &&
              (l == 0 || hand_job[l - 1] != job[l] || hand_unit[l - 1] != first[l]))
            return 2;
        }
        if (cur[l] < 0) continue;
        int64_t& nx = next[j0 + cur[l]];
        if (unit[l] != nx) return 3;
        if (l == 0) {
          lane0_job = cur[l];
          lane0_unit = unit[l];
        }
        ++nx;
        ++unit[l];
        if (step + 1 == t1[l]) {
          if (kind[l] == 0) {
            hand_job[l] = cur[l];
            hand_unit[l] = unit[l];
          } else if (unit[l] != U || done[j0 + cur[l]]++ != 0) {
            return 1;
          }
          cur[l] = -1;
          if (++sidx[l] < ls[l].nseg) ls[l].segment(sidx[l], kind[l], job[l], first[l], t0[l], t1[l]);
        }
      }
      if (lane0_job != job0 || lane0_unit != u0) return 6;   // the shadows' unit
    }
  }
  for (int64_t i = 0; i < n_int; ++i)
    if (next[i] != U || done[i] != 1) return 4;
  return 0;
}

int64_t smx_library_vjp_workspace_bytes(int64_t n_int, int32_t ckpt_every) {
  if (n_int < 0 || ckpt_every < 1) return -1;
  return (int64_t)sizeof(double) * (n_int * SEG_W * (int64_t)ckpt_every + 2 * n_int) + 256;
}

int smx_library_vjp(const smx_potential* p, const smx_grids* g, const smx_schedule* s,
                    const double* wb, const double* dt0, int64_t n, int32_t r,
                    const int32_t* bin_mapping, const double* num_per_bin, const double* v0,
                    const double* sdisp, const void* fwd_workspace, const double* valid,
                    const double* T_int, const double* ckpt, int32_t ckpt_every,
                    const double* lib_bar, void* workspace, double* wb_bar, double* dt0_bar,
                    double* field_bar, double* omega_bar, double* rot_bar, void* stream) {
  SMX_REQUIRE(p && g && s && wb && dt0 && bin_mapping && num_per_bin && v0 && sdisp &&
                  fwd_workspace && valid && T_int && ckpt && lib_bar && workspace && wb_bar &&
                  dt0_bar && field_bar && omega_bar && rot_bar,
              "smx_library_vjp: null pointer");
  SMX_REQUIRE(n >= 0 && r >= 1 && s->N_max >= 0 && ckpt_every >= 1, "smx_library_vjp: bad sizes");
  SMX_REQUIRE(g->nR > 0 && g->nz > 0 && g->nphi > 0 && g->num_vbin > 0, "smx_library_vjp: bad grids");
  if (n == 0) return 0;
  const int64_t n_int = n * r;
  const int P = g->nR * g->nz * g->nphi;
  hipStream_t st = as_stream(stream);
  double* seg = static_cast<double*>(workspace);
  double* scale = seg + n_int * SEG_W * (int64_t)ckpt_every;
  double* tbar = scale + n_int;
  stage_begin(ST_ADJOINT, st);
  hipLaunchKernelGGL(tdot_kernel, dim3(blocks_for(n_int, 4)), dim3(256), 0, st,
                     static_cast<const long long*>(fwd_workspace), n_int, (int)r, P, g->num_vbin,
                     valid, T_int, num_per_bin, g->area_pixel, dt0, *s, g->symmetry ? 4 : 1, sdisp,
                     lib_bar, scale, tbar);
  SMX_LAUNCHED();
  hipLaunchKernelGGL(library_vjp_kernel, dim3((unsigned)vjp_waves(n_int)), dim3(64), 0, st, *p, *g, *s,
                     wb, dt0, n_int, (int)r, bin_mapping, num_per_bin, v0, sdisp, lib_bar, scale,
                     tbar, ckpt, (int)ckpt_every, seg, wb_bar, dt0_bar, field_bar, omega_bar,
                     rot_bar);
  SMX_LAUNCHED();
  stage_end(ST_ADJOINT, st, 2);
  return 0;
}

int smx_initial_conditions_jac(const smx_potential* p, const double* w0, const double* normals,
                               const double* uniforms, int64_t n, int32_t r, int32_t r_total,
                               double N_step, double* wb_jac, double* dt0_jac, void* stream) {
  SMX_REQUIRE(p && w0 && normals && uniforms && wb_jac && dt0_jac,
              "smx_initial_conditions_jac: null pointer");
  SMX_REQUIRE(n >= 0 && r >= 1 && r_total >= r && N_step > 0.0,
              "smx_initial_conditions_jac: bad sizes");
  if (n == 0) return 0;
  hipStream_t st = as_stream(stream);
  const int waves_per_block = 4;
  stage_begin(ST_JACOBIAN, st);
  hipLaunchKernelGGL(ic_jac_kernel<JAC_SLOTS>, dim3(blocks_for(n, waves_per_block)),
                     dim3(64 * waves_per_block), 0, st, *p, w0, normals, uniforms, n, (int)r,
                     (int)r_total, N_step, wb_jac, dt0_jac);
  SMX_LAUNCHED();
  stage_end(ST_JACOBIAN, st, 1);
  return 0;
}

int smx_density_target_jac(const smx_potential* p, const double* R_grid, const double* z_grid,
                           const double* phi_grid, int64_t P, double dR, double dz, double dphi,
                           const double* sobol, int64_t S, double* jac, void* stream) {
  SMX_REQUIRE(p && R_grid && z_grid && phi_grid && sobol && jac,
              "smx_density_target_jac: null pointer");
  SMX_REQUIRE(P >= 0 && S > 0, "smx_density_target_jac: bad sizes");
  if (P == 0) return 0;
  hipStream_t st = as_stream(stream);
  stage_begin(ST_JACOBIAN, st);
  for (int f0 = 0; f0 < PF_COUNT; f0 += JAC_SLOTS) {
    hipLaunchKernelGGL(density_jac_kernel<JAC_SLOTS>, dim3((unsigned)P), dim3(256), 0, st, *p,
                       R_grid, z_grid, phi_grid, P, dR, dz, dphi, sobol, S, f0, jac);
    SMX_LAUNCHED();
  }
  stage_end(ST_JACOBIAN, st, 3);
  return 0;
}

}  // extern "C"
