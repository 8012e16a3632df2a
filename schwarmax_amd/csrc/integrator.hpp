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
// kApRec = 6 doubles [dt, g dt H0, ..., g dt H4] (48 bytes, 16-byte aligned)
constexpr int kApRec = 6;
__host__ __device__ __forceinline__ int64_t slab_rz_len(int P) { return (P + 7) & ~7; }
__host__ __device__ __forceinline__ int64_t slab_stride(int P, int B) {
  return slab_rz_len(P) + kApRec * (int64_t)B;
}

// Register cache of the (R, z, phi) cell each symmetry copy last hit.  An
// orbit's consecutive accepted steps mostly stay in one cell, so a copy's
// weights are summed in a register and added to the slab when the copy moves
// to another cell, and at the end of the integration: the slab receives the
// same sums, with the additions grouped per visit.  Most of the binning's
// cost is slab cache-line traffic (12.9 GB of slabs at 80,000 integrations),
// which this removes for the cells.  Out-of-grid copies (id -1) are never
// written.
template <int NC>
struct CellCache {
  int cell[NC];
  double w[NC];
  __device__ __forceinline__ void init() {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      cell[c] = -1;
      w[c] = 0.0;
    }
  }
  __device__ __forceinline__ void add(double* __restrict__ rz, int c, int id, double wt) {
    if (id == cell[c]) {
      w[c] += wt;
      return;
    }
    if (cell[c] >= 0) rz[cell[c]] += w[c];
    cell[c] = id;
    w[c] = wt;
  }
  __device__ __forceinline__ void flush(double* __restrict__ rz) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      if (cell[c] >= 0) rz[cell[c]] += w[c];
      cell[c] = -1;
      w[c] = 0.0;
    }
  }
};

// Cache of the aperture record each symmetry copy last hit, the same scheme
// as CellCache for the six record terms [dt, g dt H0..H4]: a copy's terms
// are summed while its samples fall in one aperture and added to the slab's
// record when it moves to another aperture, and at the end of the
// integration.  A copy in no aperture (id -1) leaves its cached record as it
// is.  The ids live in registers, the sums in LDS (the integrator's registers
// are full): lane l's term k of copy c at sum[(6 c + k) * 64 + l], so a
// wave's accesses are consecutive doubles.
template <int NC>
struct ApCache {
  int id[NC];
  double* sum;   // this lane's column of a __shared__ (6 NC) x 64 array
  __device__ __forceinline__ void init(double* lds_col) {
    sum = lds_col;
#pragma unroll
    for (int c = 0; c < NC; ++c) id[c] = -1;
  }
  __device__ __forceinline__ void write(double* __restrict__ ap, int c) const {
    if (id[c] < 0) return;
    double* rec = ap + kApRec * (int64_t)id[c];
#pragma unroll
    for (int k = 0; k < 6; ++k) rec[k] += sum[(6 * c + k) * 64];
  }
  __device__ __forceinline__ void add(double* __restrict__ ap, int c, int a, double w, double gd,
                                      const double (&H)[5]) {
    if (a != id[c]) {
      write(ap, c);
      id[c] = a;
      sum[(6 * c) * 64] = w;
#pragma unroll
      for (int k = 0; k < 5; ++k) sum[(6 * c + 1 + k) * 64] = gd * H[k];
      return;
    }
    sum[(6 * c) * 64] += w;
#pragma unroll
    for (int k = 0; k < 5; ++k) sum[(6 * c + 1 + k) * 64] += gd * H[k];
  }
  __device__ __forceinline__ void flush(double* __restrict__ ap) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      write(ap, c);
      id[c] = -1;
    }
  }
};

// Binning of one accepted sample's NC symmetry copies with weight w: cells
// and aperture records through their register caches.
template <int NC>
__device__ __forceinline__ void bin_sample(const smx_grids& g, const int32_t* __restrict__ bm,
                                           const double* __restrict__ v0,
                                           const double* __restrict__ sd, const double* y,
                                           double w, CellCache<NC>& cache, ApCache<NC>& acache,
                                           double* __restrict__ rz, double* __restrict__ ap) {
  SampleProj sp;
  project(g, y, NC > 2, sp);
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    double zs, phi, xr, yr, vlos;
    copy_coords(sp, c, zs, phi, xr, yr, vlos);
    cache.add(rz, c, rzphi_cell(g, sp.R, zs * y[2], phi), w);
    const int a = aperture(g, bm, xr, yr);
    if (a >= 0) {
      double gw, H[5];
      gh_terms(vlos, v0[a], sd[a], gw, H);
      acache.add(ap, c, a, w, gw * w, H);
    }
  }
}

// The accepted-step loop of one integration with binning: N_max BS3 attempts
// (integrates.py:181-205), every accepted state binned with weight dt in its
// NC copies; returns the accumulated time (NC x the integrated time).  With
// ckpt, the state (y, dt, t) before every ckpt_every-th attempt is stored at
// ckpt + (attempt / ckpt_every) * ckpt_stride (8 doubles).  ap_lds: this
// lane's column of a __shared__ array of at least 6 NC x 64 doubles (the
// aperture-record cache; the block is one wave of 64 lanes).
template <int NC>
__device__ __forceinline__ double integrate_and_bin(const smx_potential& p, const smx_grids& g,
                                                    const smx_schedule& s, const int32_t* bm,
                                                    const double* v0, const double* sd, double& t,
                                                    double* y, double* k1, double& dt,
                                                    double* __restrict__ rz,
                                                    double* __restrict__ ap, double* ckpt,
                                                    int64_t ckpt_stride, int ckpt_every,
                                                    double* ap_lds) {
  CellCache<NC> cache;
  cache.init();
  ApCache<NC> acache;
  acache.init(ap_lds);
  double T_acc = 0.0;
  for (int step = 0; step < s.N_max; ++step) {
    if (ckpt != nullptr && step % ckpt_every == 0) {
      double* c = ckpt + (int64_t)(step / ckpt_every) * ckpt_stride;
#pragma unroll
      for (int j = 0; j < 6; ++j) c[j] = y[j];
      c[6] = dt;
      c[7] = t;
    }
    const double used = bs3_attempt(p, s, t, y, k1, dt);
    if (used != 0.0) {
      bin_sample<NC>(g, bm, v0, sd, y, used, cache, acache, rz, ap);
      T_acc += NC * used;
    }
  }
  cache.flush(rz);
  acache.flush(ap);
  return T_acc;
}

// Work schedule of the balanced integrator and adjoint kernels (McNaughton's
// wrap-around rule within one wave): m jobs of L equal units each, spread over
// the wave's 64 lanes.  Lane l covers units [l C, (l + 1) C) of the jobs'
// concatenation, C = max(L, ceil(m L / 64)); a job cut at a lane boundary
// runs its first units (its head) on lane l at the start of the loop and the
// rest (its tail) on lane l + 1 at the end, ending at step C, so the parts
// never overlap (C >= L).
// One lane's schedule: segments in loop-step order.  m integrations of L
// attempts on the wave's 64 lanes; lane `lane`'s share of the attempts.
struct LaneSchedule {
  int64_t C, L;
  int64_t jb, hb;          // head: integration jb, attempts [0, hb)
  int64_t jf, nfull;       // whole integrations jf .. jf + nfull - 1 from step hb
  int64_t ja, ta;          // tail: integration ja, attempts [ta, L), steps [C - L + ta, C)
  int nseg;                // segments of this lane
  __host__ __device__ void init(int lane, int64_t m, int64_t L_) {
    L = L_;
    C = (m * L + 63) / 64 > L ? (m * L + 63) / 64 : L;
    const int64_t a0 = lane * C < m * L ? lane * C : m * L;
    const int64_t a1 = (lane + 1) * C < m * L ? (lane + 1) * C : m * L;
    jb = a1 / L;
    hb = jb < m ? a1 - jb * L : 0;
    jf = (a0 + L - 1) / L;
    nfull = a1 / L - jf > 0 ? a1 / L - jf : 0;
    ja = a0 / L;
    ta = ja < m ? a0 - ja * L : 0;
    nseg = a1 > a0 ? (hb > 0 ? 1 : 0) + (int)nfull + (ta > 0 ? 1 : 0) : 0;
  }
  // segment s (0 .. nseg-1) -> (kind, integration, first attempt, start step, end step)
  __host__ __device__ void segment(int sidx, int& kind, int64_t& job, int64_t& first, int64_t& t0,
                                   int64_t& t1) const {
    int q = sidx;
    if (hb > 0) {
      if (q == 0) { kind = 0; job = jb; first = 0; t0 = 0; t1 = hb; return; }
      --q;
    }
    if (q < nfull) { kind = 1; job = jf + q; first = 0; t0 = hb + q * L; t1 = t0 + L; return; }
    q -= (int)nfull;
    kind = 2; job = ja; first = ta; t0 = C - L + ta; t1 = C;   // q == 0 and ta > 0
  }
};


}  // namespace smx
