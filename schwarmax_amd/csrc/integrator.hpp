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
// kApRec = 6 terms [dt, g dt H0, ..., g dt H4] (48 bytes, 16-byte aligned),
// every entry a 64-bit fixed-point sum (FixScale)
constexpr int kApRec = 6;
__host__ __device__ __forceinline__ int64_t slab_rz_len(int P) { return (P + 7) & ~7; }
__host__ __device__ __forceinline__ int64_t slab_stride(int P, int B) {
  return slab_rz_len(P) + kApRec * (int64_t)B;
}

// ---------------------------------------------------------------------------
// Fixed-point sums.  Each binned contribution (a step length, or a step
// length times a Gauss-Hermite term) is rounded to a multiple of a power of
// two and the multiples are added as 64-bit integers.  Integer addition is
// exact and associative, so a cell's or record term's sum does not depend on
// the order or grouping of its contributions: the fused integrator (register
// caches, one lane per integration), the persistent batched kernel and the
// LDS histograms of the sample-binning kernel all produce the same bits.
// The power of two comes from a bound on the sum, so it never overflows:
//   * step lengths: every accepted step is at most max(dt0, dt_max) (the first
//     attempt uses dt0, later ones are clamped to dt_max), so an integration's
//     NC copies sum to less than B_t = NC N_max max(dt0, dt_max) < 2^e_t;
//   * g H_k terms: g = exp(-y^2/2) / (sqrt(2 pi) s) and |exp(-y^2/2) H_k(y)|
//     <= 1 for k <= 4, so |g H_k| < 2 / (sqrt(2 pi) s) < 2^e_g(s).
// Sums are held at 2^-(61 - e) resolution (< 2^61 in magnitude): about 2^-60
// of the bound, 1e-13 of a typical step's contribution at the 20,000-orbit
// shapes, far below the FP64 library's 1e-8 parity bar.
// ---------------------------------------------------------------------------
constexpr int kFixBits = 61;

__host__ __device__ __forceinline__ int fix_exp_of(double bound) {
  int e = 0;
  frexp(bound, &e);                       // bound < 2^e for a finite bound > 0
  if (!(bound > 0.0) || !(bound < 1e300)) e = 0;   // non-finite: its integration is invalid
  return e < -900 ? -900 : (e > 900 ? 900 : e);
}

// e_t of one integration
__host__ __device__ __forceinline__ int fix_exp_time(double dt0, const smx_schedule& s, int NC) {
  const double dtm = dt0 > s.dt_max ? dt0 : s.dt_max;
  return fix_exp_of(static_cast<double>(NC) * static_cast<double>(s.N_max > 0 ? s.N_max : 1) * dtm);
}

// e_g of an aperture with dispersion s (> 0, finite: the host checks)
__device__ __forceinline__ int fix_exp_gh(double s) { return fix_exp_of(2.0 * kInvSqrt2Pi / s); }

// scale 2^(61 - e) and its inverse
__device__ __forceinline__ double fix_scale(int e) { return ldexp(1.0, kFixBits - e); }
__device__ __forceinline__ double fix_unscale(int e) { return ldexp(1.0, e - kFixBits); }

// Rounded multiple; a value off the bound (only from an integration that
// diverged, which the validity gate then drops) saturates at +-2^62 and NaN
// maps to 0, so the conversion stays defined.
__device__ __forceinline__ long long fix_q(double x, double scale) {
  const double v = rint(x * scale);
  const double lim = 4611686018427387904.0;   // 2^62
  return static_cast<long long>(v < lim ? (v > -lim ? v : -lim) : (v == v ? lim : 0.0));
}

// Area of an aperture of npb pixels (integrates.py:212-236), one rounding
// per operation wherever the library is finalized.
__host__ __device__ __forceinline__ double aperture_area(double npb, double area_pixel) {
#pragma clang fp contract(off)
  return npb * area_pixel + kEps;
}

// Register cache of the (R, z, phi) cell each symmetry copy last hit.  An
// orbit's consecutive accepted steps mostly stay in one cell, so a copy's
// fixed-point weights are summed in a register and added to the slab when the
// copy moves to another cell, and at the end of the integration: the slab
// receives the same integer sum.  Out-of-grid copies (id -1) are never
// written.
template <int NC>
struct CellCache {
  int cell[NC];
  long long w[NC];
  __device__ __forceinline__ void init() {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      cell[c] = -1;
      w[c] = 0;
    }
  }
  __device__ __forceinline__ void add(long long* __restrict__ rz, int c, int id, long long q) {
    if (id == cell[c]) {
      w[c] += q;
      return;
    }
    if (cell[c] >= 0) rz[cell[c]] += w[c];
    cell[c] = id;
    w[c] = q;
  }
  __device__ __forceinline__ void flush(long long* __restrict__ rz) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      if (cell[c] >= 0) rz[cell[c]] += w[c];
      cell[c] = -1;
      w[c] = 0;
    }
  }
};

// Cache of the aperture record each symmetry copy last hit, the same scheme
// as CellCache for the six record terms: a copy's terms are summed while its
// samples fall in one aperture and added to the slab's record when it moves
// to another aperture, and at the end of the integration.  A copy in no
// aperture (id -1) leaves its cached record as it is.  The ids live in
// registers, the sums in LDS (the integrator's registers are full): lane l's
// term k of copy c at sum[(6 c + k) * 64 + l], so a wave's accesses are
// consecutive.
template <int NC>
struct ApCache {
  int id[NC];
  long long* sum;   // this lane's column of a __shared__ (6 NC) x 64 array
  __device__ __forceinline__ void init(long long* lds_col) {
    sum = lds_col;
#pragma unroll
    for (int c = 0; c < NC; ++c) id[c] = -1;
  }
  __device__ __forceinline__ void write(long long* __restrict__ ap, int c) const {
    if (id[c] < 0) return;
    long long* rec = ap + kApRec * (int64_t)id[c];
#pragma unroll
    for (int k = 0; k < kApRec; ++k) rec[k] += sum[(6 * c + k) * 64];
  }
  __device__ __forceinline__ void add(long long* __restrict__ ap, int c, int a, const long long (&q)[kApRec]) {
    if (a != id[c]) {
      write(ap, c);
      id[c] = a;
#pragma unroll
      for (int k = 0; k < kApRec; ++k) sum[(6 * c + k) * 64] = q[k];
      return;
    }
#pragma unroll
    for (int k = 0; k < kApRec; ++k) sum[(6 * c + k) * 64] += q[k];
  }
  __device__ __forceinline__ void flush(long long* __restrict__ ap) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      write(ap, c);
      id[c] = -1;
    }
  }
};

// One realization's term of a library entry: the fixed-point sum S (or the
// FP64 sum of the trajectory path) over T, the count term also over the
// aperture's area, added into the realization average with weight wt when
// the integration is valid.  Shared by the library finalizers (fused, batched
// and sample-binning kernels), so all round alike.
__device__ __forceinline__ void lib_add(double& out, double sum, double T, bool count, double area,
                                        bool ok, double wt) {
#pragma clang fp contract(off)
  double v = sum / T;
  if (count) v = v / area;
  out += (ok ? v : 0.0) * wt;
}

// Library row of slab entry (aperture a, record term comp): rows
// [Rzphi (P) | counts | h1 | h2 | h3 | h4 | h0] (B each); record terms
// [dt, g dt H0, g dt H1..H4].
__host__ __device__ __forceinline__ int rec_block(int comp) {
  return comp == 0 ? 0 : (comp == 1 ? 5 : comp - 1);
}
__host__ __device__ __forceinline__ int block_comp(int blk) {
  return blk == 0 ? 0 : (blk == 5 ? 1 : blk + 1);
}

// The grid cell and aperture of copy c of a projected sample, with the
// fixed-point record terms when the copy falls in an aperture (a >= 0).
// Shared by the fused integrator and the sample-binning kernel, so both add
// the same integers.  w: the step length; e_t: the integration's exponent
// (fix_exp_time).
__device__ __forceinline__ void copy_bins(const smx_grids& g, const int32_t* __restrict__ bm,
                                          const double* __restrict__ v0,
                                          const double* __restrict__ sd, const SampleProj& sp,
                                          double z, int c, double w, int e_t, int& cell, int& a,
                                          long long (&q)[kApRec]) {
  double zs, phi, xr, yr, vlos;
  copy_coords(sp, c, zs, phi, xr, yr, vlos);
  cell = rzphi_cell(g, sp.R, zs * z, phi);
  a = aperture(g, bm, xr, yr);
  if (a >= 0) {
    double gw, H[5];
    gh_terms(vlos, v0[a], sd[a], gw, H);
    const double gd = gw * w;
    const double sg = fix_scale(e_t + fix_exp_gh(sd[a]));
    q[0] = fix_q(w, fix_scale(e_t));
#pragma unroll
    for (int k = 0; k < 5; ++k) q[1 + k] = fix_q(gd * H[k], sg);
  }
}

// Binning of one accepted sample's NC symmetry copies with weight w: cells
// and aperture records through their register caches.
template <int NC>
__device__ __forceinline__ void bin_sample(const smx_grids& g, const int32_t* __restrict__ bm,
                                           const double* __restrict__ v0,
                                           const double* __restrict__ sd, const double* y,
                                           double w, int e_t, CellCache<NC>& cache,
                                           ApCache<NC>& acache, long long* __restrict__ rz,
                                           long long* __restrict__ ap) {
  SampleProj sp;
  project(g, y, NC > 2, sp);
  const long long qd = fix_q(w, fix_scale(e_t));
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    int cell, a;
    long long q[kApRec];
    copy_bins(g, bm, v0, sd, sp, y[2], c, w, e_t, cell, a, q);
    cache.add(rz, c, cell, qd);
    if (a >= 0) acache.add(ap, c, a, q);
  }
}

// The accepted-step loop of one integration with binning: N_max BS3 attempts
// (integrates.py:181-205), every accepted state binned with weight dt in its
// NC copies; returns the accumulated time (NC x the integrated time).  With
// ckpt, the state (y, dt, t) before every ckpt_every-th attempt is stored at
// ckpt + (attempt / ckpt_every) * ckpt_stride (8 doubles).  e_t: the
// integration's fixed-point exponent (fix_exp_time).  ap_lds: this lane's
// column of a __shared__ array of at least 6 NC x 64 integers (the
// aperture-record cache; the block is one wave of 64 lanes).
template <int NC>
__device__ __forceinline__ double integrate_and_bin(const smx_potential& p, const smx_grids& g,
                                                    const smx_schedule& s, const int32_t* bm,
                                                    const double* v0, const double* sd, double& t,
                                                    double* y, double* k1, double& dt, int e_t,
                                                    long long* __restrict__ rz,
                                                    long long* __restrict__ ap, double* ckpt,
                                                    int64_t ckpt_stride, int ckpt_every,
                                                    long long* ap_lds) {
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
      bin_sample<NC>(g, bm, v0, sd, y, used, e_t, cache, acache, rz, ap);
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
