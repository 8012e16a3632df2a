// Orbit library kernels: potential evaluation, Jeans initial conditions,
// fused BS3 integration + binning, realization averaging, trajectory path,
// the binning test hook and the density target.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include <cstdarg>
#include <mutex>
#include <vector>
#include <string>

#include "common.hpp"
#include "potential.hpp"
#include "binning.hpp"
#include "integrator.hpp"
#include "tuning.hpp"
#include "ics.hpp"
#include "timing.hpp"

namespace smx {

// ---------------------------------------------------------------------------
// error state
// ---------------------------------------------------------------------------
static std::mutex g_err_mu;
static std::string g_err;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  std::lock_guard<std::mutex> lk(g_err_mu);
  g_err = buf;
}

// ---------------------------------------------------------------------------
// potential / density at points
// ---------------------------------------------------------------------------
__global__ void potential_eval_kernel(smx_potential p, const double* __restrict__ xyz, int64_t n,
                                      double* __restrict__ phi, double* __restrict__ acc) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
  if (acc) {
    double ax, ay, az;
    accel_fast(p, x, y, z, ax, ay, az);
    acc[3 * i] = ax;
    acc[3 * i + 1] = ay;
    acc[3 * i + 2] = az;
  }
  if (phi) phi[i] = phi_total(p, x, y, z);
}

__global__ void density_eval_kernel(smx_potential p, const double* __restrict__ xyz, int64_t n,
                                    double* __restrict__ rho) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  rho[i] = rho_total(p, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
}

// ---------------------------------------------------------------------------
// Jeans initial conditions (jeans.py:9-94) + jitter / time scales
// (model.py:117-169).  One 64-lane wavefront per star.
// ---------------------------------------------------------------------------
#ifndef SMX_IC_WAVES
#define SMX_IC_WAVES 2   // waves per SIMD the IC kernel's registers are sized for (2: 10.0 ms, 1: 11.0 ms at plain-20k)
#endif
__global__ __launch_bounds__(256, SMX_IC_WAVES) void initial_conditions_kernel(
    smx_potential p, const double* __restrict__ w0, const double* __restrict__ normals,
    const double* __restrict__ uniforms, int64_t n, int r, int r_total, double N_step,
    double N_dyn, double* __restrict__ wb, double* __restrict__ T_total, double* __restrict__ dt0) {
  const int lane = threadIdx.x & 63;
  const int64_t i = blockIdx.x * (int64_t)(blockDim.x >> 6) + (threadIdx.x >> 6);
  if (i >= n) return;   // whole wave exits together
  StarIC<double> o;
  star_ic<double>(p, fields_const<double>(p), w0, normals, i, lane, o);
  if (lane != 0) return;
  const double scale[6] = {0.1, 0.1, 0.1, o.vsc, o.vsc, o.vsc};
  for (int k = 0; k < r; ++k) {
    const int64_t q = i * r + k;
    for (int j = 0; j < 6; ++j)
      wb[6 * q + j] = o.wn[j] + (uniforms[(i * r_total + k) * 6 + j] - 0.5) * scale[j];
    T_total[q] = o.T_orb * N_dyn;
    dt0[q] = o.T_orb / N_step;
  }
}

// ---------------------------------------------------------------------------
// Fused BS3 integration + binning (integrates.py:37-206), samples accumulated
// into each integration's own slab.  With ckpt, the state (y, dt, t) before
// every ckpt_every-th attempt is stored for the adjoint.
//
// Balanced over the resident waves.  The kernel runs W waves of 64 lanes,
// W = the waves resident at once (one per SIMD at the integrator's register
// count), and every integration makes the same L = N_max attempts.  With
// more integrations than lanes a second round of whole integrations would
// leave most SIMDs idle (80,000 integrations on 65,536 lanes: 1.22 rounds of
// work in 2 rounds of time).  Instead wave w takes the contiguous share of m
// integrations [w n / W, (w + 1) n / W) and spreads their m L attempts over
// its lanes by McNaughton's wrap-around rule: lane l covers attempts
// [l C, (l + 1) C) of the share's concatenation, C = max(L, ceil(m L / 64)).
// An integration cut at a lane boundary runs its first part (its head) on
// lane l + 1 at the start of the loop and its second part (its tail) on lane
// l at the end of the loop, ending at step C, so the two parts never overlap
// (C >= L); the state in between passes through LDS (the workgroup is the
// one wave, which keeps its LDS operations in order).  Each lane runs, in
// loop steps: its head [0, hb), its whole integrations from hb on, its tail
// [C - (L - ta), C).  Every integration's attempts, binning and checkpoints
// are those of integrate_and_bin, in the same order, into its own slab (its
// fixed-point sums do not depend on how the register caches group them at a
// head / tail boundary).  In sample mode the integrator does not bin: every
// attempt stores its state and step length (0 when rejected) in the
// integration's sample rows, and bin_library_kernel bins them.
// ---------------------------------------------------------------------------
#ifndef SMX_INTEGRATOR_WAVES
#define SMX_INTEGRATOR_WAVES 1   // waves per SIMD the integrator's registers are sized for
#endif

struct IntegState {
  double y[6], k1[6];
  double t, dt, T_acc, EJ0;
  int64_t job;       // integration index, or -1
  int64_t attempt;   // attempts made so far
  int e_t;           // fixed-point exponent of the integration (fix_exp_time)
};

// Sample rows: integration i's attempt a at samples + (i N_max + a) * kSampleW,
// [y (6), step length (0 when rejected), unused]; one 64-byte line per attempt.
constexpr int kSampleW = 8;

// Bytes of the sample rows of n_int integrations (sample mode).
static int64_t sample_bytes(const smx_schedule* s, int64_t n_int) {
  return n_int * (int64_t)(s->N_max > 0 ? s->N_max : 0) * kSampleW * (int64_t)sizeof(double);
}
constexpr int kHandoff = 16;   // doubles per handed-over state

template <int NC, bool SAMPLES>
__device__ __forceinline__ void integrate_balanced(
    const smx_potential& p, const smx_grids& g, const smx_schedule& s,
    const double* __restrict__ wb, const double* __restrict__ T_total,
    const double* __restrict__ dt0, int64_t n_int, int64_t W, const int32_t* __restrict__ bm,
    const double* __restrict__ v0, const double* __restrict__ sd, long long* __restrict__ ws,
    double* __restrict__ samples, double* __restrict__ valid_out, double* __restrict__ T_out,
    double* __restrict__ tfin_out, double* __restrict__ ckpt, int ckpt_every, long long* ap_col,
    double* handoff) {
  const int lane = threadIdx.x;
  const int64_t w = blockIdx.x;
  const int64_t j0 = w * n_int / W, m = (w + 1) * n_int / W - j0;
  const int64_t L = s.N_max;
  if (m <= 0 || L <= 0) {
    // no attempts: every integration of the share is its initial state
    for (int64_t j = lane; j < m; j += 64) {
      const int64_t i = j0 + j;
      valid_out[i] = (0.0 > T_total[i] / 10.0) ? 1.0 : 0.0;
      T_out[i] = kEps;
      tfin_out[i] = 0.0;
    }
    return;
  }
  LaneSchedule ls;
  ls.init(lane, m, L);
  const int nseg = ls.nseg;

  const int P = g.nR * g.nz * g.nphi;
  const int64_t stride = slab_stride(P, g.num_vbin);
  const int64_t ckpt_stride = n_int * 8;
  IntegState st;
  st.job = -1;
  st.attempt = 0;
  CellCache<NC> cache;
  ApCache<NC> acache;
  cache.init();
  acache.init(ap_col);
  long long* rz = nullptr;
  long long* ap = nullptr;
  st.e_t = 0;
  int sidx = 0, kind = 0;
  int64_t job = 0, first = 0, t0 = 0, t1 = 0;
  if (nseg > 0) ls.segment(0, kind, job, first, t0, t1);

  for (int64_t step = 0; step <= ls.C; ++step) {
    // segment ends (a head hands its state over; a whole integration or a
    // tail is finished) for every lane, then starts, so that a tail that
    // starts at the step its head ends reads the handed-over state
    const bool ends = sidx < nseg && st.job >= 0 && step == t1;
    const bool starts_now = sidx < nseg && st.job < 0 && step == t0;
    if (__ballot(ends || starts_now) != 0) {   // wave-uniform
      bool start = starts_now;
      if (ends) {
        if (!SAMPLES) {
          cache.flush(rz);
          acache.flush(ap);
        }
        const int64_t i = j0 + st.job;
        if (kind == 0) {
          double* h = handoff + lane;
#pragma unroll
          for (int q = 0; q < 6; ++q) {
            h[q * 64] = st.y[q];
            h[(6 + q) * 64] = st.k1[q];
          }
          h[12 * 64] = st.t;
          h[13 * 64] = st.dt;
          h[14 * 64] = st.T_acc;
          h[15 * 64] = st.EJ0;
        } else {
          const double EJ = jacobi(p, s.Omega, st.y);
          const bool ok = (fabs(EJ / st.EJ0 - 1.0) < 0.1) && (st.t > T_total[i] / 10.0);
          valid_out[i] = ok ? 1.0 : 0.0;
          T_out[i] = st.T_acc + kEps;
          tfin_out[i] = st.t;
        }
        st.job = -1;
        ++sidx;
        if (sidx < nseg) {
          ls.segment(sidx, kind, job, first, t0, t1);
          start = step == t0;
        }
      }
      // handed-over states written above are read below by the next lane
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
      if (start) {
        const int64_t i = j0 + job;
        if (kind == 2) {
          const double* h = handoff + (lane - 1);
#pragma unroll
          for (int q = 0; q < 6; ++q) {
            st.y[q] = h[q * 64];
            st.k1[q] = h[(6 + q) * 64];
          }
          st.t = h[12 * 64];
          st.dt = h[13 * 64];
          st.T_acc = h[14 * 64];
          st.EJ0 = h[15 * 64];
        } else {
#pragma unroll
          for (int q = 0; q < 6; ++q) st.y[q] = wb[6 * i + q];
          deriv(p, s.Omega, st.y, st.k1);
          st.EJ0 = jacobi(p, s.Omega, st.y);
          st.t = 0.0;
          st.dt = dt0[i];
          st.T_acc = 0.0;
        }
        st.job = job;
        st.attempt = first;
        st.e_t = fix_exp_time(dt0[i], s, NC);
        if (!SAMPLES) {
          rz = ws + i * stride;
          ap = rz + slab_rz_len(P);
        }
      }
    }
    if (st.job >= 0 && step < ls.C) {
      const int64_t i = j0 + st.job;
      if (ckpt != nullptr && st.attempt % ckpt_every == 0) {
        double* c = ckpt + (st.attempt / ckpt_every) * ckpt_stride + i * 8;
#pragma unroll
        for (int q = 0; q < 6; ++q) c[q] = st.y[q];
        c[6] = st.dt;
        c[7] = st.t;
      }
      const double used = bs3_attempt(p, s, st.t, st.y, st.k1, st.dt);
      if (SAMPLES) {
        double* row = samples + (i * L + st.attempt) * kSampleW;
        *reinterpret_cast<double2*>(row) = make_double2(st.y[0], st.y[1]);
        *reinterpret_cast<double2*>(row + 2) = make_double2(st.y[2], st.y[3]);
        *reinterpret_cast<double2*>(row + 4) = make_double2(st.y[4], st.y[5]);
        *reinterpret_cast<double2*>(row + 6) = make_double2(used, 0.0);
        if (used != 0.0) st.T_acc += NC * used;
      } else if (used != 0.0) {
        bin_sample<NC>(g, bm, v0, sd, st.y, used, st.e_t, cache, acache, rz, ap);
        st.T_acc += NC * used;
      }
      ++st.attempt;
    }
  }
}

// SAMPLES: store every attempt's state in `samples` instead of binning into
// the slabs `ws` (one of the two is null).
template <bool SAMPLES>
__global__ __launch_bounds__(64, SMX_INTEGRATOR_WAVES) void integrate_library_kernel(
    smx_potential p, smx_grids g, smx_schedule s, const double* __restrict__ wb,
    const double* __restrict__ T_total, const double* __restrict__ dt0, int64_t n_int,
    const int32_t* __restrict__ bm, const double* __restrict__ v0, const double* __restrict__ sd,
    long long* __restrict__ ws, double* __restrict__ samples, double* __restrict__ valid_out,
    double* __restrict__ T_out, double* __restrict__ tfin_out, double* __restrict__ ckpt,
    int ckpt_every) {
  __shared__ long long ap_lds[SAMPLES ? 1 : 24 * 64];   // aperture-record cache, 4 copies x 6 terms
  __shared__ double handoff[kHandoff * 64];            // head -> tail states, one per lane
  const int64_t W = gridDim.x;
  if (g.symmetry)
    integrate_balanced<4, SAMPLES>(p, g, s, wb, T_total, dt0, n_int, W, bm, v0, sd, ws, samples,
                                   valid_out, T_out, tfin_out, ckpt, ckpt_every,
                                   ap_lds + (SAMPLES ? 0 : threadIdx.x), handoff);
  else
    integrate_balanced<1, SAMPLES>(p, g, s, wb, T_total, dt0, n_int, W, bm, v0, sd, ws, samples,
                                   valid_out, T_out, tfin_out, ckpt, ckpt_every,
                                   ap_lds + (SAMPLES ? 0 : threadIdx.x), handoff);
}

// Waves of the balanced integrator: the waves resident at once when there
// are more integrations than their lanes, else one lane per integration.
template <bool SAMPLES>
static int64_t integrator_waves(int64_t n_int) {
  static int resident = 0;
  static std::mutex mu;
  {
    std::lock_guard<std::mutex> lock(mu);
    if (resident == 0) {
      int dev = 0, cus = 0, per_cu = 0;
      if (hipGetDevice(&dev) == hipSuccess &&
          hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
          hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, integrate_library_kernel<SAMPLES>,
                                                       64, 0) == hipSuccess &&
          cus > 0 && per_cu > 0)
        resident = cus * per_cu;
      else
        resident = 1024;
    }
  }
  const int64_t lanes = (n_int + 63) / 64;
  return lanes < resident ? lanes : resident;
}

// lib[c, :] = sum_r valid_r * (slab_r / T_r) / (sum_r valid_r + 0.1)
// (integrates.py:212-236, 287-300); output row order
// [Rzphi | surface density | h1 | h2 | h3 | h4 | h0].
// One work item per (column, cell) and per (column, aperture): an aperture
// item reads its 48-byte record once per realization and writes the six
// aperture rows (counts, h1..h4, h0), rather than six items each reading one
// value of the record (whose other values then come back from beyond L2 for
// the other five).  FIXED: fixed-point slabs of the library paths (scales
// from dt0, the schedule and the apertures' dispersions); otherwise FP64
// slabs of the trajectory path, already divided by T.
template <bool FIXED>
__global__ void finalize_library_kernel(const void* __restrict__ ws_, int64_t n, int r, int P,
                                        int B, const double* __restrict__ valid,
                                        const double* __restrict__ T_int,
                                        const double* __restrict__ num_per_bin, double area_pixel,
                                        const double* __restrict__ dt0, smx_schedule s, int NC,
                                        const double* __restrict__ sd, double* __restrict__ lib) {
  const long long* wsq = static_cast<const long long*>(ws_);
  const double* wsd = static_cast<const double*>(ws_);
  const int64_t rows = (int64_t)P + 6 * (int64_t)B;
  const int64_t items = (int64_t)P + (int64_t)B;
  const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= n * items) return;
  const int64_t c = idx / items;
  const int64_t j = idx - c * items;
  const int64_t stride = slab_stride(P, B);
  double nv = 0.0;
  for (int k = 0; k < r; ++k) nv += valid[c * r + k];
  const double wt = 1.0 / (nv + 0.1);
  if (j < P) {
    double out = 0.0;
    for (int k = 0; k < r; ++k) {
      const int64_t it = c * r + k;
      const int64_t e = it * stride + j;
      const double sum = FIXED ? static_cast<double>(wsq[e]) * fix_unscale(fix_exp_time(dt0[it], s, NC))
                               : wsd[e];
      lib_add(out, sum, FIXED ? T_int[it] : 1.0, false, 1.0, valid[it] > 0.5, wt);
    }
    lib[c * rows + j] = out;
    return;
  }
  const int64_t a = j - P;
  const double area = aperture_area(num_per_bin[a], area_pixel);
  const int e_g = FIXED ? fix_exp_gh(sd[a]) : 0;
  double out[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int k = 0; k < r; ++k) {
    const int64_t it = c * r + k;
    const int64_t e = it * stride + slab_rz_len(P) + kApRec * a;
    const int e_t = FIXED ? fix_exp_time(dt0[it], s, NC) : 0;
    const bool ok = valid[it] > 0.5;
#pragma unroll
    for (int blk = 0; blk < 6; ++blk) {
      const int comp = block_comp(blk);
      const double sum = FIXED ? static_cast<double>(wsq[e + comp]) *
                                     fix_unscale(comp == 0 ? e_t : e_t + e_g)
                               : wsd[e + comp];
      lib_add(out[blk], sum, FIXED ? T_int[it] : 1.0, blk == 0, area, ok, wt);
    }
  }
#pragma unroll
  for (int blk = 0; blk < 6; ++blk) lib[c * rows + P + blk * (int64_t)B + a] = out[blk];
}

// ---------------------------------------------------------------------------
// Binning of stored samples into LDS-privatised histograms, one workgroup per
// orbit column (integrates.py:115-174, 207-238, 287-300).  The workgroup's
// LDS holds a window of the column's library rows as fixed-point sums (the
// whole row set, 146 KB, at 12,288 cells and 1,000 apertures; larger grids
// take several windows, each pass re-binning the samples).  For each valid
// realization in order: the window is cleared, every thread bins its share
// of the integration's accepted samples (projection, four symmetry copies,
// cell and aperture ids, Gauss-Hermite terms) and adds their fixed-point
// contributions with 64-bit LDS atomics, then each thread converts the
// entries it owns and adds them into its register copy of the realization
// average, in realization order as finalize_library_kernel does; the rows
// are written once at the end.  The integer sums equal the fused
// integrator's slab sums bit for bit, so the library does too.
// Thread (wave w, lane l) takes attempts l L + j, j = w, w + waves, ...
// (L = ceil(N / 64)): a wave's lanes are L attempts apart on the orbit, in
// different cells mostly, so its atomics rarely collide.
// ---------------------------------------------------------------------------
constexpr int BL_THREADS = 512;
constexpr int BL_PER_THREAD = 40;   // entries per thread: windows of up to 20,480 rows (160 KB)

template <int NC>
__global__ __launch_bounds__(BL_THREADS) void bin_library_kernel(
    smx_grids g, smx_schedule s, const double* __restrict__ samples, int64_t n, int r,
    const double* __restrict__ dt0, const double* __restrict__ valid,
    const double* __restrict__ T_int, const int32_t* __restrict__ bm,
    const double* __restrict__ v0, const double* __restrict__ sd,
    const double* __restrict__ num_per_bin, int64_t win, int stage, double* __restrict__ lib,
    long long* __restrict__ slabs) {
  extern __shared__ long long hist[];
  const int64_t c = blockIdx.x;
  const int tid = threadIdx.x;
  const int P = g.nR * g.nz * g.nphi;
  const int B = g.num_vbin;
  // the apertures' v0 and s in LDS after the window when they fit: the
  // binning's gathers of them then stay on the CU
  const double* v0p = v0;
  const double* sdp = sd;
  if (stage) {
    double* sv = reinterpret_cast<double*>(hist + win);
    for (int a = tid; a < B; a += BL_THREADS) {
      sv[a] = v0[a];
      sv[B + a] = sd[a];
    }
    __syncthreads();
    v0p = sv;
    sdp = sv + B;
  }
  const int64_t rows = (int64_t)P + 6 * (int64_t)B;
  const int64_t N = s.N_max;
  const int64_t Ls = (N + 63) / 64;
  const int lane = tid & 63, wave = tid >> 6;
  double nv = 0.0;
  for (int k = 0; k < r; ++k) nv += valid[c * r + k];
  const double wt = 1.0 / (nv + 0.1);
  for (int64_t e0 = 0; e0 < rows; e0 += win) {
    const int64_t e1 = rows < e0 + win ? rows : e0 + win;
    double out[BL_PER_THREAD];
#pragma unroll
    for (int q = 0; q < BL_PER_THREAD; ++q) out[q] = 0.0;
    for (int k = 0; k < r; ++k) {
      const int64_t it = c * r + k;
      if (!(valid[it] > 0.5)) continue;   // uniform over the workgroup
      for (int64_t e = e0 + tid; e < e1; e += BL_THREADS) hist[e - e0] = 0;
      __syncthreads();
      const int e_t = fix_exp_time(dt0[it], s, NC);
      const double sc_t = fix_scale(e_t);
      const double* base = samples + it * N * kSampleW;
      // attempts l Ls + j of this lane, the next one's row loaded while the
      // current one is binned
      constexpr int WV = BL_THREADS / 64;
      double2 cur[4], nxt[4];
      auto load_row = [&](int64_t jj, double2 (&r)[4]) {
        const int64_t a_idx = lane * Ls + jj;
        if (a_idx < N) {
          const double2* row = reinterpret_cast<const double2*>(base + a_idx * kSampleW);
#pragma unroll
          for (int h = 0; h < 4; ++h) r[h] = row[h];
        } else {
          r[3] = make_double2(0.0, 0.0);   // no attempt: step length 0
        }
      };
      int64_t j = wave;
      if (j < Ls) load_row(j, cur);
      for (; j < Ls; j += WV) {   // wave-uniform
        if (j + WV < Ls) load_row(j + WV, nxt);
        const double used = cur[3].x;
        if (used != 0.0) {
          const double y[6] = {cur[0].x, cur[0].y, cur[1].x, cur[1].y, cur[2].x, cur[2].y};
          SampleProj sp;
          project(g, y, NC > 2, sp);
          const long long qd = fix_q(used, sc_t);
#pragma unroll
          for (int cp = 0; cp < NC; ++cp) {
            int cell, a;
            long long q[kApRec];
            copy_bins(g, bm, v0p, sdp, sp, y[2], cp, used, e_t, cell, a, q);
            if (cell >= e0 && cell < e1)
              atomicAdd(reinterpret_cast<unsigned long long*>(&hist[cell - e0]),
                        static_cast<unsigned long long>(qd));
            if (a >= 0) {
#pragma unroll
              for (int comp = 0; comp < kApRec; ++comp) {
                const int64_t e = P + rec_block(comp) * (int64_t)B + a;
                if (e >= e0 && e < e1)
                  atomicAdd(reinterpret_cast<unsigned long long*>(&hist[e - e0]),
                            static_cast<unsigned long long>(q[comp]));
              }
            }
          }
        }
#pragma unroll
        for (int h = 0; h < 4; ++h) cur[h] = nxt[h];
      }
      __syncthreads();
      if (slabs != nullptr) {
        // the integration's sums in the fused integrator's slab layout, for
        // the adjoint's T-cotangent pre-pass (valid integrations only: it
        // reads no other)
        long long* sl = slabs + it * slab_stride(P, B);
        for (int64_t e = e0 + tid; e < e1; e += BL_THREADS) {
          int64_t o = e;
          if (e >= P) {
            const int64_t blk = (e - P) / B, a = (e - P) - blk * B;
            o = slab_rz_len(P) + kApRec * a + block_comp((int)blk);
          }
          sl[o] = hist[e - e0];
        }
      }
      const double T = T_int[it];
      const double un_t = fix_unscale(e_t);
#pragma unroll
      for (int q = 0; q < BL_PER_THREAD; ++q) {
        // the entry's offset, opaque to the compiler: otherwise it hoists the
        // per-entry areas and exponents out of the realization loop and keeps
        // all of them live through the binning
        int off = tid + q * BL_THREADS;
        asm volatile("" : "+v"(off));
        const int64_t e = e0 + off;
        // an empty entry adds +0 (S = 0 gives v = +0, and the average is
        // never -0), so it is skipped: most cells see no sample of one orbit
        const long long Sq = e < e1 ? hist[e - e0] : 0;
        if (Sq != 0) {
          const double S = static_cast<double>(Sq);
          if (e < P) {
            lib_add(out[q], S * un_t, T, false, 1.0, true, wt);
          } else {
            const int64_t blk = (e - P) / B, a = (e - P) - blk * B;
            const double un = blk == 0 ? un_t : fix_unscale(e_t + fix_exp_gh(sdp[a]));
            lib_add(out[q], S * un, T, blk == 0, aperture_area(num_per_bin[a], g.area_pixel), true, wt);
          }
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < BL_PER_THREAD; ++q) {
      const int64_t e = e0 + tid + (int64_t)q * BL_THREADS;
      if (e < e1) lib[c * rows + e] = out[q];
    }
  }
}

// Trajectory path (integrates.py:319-512): integrate storing every attempt's
// state and dt; binning happens after, with dt / T.
__global__ __launch_bounds__(64) void integrate_traj_kernel(
    smx_potential p, smx_schedule s, const double* __restrict__ wb,
    const double* __restrict__ T_total, const double* __restrict__ dt0, int64_t n_int,
    double* __restrict__ traj, double* __restrict__ t_traj, double* __restrict__ valid_out,
    double* __restrict__ T_out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n_int) return;
  double y[6], k1[6];
#pragma unroll
  for (int j = 0; j < 6; ++j) y[j] = wb[6 * i + j];
  deriv(p, s.Omega, y, k1);
  const double EJ0 = jacobi(p, s.Omega, y);
  double t = 0.0, dt = dt0[i], csum = 0.0;
  double* tr = traj + i * (int64_t)s.N_max * 6;
  double* tt = t_traj + i * (int64_t)s.N_max;
  for (int step = 0; step < s.N_max; ++step) {
    const double used = bs3_attempt(p, s, t, y, k1, dt);
#pragma unroll
    for (int j = 0; j < 6; ++j) tr[6 * step + j] = y[j];
    csum += used;
    tt[step] = used;     // dt_traj; turned into its cumulative sum by bin_traj_kernel
  }
  const double EJ = jacobi(p, s.Omega, y);
  const bool ok = (fabs(EJ / EJ0 - 1.0) < 0.1) && (t > T_total[i] / 10.0);
  valid_out[i] = ok ? 1.0 : 0.0;
  T_out[i] = csum + kEps;
}

// Post-scan binning of stored trajectories, no symmetry, weights dt / T
// (integrates.py:438-501); converts velocities to km/s afterwards (426-429).
__global__ __launch_bounds__(64) void bin_traj_kernel(
    smx_grids g, int64_t n_int, int N, double* __restrict__ traj, double* __restrict__ t_traj,
    const double* __restrict__ T_int, const int32_t* __restrict__ bm, const double* __restrict__ v0,
    const double* __restrict__ sd, double* __restrict__ ws) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n_int) return;
  const int P = g.nR * g.nz * g.nphi;
  double* rz = ws + i * slab_stride(P, g.num_vbin);
  double* ap = rz + slab_rz_len(P);
  double* tr = traj + i * (int64_t)N * 6;
  double* tt = t_traj + i * (int64_t)N;
  const double T = T_int[i];
  double csum = 0.0;
  for (int step = 0; step < N; ++step) {
    const double used = tt[step];
    csum += used;
    tt[step] = csum;
    double y[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) y[j] = tr[6 * step + j];
    if (used != 0.0) {
      SampleProj sp;
      project(g, y, false, sp);
      double zs, phi, xr, yr, vlos;
      copy_coords(sp, 0, zs, phi, xr, yr, vlos);
      const double w = used / T;
      const int cell = rzphi_cell(g, sp.R, y[2], phi);
      if (cell >= 0) rz[cell] += w;
      const int a = aperture(g, bm, xr, yr);
      if (a >= 0) {
        double gw, H[5];
        gh_terms(vlos, v0[a], sd[a], gw, H);
        double* rec = ap + kApRec * (int64_t)a;
        rec[0] += w;
        const double wg = w * gw;
#pragma unroll
        for (int k = 0; k < 5; ++k) rec[1 + k] += wg * H[k];
      }
    }
#pragma unroll
    for (int j = 3; j < 6; ++j) tr[6 * step + j] = y[j] * kKpcGyrToKms;
  }
}

__global__ void fill_kernel(double* __restrict__ p, int64_t n, double v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Bit-exactness hook: one histogram over externally supplied samples.
__global__ void bin_samples_kernel(smx_grids g, const double* __restrict__ samples,
                                   const double* __restrict__ dt, int64_t S,
                                   const int32_t* __restrict__ bm, const double* __restrict__ v0,
                                   const double* __restrict__ sd, int32_t* __restrict__ idx_rz,
                                   int32_t* __restrict__ idx_ap, double* __restrict__ acc_rz,
                                   double* __restrict__ acc_ap) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= S) return;
  const int nsym = g.symmetry ? 4 : 1;
  double y[6];
  for (int j = 0; j < 6; ++j) y[j] = samples[6 * i + j];
  SampleProj sp;
  project(g, y, nsym > 2, sp);
  const double w = dt[i];
  for (int c = 0; c < nsym; ++c) {
    double zs, phi, xr, yr, vlos;
    copy_coords(sp, c, zs, phi, xr, yr, vlos);
    const int cell = rzphi_cell(g, sp.R, zs * y[2], phi);
    const int a = aperture(g, bm, xr, yr);
    idx_rz[i * nsym + c] = cell;
    idx_ap[i * nsym + c] = a;
    if (cell >= 0) atomicAdd(acc_rz + cell, w);
    if (a >= 0) {
      double gw, H[5];
      gh_terms(vlos, v0[a], sd[a], gw, H);
      atomicAdd(acc_ap + 6 * (int64_t)a, w);
      const double gd = gw * w;
      for (int k = 0; k < 5; ++k) atomicAdd(acc_ap + 6 * (int64_t)a + 1 + k, gd * H[k]);
    }
  }
}

// Cell masses by quasi-Monte-Carlo (model.py:192-213): one block per cell.
__global__ __launch_bounds__(256) void density_target_kernel(
    smx_potential p, const double* __restrict__ Rg, const double* __restrict__ zg,
    const double* __restrict__ pg, int64_t P, double dR, double dz, double dphi,
    const double* __restrict__ smp, int64_t S, double* __restrict__ out) {
  __shared__ double red[4];
  const int64_t c = blockIdx.x;
  if (c >= P) return;
  double acc = cell_mass_part<double>(p, fields_const<double>(p), Rg[c], zg[c], pg[c], dR, dz, dphi,
                                     smp, S, threadIdx.x, blockDim.x);
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) tot += red[w];
    out[c] = ((tot / static_cast<double>(S)) * dR * dz) * dphi;
  }
}

}  // namespace smx

using namespace smx;

// Host replay of the balanced schedule's events (the kernel's loop without
// the integration): every integration of every wave must make attempts
// 0 .. L-1 exactly once and in order, a tail must start at or after its head
// has handed over, no lane may run two segments at once, and every
// integration is finished once.  Returns 0 when all hold; *steps_out = the
// loop length C of the first wave.
int smx_integrator_schedule_check(int64_t n_int, int64_t waves, int64_t L, int64_t* steps_out) {
  SMX_REQUIRE(n_int >= 0 && waves >= 1 && L >= 1 && steps_out,
              "smx_integrator_schedule_check: bad arguments");
  std::vector<int64_t> next(n_int, 0), done(n_int, 0);
  *steps_out = 0;
  for (int64_t w = 0; w < waves; ++w) {
    const int64_t j0 = w * n_int / waves, m = (w + 1) * n_int / waves - j0;
    if (m <= 0) continue;
    LaneSchedule ls[64];
    int sidx[64], kind[64];
    int64_t job[64], first[64], t0[64], t1[64], cur[64], att[64], hand_job[64], hand_att[64];
    for (int l = 0; l < 64; ++l) {
      ls[l].init(l, m, L);
      sidx[l] = 0;
      cur[l] = -1;
      hand_job[l] = -1;
      if (ls[l].nseg > 0) ls[l].segment(0, kind[l], job[l], first[l], t0[l], t1[l]);
    }
    const int64_t C = ls[0].C;
    if (w == 0) *steps_out = C;
    for (int64_t step = 0; step <= C; ++step) {
      bool start[64];
      for (int l = 0; l < 64; ++l) {
        const bool ends = sidx[l] < ls[l].nseg && cur[l] >= 0 && step == t1[l];
        start[l] = sidx[l] < ls[l].nseg && cur[l] < 0 && step == t0[l];
        if (ends) {
          if (kind[l] == 0) {
            hand_job[l] = cur[l];
            hand_att[l] = att[l];
          } else {
            if (att[l] != L || done[j0 + cur[l]]++ != 0) return 1;
          }
          cur[l] = -1;
          if (++sidx[l] < ls[l].nseg) {
            ls[l].segment(sidx[l], kind[l], job[l], first[l], t0[l], t1[l]);
            start[l] = step == t0[l];
          }
        }
      }
      for (int l = 0; l < 64; ++l) {
        if (!start[l]) continue;
        if (kind[l] == 2 && (l == 0 || hand_job[l - 1] != job[l] || hand_att[l - 1] != first[l]))
          return 2;
        cur[l] = job[l];
        att[l] = first[l];
      }
      for (int l = 0; l < 64; ++l) {
        if (cur[l] < 0 || step >= C) continue;
        int64_t& nx = next[j0 + cur[l]];
        if (att[l] != nx) return 3;
        ++nx;
        ++att[l];
      }
    }
  }
  for (int64_t i = 0; i < n_int; ++i)
    if (next[i] != L || done[i] != 1) return 4;
  return 0;
}


extern "C" {

int smx_last_error(char* buf, int64_t len) {
  std::lock_guard<std::mutex> lk(g_err_mu);
  if (buf && len > 0) {
    std::strncpy(buf, g_err.c_str(), static_cast<size_t>(len - 1));
    buf[len - 1] = 0;
  }
  return static_cast<int>(g_err.size());
}

int smx_version(void) { return 1; }

int smx_potential_eval(const smx_potential* p, const double* xyz, int64_t n, double* phi,
                       double* acc, void* stream) {
  SMX_REQUIRE(p && xyz && n >= 0, "smx_potential_eval: bad arguments");
  if (n == 0) return 0;
  hipLaunchKernelGGL(potential_eval_kernel, dim3(blocks_for(n, 128)), dim3(128), 0,
                     as_stream(stream), *p, xyz, n, phi, acc);
  SMX_LAUNCHED();
  return 0;
}

int smx_density_eval(const smx_potential* p, const double* xyz, int64_t n, double* rho,
                     void* stream) {
  SMX_REQUIRE(p && xyz && rho && n >= 0, "smx_density_eval: bad arguments");
  if (n == 0) return 0;
  hipLaunchKernelGGL(density_eval_kernel, dim3(blocks_for(n, 128)), dim3(128), 0,
                     as_stream(stream), *p, xyz, n, rho);
  SMX_LAUNCHED();
  return 0;
}

int smx_initial_conditions(const smx_potential* p, const double* w0, const double* normals,
                           const double* uniforms, int64_t n, int32_t r, int32_t r_total,
                           double N_step_per_orb, double N_dyn, double* wb, double* T_total,
                           double* dt0, void* stream) {
  SMX_REQUIRE(p && w0 && normals && uniforms && wb && T_total && dt0,
              "smx_initial_conditions: null pointer");
  SMX_REQUIRE(n >= 0 && r >= 1 && r <= r_total, "smx_initial_conditions: bad sizes");
  if (n == 0) return 0;
  const int waves_per_block = 4;
  stage_begin(ST_ICS, as_stream(stream));
  hipLaunchKernelGGL(initial_conditions_kernel, dim3(blocks_for(n, waves_per_block)),
                     dim3(64 * waves_per_block), 0, as_stream(stream), *p, w0, normals, uniforms,
                     n, r, r_total, N_step_per_orb, N_dyn, wb, T_total, dt0);
  SMX_LAUNCHED();
  stage_end(ST_ICS, as_stream(stream), 1);
  return 0;
}

int64_t smx_library_workspace_bytes(const smx_grids* g, int64_t n_int) {
  if (!g) return -1;
  return slab_stride(g->nR * g->nz * g->nphi, g->num_vbin) * n_int * (int64_t)sizeof(double);
}

int64_t smx_library_ckpt_bytes(const smx_grids* g, const smx_schedule* s, int64_t n_int) {
  if (!g || !s || n_int < 0) return -1;
  return smx_library_workspace_bytes(g, n_int) + sample_bytes(s, n_int);
}

int64_t smx_library_build_bytes(const smx_grids* g, const smx_schedule* s, int64_t n_int) {
  if (!g || !s || n_int < 0) return -1;
  const int64_t slabs = smx_library_workspace_bytes(g, n_int);
  const int64_t smp = sample_bytes(s, n_int);
  return slabs > smp ? slabs : smp;
}

static int check_grids(const smx_grids* g) {
  SMX_REQUIRE(g->nR > 0 && g->nz > 0 && g->nphi > 0 && g->nX > 0 && g->nY > 0 && g->num_vbin > 0,
              "grid sizes must be positive");
  return 0;
}

// Rows per bin_library_kernel window: the whole row set when it fits the
// workgroup's LDS, else as many as fit (at most BL_THREADS x BL_PER_THREAD).
static int library_window(int64_t rows, int64_t B, int64_t* win, int* stage, size_t* bytes) {
  static int lds = 0;
  static std::mutex mu;
  std::lock_guard<std::mutex> lock(mu);
  if (lds == 0) {
    int dev = 0, v = 0;
    SMX_HIP(hipGetDevice(&dev));
    SMX_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    SMX_REQUIRE(v >= 8 * BL_THREADS, "bin_library_kernel: %d bytes of LDS per workgroup", v);
    SMX_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(bin_library_kernel<1>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, v));
    SMX_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(bin_library_kernel<4>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, v));
    lds = v;
  }
  int64_t cap = lds / (int64_t)sizeof(long long);
  if (cap > (int64_t)BL_THREADS * BL_PER_THREAD) cap = (int64_t)BL_THREADS * BL_PER_THREAD;
  *win = rows < cap ? rows : cap;
  const int64_t need = (*win + 2 * B) * (int64_t)sizeof(double);
  *stage = need <= lds ? 1 : 0;
  *bytes = (size_t)(*stage ? need : *win * (int64_t)sizeof(long long));
  return 0;
}

// The library of n columns x r realizations.  samples: the sample mode
// (integrate_library_kernel<true>, then bin_library_kernel), whose workspace
// holds the sample rows; otherwise the fused integrator with fixed-point
// slabs in the workspace, then finalize_library_kernel.  Both give the same
// bits.  With checkpoints (the adjoint's) the integrations' slabs are kept
// for the T-cotangent pre-pass: the binning kernel writes them too, the
// sample rows following them in the workspace.
static int build_library(const smx_potential* p, const smx_grids* g, const smx_schedule* s,
                         const double* wb, const double* T_total, const double* dt0, int64_t n,
                         int32_t r, const int32_t* bin_mapping, const double* num_per_bin,
                         const double* v0, const double* sdisp, void* workspace,
                         int64_t workspace_bytes, double* lib, double* valid, double* T_int,
                         double* t_final, double* ckpt, int32_t ckpt_every, bool samples,
                         void* stream) {
  SMX_REQUIRE(p && g && s && wb && T_total && dt0 && bin_mapping && num_per_bin && v0 && sdisp &&
                  workspace && lib && valid && T_int && t_final,
              "smx_build_library: null pointer");
  SMX_REQUIRE(n >= 0 && r >= 1 && s->N_max >= 0, "smx_build_library: bad sizes");
  SMX_REQUIRE(ckpt == nullptr || ckpt_every >= 1, "smx_build_library_ckpt: ckpt_every must be >= 1");
  if (int e = check_grids(g)) return e;
  if (n == 0) return 0;
  const int64_t n_int = n * r;
  const int P = g->nR * g->nz * g->nphi;
  const int B = g->num_vbin;
  const int NC = g->symmetry ? 4 : 1;
  // the sample binning when the workspace holds its rows (after the slabs
  // with checkpoints), else the fused form when it holds the slabs
  const int64_t slab_need = smx_library_workspace_bytes(g, n_int);
  const int64_t smp_need = (ckpt ? slab_need : 0) + sample_bytes(s, n_int);
  samples = samples && workspace_bytes >= smp_need;
  SMX_REQUIRE(samples || workspace_bytes >= slab_need,
              "smx_build_library: workspace of %lld bytes, the fused form needs %lld and the "
              "sample binning %lld", (long long)workspace_bytes, (long long)slab_need,
              (long long)smp_need);
  hipStream_t st = as_stream(stream);
  if (samples) {
    int64_t win = 0;
    int stage = 0;
    size_t lds_bytes = 0;
    if (int e = library_window((int64_t)P + 6 * (int64_t)B, B, &win, &stage, &lds_bytes)) return e;
    // with checkpoints the workspace holds the slabs (kept for the adjoint)
    // and then the sample rows
    long long* slabs = ckpt ? static_cast<long long*>(workspace) : nullptr;
    double* smp = ckpt ? reinterpret_cast<double*>(slabs + slab_stride(P, B) * n_int)
                       : static_cast<double*>(workspace);
    stage_begin(ST_INTEGRATE, st);
    hipLaunchKernelGGL(integrate_library_kernel<true>, dim3((unsigned)integrator_waves<true>(n_int)),
                       dim3(64), 0, st, *p, *g, *s, wb, T_total, dt0, n_int, bin_mapping, v0, sdisp,
                       nullptr, smp, valid, T_int, t_final, ckpt, ckpt ? ckpt_every : 1);
    SMX_LAUNCHED();
    stage_end(ST_INTEGRATE, st, 1);
    stage_begin(ST_FINALIZE, st);
    if (NC == 4)
      hipLaunchKernelGGL(bin_library_kernel<4>, dim3((unsigned)n), dim3(BL_THREADS), lds_bytes, st,
                         *g, *s, smp, n, (int)r, dt0, valid, T_int, bin_mapping, v0, sdisp,
                         num_per_bin, win, stage, lib, slabs);
    else
      hipLaunchKernelGGL(bin_library_kernel<1>, dim3((unsigned)n), dim3(BL_THREADS), lds_bytes, st,
                         *g, *s, smp, n, (int)r, dt0, valid, T_int, bin_mapping, v0, sdisp,
                         num_per_bin, win, stage, lib, slabs);
    SMX_LAUNCHED();
    stage_end(ST_FINALIZE, st, 1);
    return 0;
  }
  const int64_t words = slab_stride(P, B) * n_int;
  long long* ws = static_cast<long long*>(workspace);
  hipLaunchKernelGGL(fill_kernel, dim3(4096), dim3(256), 0, st, reinterpret_cast<double*>(ws), words,
                     0.0);
  SMX_LAUNCHED();
  stage_begin(ST_INTEGRATE, st);
  hipLaunchKernelGGL(integrate_library_kernel<false>, dim3((unsigned)integrator_waves<false>(n_int)),
                     dim3(64), 0, st, *p, *g, *s, wb, T_total, dt0, n_int, bin_mapping, v0, sdisp,
                     ws, nullptr, valid, T_int, t_final, ckpt, ckpt_every);
  SMX_LAUNCHED();
  stage_end(ST_INTEGRATE, st, 1);
  const int64_t items = (int64_t)P + B;   // finalize work items per column
  stage_begin(ST_FINALIZE, st);
  hipLaunchKernelGGL(finalize_library_kernel<true>, dim3(blocks_for(n * items, 256)), dim3(256), 0, st,
                     ws, n, r, P, B, valid, T_int, num_per_bin, g->area_pixel, dt0, *s, NC, sdisp,
                     lib);
  SMX_LAUNCHED();
  stage_end(ST_FINALIZE, st, 1);
  return 0;
}

int smx_build_library(const smx_potential* p, const smx_grids* g, const smx_schedule* s,
                      const double* wb, const double* T_total, const double* dt0, int64_t n,
                      int32_t r, const int32_t* bin_mapping, const double* num_per_bin,
                      const double* v0, const double* sdisp, void* workspace,
                      int64_t workspace_bytes, double* lib, double* valid, double* T_int,
                      double* t_final, void* stream) {
  return build_library(p, g, s, wb, T_total, dt0, n, r, bin_mapping, num_per_bin, v0, sdisp,
                       workspace, workspace_bytes, lib, valid, T_int, t_final, nullptr, 0,
                       knob(KN_LIB_BINNING) == 1, stream);
}

// ---------------------------------------------------------------------------
// smx_build_library without device allocation: the per-integration
// accumulator slabs (the workspace of smx_build_library) live in a fixed
// file-scope device array, `scratch`, of SMX_NOALLOC_SCRATCH_BYTES.  Each GPU
// the library runs on holds its own copy for the life of the process.  Its
// device address is looked up once per device (hipGetSymbolAddress, with the
// device current) and kept; the array's name is only a host-side stand-in.
// Every call on one GPU shares that GPU's array: the slabs are cleared by a
// fill kernel at the start of each call, and calls that overlap on one GPU
// (other streams or threads) race on it.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[SMX_NOALLOC_SCRATCH_BYTES / sizeof(double)];

static std::mutex g_scratch_mu;
static double* g_scratch_dev[64] = {};

static int scratch_for_current_device(double** out) {
  int dev = -1;
  SMX_HIP(hipGetDevice(&dev));
  SMX_REQUIRE(dev >= 0 && dev < 64, "smx_build_library_noalloc: device id %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* ptr = nullptr;
    SMX_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    SMX_REQUIRE(ptr != nullptr, "smx_build_library_noalloc: no device address for the scratch array");
    g_scratch_dev[dev] = static_cast<double*>(ptr);
  }
  *out = g_scratch_dev[dev];
  return 0;
}

int smx_build_library_noalloc(const smx_potential* p, const smx_grids* g, const smx_schedule* s,
                              const double* wb, const double* T_total, const double* dt0,
                              int64_t n, int32_t r, const int32_t* bin_mapping,
                              const double* num_per_bin, const double* v0, const double* sdisp,
                              double* lib, double* valid, double* T_int, double* t_final,
                              void* stream) {
  SMX_REQUIRE(g, "smx_build_library_noalloc: null pointer");
  SMX_REQUIRE(n >= 0 && r >= 1, "smx_build_library_noalloc: bad sizes");
  if (int e = check_grids(g)) return e;
  const int64_t need = smx_library_workspace_bytes(g, n * r);
  SMX_REQUIRE(need <= SMX_NOALLOC_SCRATCH_BYTES,
              "smx_build_library_noalloc: %lld integrations need %lld bytes of accumulator slabs, "
              "more than the %lld-byte scratch array; use smx_build_library with a workspace",
              (long long)(n * r), (long long)need, (long long)SMX_NOALLOC_SCRATCH_BYTES);
  double* ws = nullptr;
  if (int e = scratch_for_current_device(&ws)) return e;
  return build_library(p, g, s, wb, T_total, dt0, n, r, bin_mapping, num_per_bin, v0, sdisp, ws,
                       SMX_NOALLOC_SCRATCH_BYTES, lib, valid, T_int, t_final, nullptr, 0, false, stream);
}

int smx_build_library_ckpt(const smx_potential* p, const smx_grids* g, const smx_schedule* s,
                           const double* wb, const double* T_total, const double* dt0, int64_t n,
                           int32_t r, const int32_t* bin_mapping, const double* num_per_bin,
                           const double* v0, const double* sdisp, void* workspace,
                           int64_t workspace_bytes, double* lib, double* valid, double* T_int,
                           double* t_final, double* ckpt, int32_t ckpt_every, void* stream) {
  SMX_REQUIRE(ckpt != nullptr, "smx_build_library_ckpt: null checkpoint buffer");
  return build_library(p, g, s, wb, T_total, dt0, n, r, bin_mapping, num_per_bin, v0, sdisp,
                       workspace, workspace_bytes, lib, valid, T_int, t_final, ckpt, ckpt_every,
                       knob(KN_LIB_BINNING) == 1, stream);
}

int smx_build_library_traj(const smx_potential* p, const smx_grids* g, const smx_schedule* s,
                           const double* wb, const double* T_total, const double* dt0, int64_t n,
                           int32_t r, const int32_t* bin_mapping, const double* num_per_bin,
                           const double* v0, const double* sdisp, void* workspace, double* lib,
                           double* o_traj, double* t_traj, double* valid, double* T_int,
                           void* stream) {
  SMX_REQUIRE(p && g && s && wb && T_total && dt0 && bin_mapping && num_per_bin && v0 && sdisp &&
                  workspace && lib && o_traj && t_traj && valid && T_int,
              "smx_build_library_traj: null pointer");
  SMX_REQUIRE(n >= 0 && r >= 1 && s->N_max >= 0, "smx_build_library_traj: bad sizes");
  SMX_REQUIRE(g->symmetry == 0, "smx_build_library_traj: the trajectory path has no symmetry");
  if (int e = check_grids(g)) return e;
  if (n == 0) return 0;
  const int64_t n_int = n * r;
  const int P = g->nR * g->nz * g->nphi;
  const int64_t words = slab_stride(P, g->num_vbin) * n_int;
  hipStream_t st = as_stream(stream);
  double* ws = static_cast<double*>(workspace);
  hipLaunchKernelGGL(fill_kernel, dim3(4096), dim3(256), 0, st, ws, words, 0.0);
  SMX_LAUNCHED();
  hipLaunchKernelGGL(integrate_traj_kernel, dim3(blocks_for(n_int, 64)), dim3(64), 0, st, *p, *s,
                     wb, T_total, dt0, n_int, o_traj, t_traj, valid, T_int);
  SMX_LAUNCHED();
  hipLaunchKernelGGL(bin_traj_kernel, dim3(blocks_for(n_int, 64)), dim3(64), 0, st, *g, n_int,
                     s->N_max, o_traj, t_traj, T_int, bin_mapping, v0, sdisp, ws);
  SMX_LAUNCHED();
  const int64_t items = (int64_t)P + g->num_vbin;   // finalize work items per column
  hipLaunchKernelGGL(finalize_library_kernel<false>, dim3(blocks_for(n * items, 256)), dim3(256), 0,
                     st, ws, n, r, P, g->num_vbin, valid, T_int, num_per_bin, g->area_pixel, dt0, *s,
                     1, sdisp, lib);
  SMX_LAUNCHED();
  return 0;
}

int smx_bin_samples(const smx_grids* g, const double* samples, const double* dt, int64_t S,
                    const int32_t* bin_mapping, const double* v0, const double* sdisp,
                    int32_t* idx_rz, int32_t* idx_ap, double* acc_rz, double* acc_ap,
                    void* stream) {
  SMX_REQUIRE(g && samples && dt && bin_mapping && v0 && sdisp && idx_rz && idx_ap && acc_rz &&
                  acc_ap,
              "smx_bin_samples: null pointer");
  if (int e = check_grids(g)) return e;
  if (S <= 0) return 0;
  const int P = g->nR * g->nz * g->nphi;
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(fill_kernel, dim3(64), dim3(256), 0, st, acc_rz, (int64_t)P, 0.0);
  hipLaunchKernelGGL(fill_kernel, dim3(64), dim3(256), 0, st, acc_ap, 6 * (int64_t)g->num_vbin, 0.0);
  hipLaunchKernelGGL(bin_samples_kernel, dim3(blocks_for(S, 128)), dim3(128), 0, st, *g, samples,
                     dt, S, bin_mapping, v0, sdisp, idx_rz, idx_ap, acc_rz, acc_ap);
  SMX_LAUNCHED();
  return 0;
}

int smx_density_target(const smx_potential* p, const double* R_grid, const double* z_grid,
                       const double* phi_grid, int64_t P, double dR, double dz, double dphi,
                       const double* sample, int64_t S, double* out, void* stream) {
  SMX_REQUIRE(p && R_grid && z_grid && phi_grid && sample && out && P >= 0 && S > 0,
              "smx_density_target: bad arguments");
  if (P == 0) return 0;
  stage_begin(ST_DENSITY, as_stream(stream));
  hipLaunchKernelGGL(density_target_kernel, dim3((unsigned)P), dim3(256), 0, as_stream(stream), *p,
                     R_grid, z_grid, phi_grid, P, dR, dz, dphi, sample, S, out);
  SMX_LAUNCHED();
  stage_end(ST_DENSITY, as_stream(stream), 1);
  return 0;
}

}  // extern "C"
