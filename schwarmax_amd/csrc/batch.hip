// Many model evaluations in one launch: the Jeans ICs of every (model, star)
// pair in one grid, and the orbit libraries of every model in one persistent
// launch fed by a device work queue.
//
// The batch is the reference's vmap over parameter vectors (the MCMC driver
// evaluates model_likelihood per theta, likelihood.py:96-244); each model has
// its own potential, sky rotation, schedule and orbit count.  Work items are
// (model, up to 64/r columns) pieces, sorted by cost on the host; each wave
// of the persistent grid pops items from an atomic counter, integrates the
// item's r x columns integrations into a slab set it owns (so the slab pool
// is bounded by the grid size, not the batch), then writes the item's
// realization-averaged library rows and re-zeroes the slabs it touched.  The
// slabs hold the same fixed-point sums as the single-model path's, so each
// model's library equals its single-model library bit for bit.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <vector>

#include "common.hpp"
#include "potential.hpp"
#include "binning.hpp"
#include "integrator.hpp"
#include "ics.hpp"
#include "timing.hpp"
#include "../../include/smx.h"

namespace smx {

struct BatchModel {
  const double* wb;        // (n, r, 6)
  const double* T_total;   // (n, r)
  const double* dt0;       // (n, r)
  double* lib;             // (n, P + 6B)
  double* wb_out;          // IC outputs (n, r, 6), T_total, dt0
  double* T_out;
  double* dt0_out;
  int64_t n;
  int64_t star_off;        // first star of this model in the IC grid
};

struct BatchItem {
  int32_t model, col0, ncols, pad;
};

__global__ __launch_bounds__(256) void initial_conditions_batch_kernel(
    const smx_potential* __restrict__ pots, const BatchModel* __restrict__ models, int n_models,
    const double* __restrict__ N_step, const double* __restrict__ N_dyn, int64_t n_stars,
    const double* __restrict__ w0, const double* __restrict__ normals,
    const double* __restrict__ uniforms, int r, int r_total) {
  const int lane = threadIdx.x & 63;
  const int64_t gs = blockIdx.x * (int64_t)(blockDim.x >> 6) + (threadIdx.x >> 6);
  if (gs >= n_stars) return;   // whole wave exits together
  int m = 0;
  while (m + 1 < n_models && models[m + 1].star_off <= gs) ++m;
  const BatchModel& M = models[m];
  const int64_t i = gs - M.star_off;
  const smx_potential p = pots[m];
  StarIC<double> o;
  star_ic<double>(p, fields_const<double>(p), w0, normals, i, lane, o);
  if (lane != 0) return;
  const double scale[6] = {0.1, 0.1, 0.1, o.vsc, o.vsc, o.vsc};
  for (int k = 0; k < r; ++k) {
    const int64_t q = i * r + k;
    for (int j = 0; j < 6; ++j)
      M.wb_out[6 * q + j] = o.wn[j] + (uniforms[(i * r_total + k) * 6 + j] - 0.5) * scale[j];
    M.T_out[q] = o.T_orb * N_dyn[m];
    M.dt0_out[q] = o.T_orb / N_step[m];
  }
}

__global__ __launch_bounds__(64) void integrate_batch_kernel(
    const smx_potential* __restrict__ pots, const smx_grids* __restrict__ grids,
    const smx_schedule* __restrict__ scheds, const BatchModel* __restrict__ models,
    const BatchItem* __restrict__ items, int n_items, int r, const int32_t* __restrict__ bm,
    const double* __restrict__ npb, const double* __restrict__ v0, const double* __restrict__ sd,
    long long* __restrict__ pool, unsigned* __restrict__ counter) {
  __shared__ double s_valid[64];
  __shared__ double s_T[64];
  __shared__ int s_et[64];
  __shared__ long long ap_lds[24 * 64];   // aperture-record cache, 4 copies x 6 terms
  __shared__ int s_item;
  const int lane = threadIdx.x;
  const smx_grids& g0 = grids[0];
  const int P = g0.nR * g0.nz * g0.nphi;
  const int B = g0.num_vbin;
  const int64_t stride = slab_stride(P, B);
  const int64_t rows = (int64_t)P + 6 * (int64_t)B;
  long long* rz = pool + ((int64_t)blockIdx.x * 64 + lane) * stride;
  long long* ap = rz + slab_rz_len(P);
  for (;;) {
    if (lane == 0) s_item = (int)atomicAdd(counter, 1u);
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(s_item);
    __syncthreads();
    if (it >= n_items) break;
    const BatchItem w = items[it];
    const BatchModel& M = models[w.model];
    const smx_potential& p = pots[w.model];
    const smx_grids& g = grids[w.model];
    const smx_schedule& s = scheds[w.model];
    const int cc = lane / r, k = lane - cc * r;
    const bool active = cc < w.ncols && cc * r + k < 64;
    const int NC = g.symmetry ? 4 : 1;
    double valid = 0.0, T = 1.0;
    int e_t = 0;
    if (active) {
      const int64_t q = (int64_t)(w.col0 + cc) * r + k;
      double y[6], k1[6];
#pragma unroll
      for (int j = 0; j < 6; ++j) y[j] = M.wb[6 * q + j];
      deriv(p, s.Omega, y, k1);
      const double EJ0 = jacobi(p, s.Omega, y);
      double t = 0.0, dt = M.dt0[q];
      e_t = fix_exp_time(M.dt0[q], s, NC);
      const double T_acc =
          g.symmetry ? integrate_and_bin<4>(p, g, s, bm, v0, sd, t, y, k1, dt, e_t, rz, ap, nullptr,
                                            0, 1, ap_lds + lane)
                     : integrate_and_bin<1>(p, g, s, bm, v0, sd, t, y, k1, dt, e_t, rz, ap, nullptr,
                                            0, 1, ap_lds + lane);
      const double EJ = jacobi(p, s.Omega, y);
      valid = ((fabs(EJ / EJ0 - 1.0) < 0.1) && (t > M.T_total[q] / 10.0)) ? 1.0 : 0.0;
      T = T_acc + kEps;
    }
    s_valid[lane] = valid;
    s_T[lane] = T;
    s_et[lane] = e_t;
    __syncthreads();
    // realization average of the item's columns (integrates.py:212-236,
    // 287-300), rows [Rzphi | xy | h1 | h2 | h3 | h4 | h0]; the slabs read are
    // zeroed for the next item
    for (int c = 0; c < w.ncols; ++c) {
      double nv = 0.0;
      for (int kk = 0; kk < r; ++kk) nv += s_valid[c * r + kk];
      const double wt = 1.0 / (nv + 0.1);
      double* out = M.lib + (int64_t)(w.col0 + c) * rows;
      for (int64_t e = lane; e < rows; e += 64) {
        double acc = 0.0;
        for (int kk = 0; kk < r; ++kk) {
          long long* srz = pool + ((int64_t)blockIdx.x * 64 + c * r + kk) * stride;
          long long* sap = srz + slab_rz_len(P);
          const int et = s_et[c * r + kk];
          const bool ok = s_valid[c * r + kk] > 0.5;
          if (e < P) {
            lib_add(acc, static_cast<double>(srz[e]) * fix_unscale(et), s_T[c * r + kk], false, 1.0,
                    ok, wt);
            srz[e] = 0;
          } else {
            const int64_t q = e - P;
            const int64_t blk = q / B, a = q - blk * B;
            const int comp = block_comp((int)blk);
            const double un = fix_unscale(comp == 0 ? et : et + fix_exp_gh(sd[a]));
            lib_add(acc, static_cast<double>(sap[kApRec * a + comp]) * un, s_T[c * r + kk], blk == 0,
                    aperture_area(npb[a], g.area_pixel), ok, wt);
            sap[kApRec * a + comp] = 0;
          }
        }
        out[e] = acc;
      }
    }
    __syncthreads();
  }
}

__global__ void zero_kernel(long long* __restrict__ p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

constexpr int BATCH_WAVES = 1024;   // persistent waves (4 per CU on 256 CUs)

}  // namespace smx

using namespace smx;

extern "C" {

int64_t smx_batch_workspace_bytes(const smx_grids* g, int32_t n_models, int64_t total_items) {
  if (!g || n_models < 1 || total_items < 0) return -1;
  const int64_t pool = slab_stride(g->nR * g->nz * g->nphi, g->num_vbin) * 64 * BATCH_WAVES;
  return (int64_t)sizeof(double) * pool + 256 +
         (int64_t)n_models * (sizeof(smx_potential) + sizeof(smx_grids) + sizeof(smx_schedule) +
                              sizeof(BatchModel) + 2 * sizeof(double)) +
         total_items * (int64_t)sizeof(BatchItem) + 4096;
}

static size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

int smx_build_library_batch(int32_t n_models, const smx_potential* pots, const smx_grids* grids,
                            const smx_schedule* scheds, const double* w0, const double* normals,
                            const double* uniforms, const int64_t* n_orbits, int32_t r,
                            int32_t r_total, const double* N_step_per_orb,
                            const double* N_dynamical_time, const int32_t* bin_mapping,
                            const double* num_per_bin, const double* v0, const double* sdisp,
                            double* const* wb, double* const* T_total, double* const* dt0,
                            double* const* lib, void* workspace, int64_t workspace_bytes,
                            void* stream) {
  SMX_REQUIRE(n_models >= 1 && pots && grids && scheds && w0 && normals && uniforms && n_orbits &&
                  N_step_per_orb && N_dynamical_time && bin_mapping && num_per_bin && v0 && sdisp &&
                  wb && T_total && dt0 && lib && workspace,
              "smx_build_library_batch: null pointer");
  SMX_REQUIRE(r >= 1 && r <= 64 && r <= r_total, "smx_build_library_batch: bad r");
  const int P = grids[0].nR * grids[0].nz * grids[0].nphi;
  const int B = grids[0].num_vbin;
  for (int m = 0; m < n_models; ++m) {
    SMX_REQUIRE(grids[m].nR * grids[m].nz * grids[m].nphi == P && grids[m].num_vbin == B &&
                    grids[m].nX == grids[0].nX && grids[m].nY == grids[0].nY,
                "smx_build_library_batch: every model must share the binning grids");
    SMX_REQUIRE(n_orbits[m] >= 0 && wb[m] && T_total[m] && dt0[m] && lib[m],
                "smx_build_library_batch: bad model %d", m);
    SMX_REQUIRE(scheds[m].N_max >= 0, "smx_build_library_batch: bad schedule %d", m);
  }
  // host tables: models and work items sorted by cost (attempts), longest first
  const int cols_per_item = 64 / r;
  std::vector<BatchModel> hm(n_models);
  std::vector<BatchItem> hi;
  int64_t stars = 0;
  for (int m = 0; m < n_models; ++m) {
    hm[m] = BatchModel{wb[m], T_total[m], dt0[m], lib[m], wb[m], T_total[m], dt0[m], n_orbits[m],
                       stars};
    stars += n_orbits[m];
    for (int64_t c = 0; c < n_orbits[m]; c += cols_per_item) {
      const int nc = (int)((n_orbits[m] - c < cols_per_item) ? n_orbits[m] - c : cols_per_item);
      hi.push_back(BatchItem{m, (int32_t)c, nc, 0});
    }
  }
  std::stable_sort(hi.begin(), hi.end(), [&](const BatchItem& a, const BatchItem& b) {
    return scheds[a.model].N_max > scheds[b.model].N_max;
  });
  const int64_t need = smx_batch_workspace_bytes(&grids[0], n_models, (int64_t)hi.size());
  SMX_REQUIRE(workspace_bytes >= need, "smx_build_library_batch: workspace %lld < %lld bytes",
              (long long)workspace_bytes, (long long)need);
  if (stars == 0) return 0;
  hipStream_t st = as_stream(stream);
  char* base = static_cast<char*>(workspace);
  const int64_t pool_words = slab_stride(P, B) * 64 * BATCH_WAVES;
  long long* pool = reinterpret_cast<long long*>(base);
  size_t off = align256(pool_words * sizeof(double));
  unsigned* counter = reinterpret_cast<unsigned*>(base + off);
  off += 256;
  smx_potential* d_pots = reinterpret_cast<smx_potential*>(base + off);
  off = align256(off + n_models * sizeof(smx_potential));
  smx_grids* d_grids = reinterpret_cast<smx_grids*>(base + off);
  off = align256(off + n_models * sizeof(smx_grids));
  smx_schedule* d_scheds = reinterpret_cast<smx_schedule*>(base + off);
  off = align256(off + n_models * sizeof(smx_schedule));
  BatchModel* d_models = reinterpret_cast<BatchModel*>(base + off);
  off = align256(off + n_models * sizeof(BatchModel));
  double* d_nstep = reinterpret_cast<double*>(base + off);
  off = align256(off + n_models * sizeof(double));
  double* d_ndyn = reinterpret_cast<double*>(base + off);
  off = align256(off + n_models * sizeof(double));
  BatchItem* d_items = reinterpret_cast<BatchItem*>(base + off);
  SMX_HIP(hipMemcpyAsync(d_pots, pots, n_models * sizeof(smx_potential), hipMemcpyHostToDevice, st));
  SMX_HIP(hipMemcpyAsync(d_grids, grids, n_models * sizeof(smx_grids), hipMemcpyHostToDevice, st));
  SMX_HIP(hipMemcpyAsync(d_scheds, scheds, n_models * sizeof(smx_schedule), hipMemcpyHostToDevice, st));
  SMX_HIP(hipMemcpyAsync(d_models, hm.data(), n_models * sizeof(BatchModel), hipMemcpyHostToDevice, st));
  SMX_HIP(hipMemcpyAsync(d_nstep, N_step_per_orb, n_models * sizeof(double), hipMemcpyHostToDevice, st));
  SMX_HIP(hipMemcpyAsync(d_ndyn, N_dynamical_time, n_models * sizeof(double), hipMemcpyHostToDevice, st));
  SMX_HIP(hipMemcpyAsync(d_items, hi.data(), hi.size() * sizeof(BatchItem), hipMemcpyHostToDevice, st));
  // the host tables must outlive the copies
  SMX_HIP(hipStreamSynchronize(st));
  stage_begin(ST_ICS, st);
  hipLaunchKernelGGL(initial_conditions_batch_kernel, dim3(blocks_for(stars, 4)), dim3(256), 0, st,
                     d_pots, d_models, n_models, d_nstep, d_ndyn, stars, w0, normals, uniforms, r,
                     r_total);
  SMX_LAUNCHED();
  stage_end(ST_ICS, st, 1);
  stage_begin(ST_INTEGRATE, st);
  hipLaunchKernelGGL(zero_kernel, dim3(4096), dim3(256), 0, st, pool, pool_words);
  hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, st, reinterpret_cast<long long*>(counter), 1);
  SMX_LAUNCHED();
  hipLaunchKernelGGL(integrate_batch_kernel, dim3(BATCH_WAVES), dim3(64), 0, st, d_pots, d_grids,
                     d_scheds, d_models, d_items, (int)hi.size(), (int)r, bin_mapping, num_per_bin,
                     v0, sdisp, pool, counter);
  SMX_LAUNCHED();
  stage_end(ST_INTEGRATE, st, 1);
  return 0;
}

}  // extern "C"
