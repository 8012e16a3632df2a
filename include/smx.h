/*
 * smx.h — C ABI of libsmx.so, the MI355X (gfx950) implementation of
 * SchwarMAX's model-evaluation path.
 *
 * The reference has no FFI: its boundary is a set of JAX functions.  Each entry
 * point below names the reference function whose computation it replaces.
 * Conventions:
 *   - every array argument is a caller-owned DEVICE pointer (torch tensors'
 *     data_ptr() on the Python side), row-major, with the shape in its comment;
 *   - scratch memory is passed in by the caller (sizes from the *_bytes queries);
 *   - `stream` is a hipStream_t passed as void*; every call is asynchronous on it;
 *   - the return value is 0 on success, non-zero on error; smx_last_error()
 *     then returns a message.  Numerical failures are not errors: as in the
 *     reference they become NaN, zero weights or invalid orbits.
 */
#ifndef SMX_H
#define SMX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Canonical potential NFW + Miyamoto-Nagai + T3 bar + V4 bulge and the
 * canonical stellar density MN + T3 + V4 (likelihood.py:13-71).  Filled on the
 * host from the reference's parameter dicts. */
typedef struct smx_potential {
  /* NFW halo (potentials.py:18-27) */
  double nfw_mGM;        /* -G * 10**logM                                 */
  double nfw_Rs, nfw_a, nfw_b, nfw_c;
  double nfw_origin[3];
  double nfw_rot[9];     /* get_mat(dirx, diry, dirz), row-major          */
  /* Miyamoto-Nagai disc (potentials.py:29-65) */
  double mn_mGM;         /* -G * 10**logM_disc                            */
  double mn_rho0;        /* Hs_disc**2 * 10**logM_disc / (4 pi)           */
  double mn_Rs, mn_Hs;
  double mn_origin[3];
  double mn_rot[9];
  /* Dehnen & Aly T3 bar and V4 bulge (dehnen_bar.py:1318-1383, 1425-1543) */
  double t3_M, t3_a, t3_b, t3_L, t3_gamma;
  double v4_M, v4_a, v4_b, v4_L, v4_gamma;
  double G;
} smx_potential;

/* Binning geometry (integrates.py:37-52, utils.py:10-42). */
typedef struct smx_grids {
  int32_t nR, nz, nphi;    /* (R, z, phi) cell grid, strides 1, nR, nR*nz   */
  int32_t nX, nY;          /* sky pixel grid                                 */
  int32_t num_vbin;        /* apertures B                                    */
  int32_t symmetry;        /* 1: fourfold bar symmetry (integrates.py:74-79) */
  int32_t pad_;
  double rz_min[3], rz_max[3];
  double xy_min[2], xy_max[2];
  double rot[9];           /* makeRotationMatrix(alpha, beta, gamma)         */
  double area_pixel;       /* pixel area in pc^2 (integrates.py:70-71)       */
} smx_grids;

/* Integration schedule (model.py:163-180). */
typedef struct smx_schedule {
  int32_t N_max;           /* BS3 attempts per integration (fixed)           */
  int32_t pad_;
  double Omega;            /* ODE pattern speed, = -Omega_bar (model.py:173) */
  double atol, rtol, dt_min, dt_max;
} smx_schedule;

/* ---- library management ---------------------------------------------- */
int smx_last_error(char* buf, int64_t len);
int smx_version(void);
/* Per-stage timing with HIP events on the launch stream (no counterpart in
 * the reference; used by the benchmark).  Stages, in order: initial
 * conditions, integration, finalize (the sample binning and realization
 * average, or the fused form's finalize), Gram, factor + inverse, right-hand
 * sides, ADMM loop, model maps, density target, and the ADMM product kernel alone
 * (one record per launch, nested inside the ADMM loop).  enable != 0 turns recording on and
 * clears previous records; read fills ms[] and launches[] (kernel launches per
 * stage) summed over the records since, then clears them. */
int smx_timing_enable(int enable);
int smx_timing_read(double* ms, int64_t* launches, int32_t max_stages);

/* ---- potential (potentials.py, dehnen_bar.py, likelihood.py:13-71) ------ */
/* Phi and acc = -grad Phi at n points xyz (n,3); acc may be NULL.
 * Replaces potential_func / acc_fn (likelihood.py:46-71, model.py:152-161). */
int smx_potential_eval(const smx_potential* p, const double* xyz, int64_t n,
                       double* phi, double* acc, void* stream);
/* Stellar density at n points.  Replaces density_func (likelihood.py:13-43). */
int smx_density_eval(const smx_potential* p, const double* xyz, int64_t n,
                     double* rho, void* stream);

/* ---- initial conditions (jeans.py:9-94, model.py:112-169) --------------- */
/* Jeans velocities from positions w0 (n,3) and normals (n,3), rotation curve,
 * jitter of r realizations from uniforms (n,r_total,6) (first r used), orbital
 * time scale.  Outputs wb (n,r,6), T_total (n,r) = N_dyn*T_orb,
 * dt0 (n,r) = T_orb/N_step.  Replaces get_w0_new + the jitter/timescale block
 * of build_model (model.py:112-169). */
int smx_initial_conditions(const smx_potential* p, const double* w0, const double* normals,
                           const double* uniforms, int64_t n, int32_t r, int32_t r_total,
                           double N_step_per_orb, double N_dyn,
                           double* wb, double* T_total, double* dt0, void* stream);

/* ---- orbit library (integrates.py:37-313) ------------------------------- */
/* Every bin_mapping argument below holds nX*nY + 1 int32 entries: the pixel
 * -> aperture map, then the entry read for samples off the sky field (the
 * reference's sentinel -1, or for a map given without one the last pixel's
 * aperture, which is where JAX's gather of index nX*nY clamps;
 * integrates.py:132-143, utils.py:429).  Ids outside [0, num_vbin) are
 * dropped, as segment_sum drops them. */
/* Bytes of the per-integration accumulator slabs of n_int integrations: the
 * workspace of smx_build_library_traj and smx_build_library_noalloc's limit. */
int64_t smx_library_workspace_bytes(const smx_grids* g, int64_t n_int);
/* Bytes of the workspace of smx_build_library_ckpt: the slabs, which the
 * adjoint reads, followed by the sample rows. */
int64_t smx_library_ckpt_bytes(const smx_grids* g, const smx_schedule* s, int64_t n_int);
/* Bytes of the workspace of smx_build_library for n_int integrations: the
 * larger of the slabs and the sample rows (N_max 64-byte rows each). */
int64_t smx_library_build_bytes(const smx_grids* g, const smx_schedule* s, int64_t n_int);
/* Integrate n_int = n*r orbits, then bin and average realizations into lib
 * (n, P + 6B) laid out as [Rzphi (P) | surface density (B) | h1 | h2 | h3 |
 * h4 | h0] per column.  The integrator stores every attempt's state in the
 * workspace and one workgroup per column bins them into LDS histograms
 * (SMX_KNOB_LIB_BINNING = 2 selects the fused integrator, which bins into
 * per-integration slabs; both give the same bits: the sums are 64-bit fixed
 * point).  valid/T_int/t_final (n_int,) are per-integration diagnostics.
 * workspace: workspace_bytes bytes; the sample binning runs when they are at
 * least smx_library_build_bytes(g, s, n*r), the fused form when they only hold
 * the slabs (smx_library_workspace_bytes); fewer are refused (code 2).
 * Replaces _integrate_adaptive_batch_chunked_vmap (integrates.py:303-313). */
int smx_build_library(const smx_potential* p, const smx_grids* g, const smx_schedule* s,
                      const double* wb, const double* T_total, const double* dt0,
                      int64_t n, int32_t r, const int32_t* bin_mapping,
                      const double* num_per_bin, const double* v0, const double* sdisp,
                      void* workspace, int64_t workspace_bytes, double* lib, double* valid,
                      double* T_int, double* t_final, void* stream);
/* No counterpart in the reference: smx_build_library for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments and results as
 * smx_build_library without the workspace: the accumulator slabs live in a
 * fixed device array of SMX_NOALLOC_SCRATCH_BYTES that each GPU holds for the
 * life of the process, so n * r * (P rounded up to 8 + 6 B) * 8 bytes
 * (smx_library_workspace_bytes) must fit in it; larger calls are rejected
 * (code 2, message in smx_last_error).  No hipMalloc / hipFree.  Calls that
 * overlap on one GPU share that array and race. */
#define SMX_NOALLOC_SCRATCH_BYTES (64LL << 20)
int smx_build_library_noalloc(const smx_potential* p, const smx_grids* g, const smx_schedule* s,
                              const double* wb, const double* T_total, const double* dt0,
                              int64_t n, int32_t r, const int32_t* bin_mapping,
                              const double* num_per_bin, const double* v0, const double* sdisp,
                              double* lib, double* valid, double* T_int, double* t_final,
                              void* stream);
/* Trajectory variant, no symmetry, dt normalised (integrates.py:319-581):
 * o_traj (n_int, N_max, 6) with velocities in km/s, t_traj (n_int, N_max)
 * the cumulative time, and the realization average lib (n, P + 6B). */
int smx_build_library_traj(const smx_potential* p, const smx_grids* g, const smx_schedule* s,
                           const double* wb, const double* T_total, const double* dt0,
                           int64_t n, int32_t r, const int32_t* bin_mapping,
                           const double* num_per_bin, const double* v0, const double* sdisp,
                           void* workspace, double* lib, double* o_traj, double* t_traj,
                           double* valid, double* T_int, void* stream);
/* Bin externally supplied samples (S,6) with weights dt (S,) into ONE
 * histogram: the bit-exactness hook for the grid rules (utils.py:10-42,
 * integrates.py:115-174).  Writes per-sample cell / aperture ids for each
 * symmetry copy (S, nsym) (-1 when dropped), and the sums acc_rz (P) and
 * acc_ap (B, 6) = [dt, g dt H0 .. g dt H4]. */
int smx_bin_samples(const smx_grids* g, const double* samples, const double* dt, int64_t S,
                    const int32_t* bin_mapping, const double* v0, const double* sdisp,
                    int32_t* idx_rz, int32_t* idx_ap, double* acc_rz, double* acc_ap,
                    void* stream);
/* Host-only check of the library integrator's work schedule (no counterpart
 * in the reference, which integrates one orbit per vmap lane): replays, with
 * the kernel's own schedule arithmetic, how `waves` waves of 64 lanes share
 * n_int integrations of L attempts (each wave's attempts spread over its
 * lanes, an integration cut at a lane boundary finished by the next lane).
 * Returns 0 when every integration makes attempts 0..L-1 exactly once, in
 * order, and is finished once; *steps_out = the per-lane loop length of
 * wave 0.  Touches no GPU. */
int smx_integrator_schedule_check(int64_t n_int, int64_t waves, int64_t L, int64_t* steps_out);
/* The same for the balanced library adjoint (no counterpart in the
 * reference, which differentiates with jax.grad): replays its step loop over
 * U checkpoint segments per integration, swept from the last, with the
 * lanes that have no unit shadowing lane 0.  Returns 0 when every
 * integration's units are swept once and in order, every tail starts from
 * its head's hand-over, lane 0 holds the shadowed unit at every step and
 * every integration is finished once.  Touches no GPU. */
int smx_adjoint_schedule_check(int64_t n_int, int64_t waves, int64_t U, int64_t* steps_out);

/* ---- density target (model.py:192-213) ---------------------------------- */
/* Cell masses by Sobol quadrature: out (P) from cell centres (P) x3 and
 * samples (S,3). */
int smx_density_target(const smx_potential* p, const double* R_grid, const double* z_grid,
                       const double* phi_grid, int64_t P, double dR, double dz, double dphi,
                       const double* sample, int64_t S, double* out, void* stream);

/* ---- derivatives (jax.grad of likelihood.py:96-244 through model.py:70-258) */
/* The potential's theta-dependent fields, in this order, are the coordinates
 * of every field cotangent / Jacobian below: nfw_mGM, nfw_Rs, mn_mGM,
 * mn_rho0, mn_Rs, mn_Hs, t3_M, t3_L, t3_a, t3_b, v4_M.  The host maps theta
 * to the fields (likelihood.py:105-165) and differentiates that map itself. */
#define SMX_POT_FIELDS 11

/* smx_build_library that also stores, for the adjoint, the state
 * ckpt[(j, i, 0:8)] = (y[6], dt, t) of integration i before attempt
 * j * ckpt_every, j < ceil(N_max / ckpt_every), and keeps each valid
 * integration's fixed-point sums in the slabs at the start of `workspace`
 * (smx_library_ckpt_bytes bytes for the sample binning, the rows following the
 * slabs; with only the slabs' bytes the fused form runs).  The slabs stay
 * valid until the caller reuses the workspace; smx_library_vjp reads them. */
int smx_build_library_ckpt(const smx_potential* p, const smx_grids* g, const smx_schedule* s,
                           const double* wb, const double* T_total, const double* dt0,
                           int64_t n, int32_t r, const int32_t* bin_mapping,
                           const double* num_per_bin, const double* v0, const double* sdisp,
                           void* workspace, int64_t workspace_bytes, double* lib, double* valid,
                           double* T_int, double* t_final, double* ckpt, int32_t ckpt_every,
                           void* stream);
/* Scratch for smx_library_vjp: per integration ckpt_every * 31 doubles of
 * replayed attempts plus two scalars. */
int64_t smx_library_vjp_workspace_bytes(int64_t n_int, int32_t ckpt_every);
/* Vector-Jacobian product of the library (integrates.py:37-300 with the
 * realization average; bin membership, accept decisions and validity fixed,
 * as jax.grad treats them).  lib_bar (n, P + 6B) is the cotangent of lib.
 * Outputs per integration i < n*r: wb_bar (i, 6), dt0_bar (i), field_bar
 * (i, SMX_POT_FIELDS), omega_bar (i) (the ODE's Omega = -Omega_bar) and
 * rot_bar (i, 3) (the sky rotation's row 2, which sets v_los); the caller
 * sums over integrations.  fwd_workspace, valid, T_int and ckpt are those of
 * the smx_build_library_ckpt call that made lib. */
int smx_library_vjp(const smx_potential* p, const smx_grids* g, const smx_schedule* s,
                    const double* wb, const double* dt0, int64_t n, int32_t r,
                    const int32_t* bin_mapping, const double* num_per_bin, const double* v0,
                    const double* sdisp, const void* fwd_workspace, const double* valid,
                    const double* T_int, const double* ckpt, int32_t ckpt_every,
                    const double* lib_bar, void* workspace, double* wb_bar, double* dt0_bar,
                    double* field_bar, double* omega_bar, double* rot_bar, void* stream);
/* Jacobians of smx_initial_conditions with respect to the fields
 * (jeans.py:9-94, model.py:117-169): wb_jac (n, r, 6, SMX_POT_FIELDS),
 * dt0_jac (n, r, SMX_POT_FIELDS).  T_total only sets validity and has none. */
int smx_initial_conditions_jac(const smx_potential* p, const double* w0, const double* normals,
                               const double* uniforms, int64_t n, int32_t r, int32_t r_total,
                               double N_step, double* wb_jac, double* dt0_jac, void* stream);
/* Jacobian of smx_density_target with respect to the fields (model.py:192-213):
 * jac (P, SMX_POT_FIELDS). */
int smx_density_target_jac(const smx_potential* p, const double* R_grid, const double* z_grid,
                           const double* phi_grid, int64_t P, double dR, double dz, double dphi,
                           const double* sample, int64_t S, double* jac, void* stream);

/* ---- NNLS operator (the jackknife's shared solve) ----------------------- */
/* The ADMM operator alone, for many solves that share one Gram up to
 * low-rank changes (model.py:490-577 re-solves the NNLS per jackknife
 * replicate): M = (U^T U + (lambda/n) I + rho I)^-1, rho = tr(.)/n
 * (nnls.py:115-121), formed once in `workspace` (sized by
 * smx_nnls_operator_bytes for at most K_max right-hand sides; U as in
 * smx_nnls_admm).  rho_out: one device double. */
int64_t smx_nnls_operator_bytes(int64_t n, int64_t m, int64_t K_max);
int smx_nnls_operator(const double* lib, int64_t ld_lib, int64_t n, int64_t m,
                      const double* row_scale, double lambda_reg, int64_t K_max, void* workspace,
                      double* rho_out, void* stream);
/* out (K, n) = R (K, n) M for K <= K_max rows, FP32 in and out (the ADMM
 * product: fp16 hi/lo splits, three MFMAs, FP32 accumulation). */
int smx_nnls_operator_apply(int64_t n, int64_t m, int64_t K_max, void* workspace, const float* R,
                            int64_t K, float* out, void* stream);
/* out (K, n) = D_hi M_hi: the fp16 hi parts only, one MFMA (the ADMM's delta
 * product).  For D = R - R_b, a change of an input R_b whose full product
 * the caller holds, out + (R_b M) equals R M to twice the full product's
 * error bound while max|D_k| <= 2^-10 max|R_k| on every row. */
int smx_nnls_operator_apply_hi(int64_t n, int64_t m, int64_t K_max, void* workspace,
                               const float* D, int64_t K, float* out, void* stream);
/* One ADMM iteration of all G jackknife replicates after the operator
 * products (model.py:490-577 re-solves nnls.py:51-66 per replicate): with
 * S (T, G, n) FP32 the products r M^(k+1) of this iteration's right-hand
 * sides r, x = sum_k delta_g^k S_k in FP64 (the shifted operator by its
 * Neumann series), V (G, d, n) the design rows each replicate drops,
 * P (G, d, n) the shifted operator applied to them and Cinv (G, d, d) the
 * inverse Woodbury capacitance matrices, w = x + P^T Cinv (V x); then, in
 * place, w_hat = alpha w + (1 - alpha) z, z = max(w_hat + u, 0),
 * u = u + w_hat - z (nnls.py:155-162), and r (G, n, FP32) = rho_g (z - u) - c
 * (G, n), the next right-hand sides.  d <= 2048, T <= 64. */
int smx_jackknife_admm_step(int64_t G, int64_t n, int64_t d, int64_t T, const float* S,
                            const double* delta, const double* V, const double* P,
                            const double* Cinv, const double* rho_g, const double* c, double alpha,
                            double* z, double* u, float* r, void* stream);

/* ---- batched evaluation (many theta, one launch per stage) -------------- */
/* Workspace for smx_build_library_batch: the persistent grid's slab pool plus
 * the model and work-item tables (total_items = sum over models of
 * ceil(n_orbits / (64 / r))). */
int64_t smx_batch_workspace_bytes(const smx_grids* g, int32_t n_models, int64_t total_items);
/* ICs and orbit libraries of n_models models in two launches (the reference
 * evaluates model.py:104-258 per theta).  Model m has its own potential,
 * grids (the sky rotation differs; the binning geometry must be shared),
 * schedule, N_step_per_orb / N_dynamical_time (model.py:163-169) and orbit
 * count n_orbits[m] <= the stars in w0 (n, 3), normals (n, 3) and uniforms
 * (n, r_total, 6), of which it uses the first n_orbits[m].  The host arrays
 * wb, T_total, dt0, lib hold per-model DEVICE pointers ((n_m, r, 6), (n_m, r),
 * (n_m, r), (n_m, P + 6B)).  Integration is one persistent launch over work
 * items of 64/r columns taken from a device queue, longest schedules first. */
int smx_build_library_batch(int32_t n_models, const smx_potential* pots, const smx_grids* grids,
                            const smx_schedule* scheds, const double* w0, const double* normals,
                            const double* uniforms, const int64_t* n_orbits, int32_t r,
                            int32_t r_total, const double* N_step_per_orb,
                            const double* N_dynamical_time, const int32_t* bin_mapping,
                            const double* num_per_bin, const double* v0, const double* sdisp,
                            double* const* wb, double* const* T_total, double* const* dt0,
                            double* const* lib, void* workspace, int64_t workspace_bytes,
                            void* stream);

/* ---- NNLS (nnls.py:6-166) ----------------------------------------------- */
/* Bytes of NNLS workspace for n orbits, m rows, K right-hand sides. */
int64_t smx_nnls_workspace_bytes(int64_t n, int64_t m, int64_t K);
/* Bootstrap ADMM NNLS.  lib (n, ld_lib) holds the library columns whose first
 * m entries are the design rows; row_scale (m) turns them into U
 * (U[j,i] = lib[i,j] * row_scale[j], FP32).  Y (K, m) are the stacked data
 * vectors.  Solves K problems with the shared factor of U^T U + lambda/n I,
 * maxiter ADMM steps (alpha = 1.6, rho = tr(Q)/n, start w = z = w_init).
 * Writes W (K, n).  tol > 0 stops each right-hand side early once an
 * iterate changes by less than tol (relative, max norm).
 * Replaces solve_nnls_admm_bootstrap (nnls.py:68-166) and, with K = 1,
 * solve_nnls_admm (nnls.py:6-66). */
int smx_nnls_admm(const double* lib, int64_t ld_lib, int64_t n, int64_t m,
                  const double* row_scale, const double* Y, int64_t K,
                  double lambda_reg, int32_t maxiter, double w_init, double tol,
                  void* workspace, double* W, int32_t* iters_done, void* stream);
/* How many of the last smx_nnls_admm's iterations on this workspace took the
 * full ADMM product; the others took the delta product on the hi plane of
 * the inverse (no reference counterpart: reported for the byte accounting of
 * the product kernel).  Synchronises the stream. */
int smx_nnls_admm_full_products(int64_t n, int64_t m, int64_t K, const void* workspace,
                                int32_t* full_products, void* stream);

/* Grouped bootstrap NNLS: the solves of several models with the same m and K
 * (the per-theta solve_nnls_admm_bootstrap calls, nnls.py:68-166, that the
 * reference's jax.vmap over theta puts in one program; example_schwarmax.ipynb
 * vmaps model_likelihood, likelihood.py:96-244).  Each model's Gram, inverse
 * and right-hand sides are formed one after another on a shared scratch;
 * then every ADMM iteration is one product launch and one update launch for
 * the whole group.  Each model's result is bit-identical to its
 * smx_nnls_admm call (tol = 0).  A problem holds what smx_nnls_admm takes for
 * its model; `state` is a device buffer of smx_nnls_group_state_bytes(n, m,
 * K) bytes that the solve keeps for that model. */
#define SMX_NNLS_GROUP_MAX 64
typedef struct smx_nnls_problem {
  const double* lib;        /* (n, ld_lib) library rows, first m entries the design rows */
  int64_t ld_lib;
  int64_t n;
  const double* row_scale;  /* (m) */
  const double* Y;          /* (K, m) */
  double w_init;            /* start w = z (nnls.py:123) */
  void* state;              /* smx_nnls_group_state_bytes(n, m, K) device bytes */
  double* W;                /* (K, n) out */
} smx_nnls_problem;
int64_t smx_nnls_group_state_bytes(int64_t n, int64_t m, int64_t K);
/* Scratch shared by a group whose largest model has n_max orbits. */
int64_t smx_nnls_group_scratch_bytes(int64_t n_max, int64_t m, int64_t K);
/* n_models in 1 .. SMX_NNLS_GROUP_MAX; problems is a host array.  Runs
 * maxiter iterations for every model (no early exit). */
int smx_nnls_admm_group(int32_t n_models, const smx_nnls_problem* problems, int64_t m, int64_t K,
                        double lambda_reg, int32_t maxiter, void* scratch, void* stream);
/* Full products taken by the last grouped solve of the model whose state
 * buffer this is.  Synchronises the stream. */
int smx_nnls_group_full_products(int64_t n, int64_t m, int64_t K, const void* state,
                                 int32_t* full_products, void* stream);

/* ---- model maps epilogue and bootstrap chi^2 (forward path) ------------- */
/* From maps (K, ldm) = W lib[:, :m] with rows [Rzphi (P) | xy (B) | h1..h4
 * (B each)]: surface_density (K, B) = (map_xy * mmpo) * lm, h (4, K, B) =
 * clip(map_h / gamma, -10, 10) with |gamma| floored at 1e-8, and V, sigma
 * (K, B) from h (model.py:14-66, 260-286; ghmoments.py:53-79). */
int smx_model_epilogue(const double* maps, int64_t ldm, int64_t K, int32_t P, int32_t B,
                       const double* gamma_kin, const double* v0, const double* s, double mmpo,
                       double lm, double* surface_density, double* h, double* V, double* sigma,
                       void* stream);
/* Bootstrap data of the chi^2: XY_density_data, its error and the (K, B)
 * standard normals; h1..h4 data, errors and normals (likelihood.py:184-244). */
typedef struct smx_bootstrap_data {
  const double* xy;
  const double* xy_err;
  const double* xy_normal;
  const double* h[4];
  const double* h_err[4];
  const double* h_normal[4];
} smx_bootstrap_data;
/* The model side of the chi^2: density_3d_model d3 (K rows, leading
 * dimension ld3 >= P), the density target y_R (P), surface_density and
 * h1..h4 (K, B each, contiguous). */
typedef struct smx_bootstrap_model {
  const double* d3;
  int64_t ld3;
  const double* y_R;
  const double* surface_density;
  const double* h[4];
} smx_bootstrap_model;
/* Bootstrap log-likelihood rows logl_rows (K) and their log-mean-exp logl (1)
 * (likelihood.py:184-244): every row adds the density term summed over all
 * rows and cells (the reference broadcasts res_Rzphi), NaN terms skipped.
 * work: 2K doubles.  The forward path's chi^2; the gradient takes its chi^2
 * through torch autograd. */
int smx_bootstrap_loglike(const smx_bootstrap_model* model, int64_t K, int32_t P, int32_t B,
                          const smx_bootstrap_data* data, double sigma_amplify, double* work,
                          double* logl_rows, double* logl, void* stream);

/* ---- ADMM product tuning (no reference counterpart: the knob the
 * autotuner in schwarmax_amd/tune.py sets) ------------------------------- */
/* 256-thread workgroups per CU that the ADMM product's split-K rule targets
 * (1 .. 8; 0 restores the default, 2; the rule keeps at least one workgroup
 * per CU).  Process-wide, atomic; applies to solves
 * started afterwards (workspaces are sized for the largest setting). */
int smx_admm_set_occupancy(int32_t workgroups_per_cu);
int smx_admm_get_occupancy(int32_t* workgroups_per_cu);
/* Launch-shape knobs, process-wide and atomic, read by each launch (value 0
 * restores a knob's default).  None changes what a kernel computes; the
 * split-K knobs change the order in which partial sums are added.
 *   SMX_KNOB_ADMM_OCCUPANCY       1..8, default 2: as smx_admm_set_occupancy.
 *   SMX_KNOB_GRAM_WAVES           4 or 8, default 8: waves per Gram workgroup
 *                                 (128 x 64 or 128 x 32 outputs per wave; two
 *                                 workgroups per CU either way).
 *   SMX_KNOB_INV_SYRKX_MIN        >= 256, default 16384: nodes of the FP64
 *                                 inverse with at least this many rows form
 *                                 their symmetric updates with rocBLAS syrkx,
 *                                 smaller ones with full dgemm products.
 *   SMX_KNOB_GEMM64_SPLIT_TARGET  256..4096, default 2048: workgroups the
 *                                 tiled FP64 GEMM's split-K rule fills
 *                                 (right-hand sides c = -Y U and the ADMM
 *                                 anchors, when they take the tiled kernel).
 *   SMX_KNOB_ADMM_LOADS           0, 1 or 2, default 0: how the ADMM product
 *                                 loads the inverse's planes: 0 non-temporal
 *                                 with one 128-row block of right-hand sides
 *                                 and a hi plane above 192 MiB, cached
 *                                 otherwise (row blocks then share M's tiles
 *                                 through L2, or the plane stays in the
 *                                 memory-side cache), 1 always non-temporal,
 *                                 2 always cached.
 *   SMX_KNOB_GEMM64_KERNEL        1 or 2, default 1: the kernel of the FP64
 *                                 products with few rows (c = -Y U, the ADMM
 *                                 anchors): 1 the thin kernel (operands
 *                                 straight into registers, one wave per SIMD),
 *                                 2 the 128 x 128-tiled kernel with split-K.
 *   SMX_KNOB_LIB_BINNING          1 or 2, default 1: how smx_build_library
 *                                 bins: 1 the integrator stores samples and
 *                                 one workgroup per column bins them into LDS
 *                                 histograms, 2 the integrator bins into
 *                                 per-integration slabs (the same bits). */
enum {
  SMX_KNOB_ADMM_OCCUPANCY = 0,
  SMX_KNOB_GRAM_WAVES = 1,
  SMX_KNOB_INV_SYRKX_MIN = 2,
  SMX_KNOB_GEMM64_SPLIT_TARGET = 3,
  SMX_KNOB_ADMM_LOADS = 4,
  SMX_KNOB_GEMM64_KERNEL = 5,
  SMX_KNOB_LIB_BINNING = 6,
  SMX_KNOB_COUNT = 7
};
int smx_tune_set(int32_t knob, int64_t value);
int smx_tune_get(int32_t knob, int64_t* value);
/* Delta-product bound tau in [0, 0.25] (default 2^-10): an ADMM iteration
 * multiplies only R - R_b by the inverse's fp16 hi plane while every row has
 * max|R_k - R_b,k| <= tau max|R_k|, and takes the full split product (and a
 * new base R_b) otherwise; 0 makes every product full.  The delta product's
 * error is about 2^-11 |M| |R - R_b|, so tau trades accuracy for bytes.
 * Process-wide, atomic; applies to solves started afterwards. */
int smx_admm_set_delta_threshold(double tau);
int smx_admm_get_delta_threshold(double* tau);
/* Mean time of `reps` ADMM product launches (full = 1: the full split
 * product; 0: the delta product) on the buffers of a workspace that a
 * smx_nnls_admm call of the same (n, m, K) has filled, with the current
 * occupancy.  Synchronises the stream. */
int smx_admm_product_time(int64_t n, int64_t m, int64_t K, void* workspace, int32_t full,
                          int32_t reps, float* ms_per_launch, void* stream);

/* ---- dense building blocks (no reference counterpart: the reference calls
 * XLA's dot / cholesky / cho_solve; exposed for tests) ------------------- */
/* C = alpha A B^T + beta C on MFMA, A (M,K) and B (N,K) row-major (k
 * contiguous), C (M,N); FP32 and FP64. */
int smx_gemm_nt_f32(int64_t M, int64_t N, int64_t K, float alpha, const float* A, int64_t lda,
                    const float* B, int64_t ldb, float beta, float* C, int64_t ldc, void* stream);
int smx_gemm_nt_f64(int64_t M, int64_t N, int64_t K, double alpha, const double* A, int64_t lda,
                    const double* B, int64_t ldb, double beta, double* C, int64_t ldc, void* stream);
/* C = alpha A B^T in FP64 as the NNLS forms c = -Y U and its anchors: with
 * work (work_elems doubles) for split-K partials, on the kernel that
 * SMX_KNOB_GEMM64_KERNEL selects (FP64 partials summed in split order). */
int smx_gemm_nt_f64_split(int64_t M, int64_t N, int64_t K, double alpha, const double* A,
                          int64_t lda, const double* B, int64_t ldb, double* C, int64_t ldc,
                          double* work, int64_t work_elems, void* stream);
/* In-place inverse of a full symmetric positive definite n x n FP64 matrix
 * (recursive Schur complements); work holds n*n + n doubles. */
int smx_spd_inverse_f64(double* A, int64_t n, double* work, void* stream);

/* ---- model maps and likelihood (model.py:14-68, likelihood.py:199-244) -- */
/* maps (K, m) = W (K, n) . lib[:, :m]^T in FP64. */
int smx_model_maps(const double* lib, int64_t ld_lib, int64_t n, int64_t m,
                   const double* W, int64_t K, double* maps, void* stream);

/* ------------------------------------------------------------------------
 * CylSpline potential (CylindricalSpline.py:269-536, CubicSpline.py).
 * ---------------------------------------------------------------------- */
#define SMX_CYLSPLINE_MAX_MODES 21

/* Mode tables of a CylSpline potential on the scaled grid u_i = asinh(R_i/R0),
 * v_j = asinh(z_j/R0).  coef[((i*nZ + j)*n_modes + m)*6 + k] holds, for mode m
 * at node (i, j), k = 0..5: Phi_m real, imaginary, their second derivatives
 * along u (the spline's M_x), then along v (M_y).  The reference keeps the
 * same numbers as separate (n_modes, nR, nZ) arrays (Phi_m_real, Phi_m_img,
 * Mx_real, Mx_img, My_real, My_img; CylindricalSpline.py:339-350).
 * n_modes_eval <= n_modes modes are summed (1: the axisymmetric part). */
typedef struct smx_cylspline {
  const double* u;
  const double* v;
  const double* coef;
  int32_t nR, nZ, n_modes, n_modes_eval;
  double R0;
  double G;
} smx_cylspline;

/* The builder's quadrature (CylindricalSpline.py:248-268, 318-330):
 * phi_m[(2m + part)*nR*nZ + i*nZ + j] = -G sum_k node_w[k]
 *   rho_nodes[(2m + part)*n_nodes + k] (Xi_m(R_i, Z_j; node_R[k], node_z[k])
 *   + Xi_m(R_i, Z_j; node_R[k], -node_z[k])), part 0 real, 1 imaginary;
 * node_w holds the Simpson weight times 2 pi R' times the map's Jacobian.
 * Xi_m = Q_{m-1/2}(chi)/(pi sqrt(R R')) as in jax_kernel_Xi_m
 * (CylindricalSpline.py:162-190). */
int smx_cylspline_phi_m(int32_t n_modes, const double* R, int32_t nR, const double* Z, int32_t nZ,
                        const double* node_R, const double* node_z, const double* node_w,
                        int64_t n_nodes, const double* rho_nodes, double G, double* phi_m,
                        void* stream);

/* Phi (n), acceleration -grad Phi (n, 3) and density nabla^2 Phi/(4 pi G) (n)
 * at points xyz (n, 3); any output may be NULL.  Replaces evaluate_phi,
 * evaluate_phi_axisymmetric, get_acc and get_density
 * (CylindricalSpline.py:355-416, 489-536). */
int smx_cylspline_eval(const smx_cylspline* t, const double* xyz, int64_t n, double* phi,
                       double* acc, double* rho, void* stream);

/* Test hook: Q_{m-1/2}(1 + delta) (jax_legendreQ, CylindricalSpline.py:145-160). */
int smx_legendre_q_half(int32_t m, const double* delta, int64_t n, double* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* SMX_H */
