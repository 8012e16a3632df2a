"""Benchmark: SchwarMAX model evaluations per second on one MI355X.

Workload (default): the full-size plain evaluation — 20,000 orbit columns x 4
realizations, 2,000 BS3 attempts each (validity horizon 100 orbital times),
1,000 Voronoi apertures on a 60"x60" field, a 32x24x16 (R, phi, z) grid,
Gauss-Hermite h1-h4, ADMM NNLS on the 17,288 x 20,000 design matrix with 100
bootstrap right-hand sides and 1,000 iterations, then the bootstrap chi^2.
One step = one model_likelihood(theta) evaluation with all inputs resident in
HBM.  Synthetic galaxy (seeded); observations made from the same model's
library with Dirichlet weights plus 2% noise.

Prints one JSON line (rank 0).  Progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {'fp64': 78.6, 'fp32': 157.3, 'hbm': 8000.0}   # MI355X: TFLOP/s, TFLOP/s, GB/s


def log(*a):
    print('[bench]', *a, file=sys.stderr, flush=True)


def make_workload(config, seed, n=None):
    from schwarmax_amd import synth
    t0 = time.time()
    theta, dd, sched, info = synth.observed_workload(config, seed, n=n)
    log(f'workload {config}: n={dd["w0"].shape[0]} B={dd["total_bins"]} P={info["P"]} '
        f'K={dd["XY_standard_normal"].shape[0]} valid={info["valid"]:.3f} '
        f'setup {time.time() - t0:.1f}s')
    return theta, dd, sched


def _cpu_columns(job):
    """Worker of the CPU baseline: the oracle's ICs, integration and binning
    for orbit columns [c0, c0 + c) of the workload's galaxy (one thread)."""
    config, seed, c0, c = job
    import threadpoolctl
    from schwarmax_amd import synth
    from oracle import model as OM, integrate as OI
    with threadpoolctl.threadpool_limits(1):
        theta, dd, sched = synth.make_galaxy(config, seed=seed)
        ph, pb = OM.theta_to_params(theta)
        s = OM.schedule_of(sched)
        sub = {k: (v[c0:c0 + c] if k in ('w0', 'jeans_normals', 'jitter_uniforms') else v)
               for k, v in dd.items()}
        t0 = time.time()
        wb, Tt, dt0 = OM.initial_conditions(sub, OM.baryon_params(pb), ph, 4, s['N_step_per_orb'],
                                            s['N_dynamical_time'])
        rot = OM.make_rotation_matrix(pb['alpha'], pb['beta'], pb['gamma'])
        Rl, Rn, xyl, xyn = OM.grid_args(dd)
        og = OI.Grids(int(dd['total_bins']), dd['bin_mapping'], dd['num_per_bin'], Rl, xyl, Rn,
                      xyn, dd['v0'], dd['s'], rot)
        acc_fn, pot_fn = OI.make_acc_pot(OM.baryon_params(pb), ph)
        OI.integrate_chunked(wb.reshape(-1, 6), acc_fn, pot_fn, s['N_max'], Tt.reshape(-1),
                             dt0.reshape(-1), -pb['Omega_bar'], s['atol'], s['rtol'], s['dt_min'],
                             s['dt_max'], og, chunk_size=s['chunk'])
        return time.time() - t0


def _cpu_model():
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown'


def cpu_baseline(config, seed, cols_per_worker=16, n_nnls=1000, repeats=3):
    """The oracle (NumPy FP64) timed on the host on a bounded sample of the
    workload: a process pool over orbit columns (one process per core of
    this process's CPU share, at most 16) for the initial conditions,
    integration and binning, and BLAS/LAPACK on the same cores for an NNLS
    of the workload's m rows, K right-hand sides and T iterations at
    n_nnls columns; 1 warm-up and `repeats` timed runs of each, median.

    Must run before the process touches the GPU: the pool forks it."""
    import multiprocessing
    import threadpoolctl
    from schwarmax_amd import synth
    from oracle import nnls as ON
    theta, dd, sched = synth.make_galaxy(config, seed=seed)
    cfg = synth.CONFIGS[config]
    n_full = dd['w0'].shape[0]
    workers = max(1, min(16, len(os.sched_getaffinity(0))))
    jobs = [(config, seed, w * cols_per_worker, cols_per_worker) for w in range(workers)]
    cols = workers * cols_per_worker

    def timed(fn):
        fn()                                       # warm-up
        return float(np.median([fn() for _ in range(repeats)]))

    with multiprocessing.get_context('fork').Pool(workers) as pool:
        def run_cols():
            t0 = time.time()
            pool.map(_cpu_columns, jobs)
            return time.time() - t0
        t_int = timed(run_cols)
    # NNLS sample: n_nnls columns, the workload's m rows, K rhs, T iterations.
    # Each part is scaled by its own order in n: Gram U^T U ~ n^2 m, Cholesky
    # ~ n^3, one ADMM iteration (two triangular solves per rhs) ~ n^2 K.
    rng = np.random.default_rng(1)
    nR, nz, nphi = (int(v) for v in dd['Rzphi_n_grid'])
    m = nR * nz * nphi + 5 * int(dd['total_bins'])
    K, T = cfg['K'], cfg['T']
    U = rng.random((m, n_nnls)) * (rng.random((m, n_nnls)) < 0.05)
    Y = rng.random((K, m))
    parts = {}

    def run_nnls():
        t0 = time.time()
        Q = U.T @ U
        t1 = time.time()
        _, rho, Lc = ON.gram_factor(U, 1.0)
        t2 = time.time()
        ON.admm(Lc, rho, -(Y @ U), 1.0, T)
        t3 = time.time()
        parts.setdefault('gram', []).append(t1 - t0)
        parts.setdefault('fac', []).append(max(t2 - t1 - (t1 - t0), 0.0))
        parts.setdefault('admm', []).append(t3 - t2)
        del Q
        return t3 - t0
    with threadpoolctl.threadpool_limits(workers):
        timed(run_nnls)
    t_gram, t_fac, t_admm = (float(np.median(parts[k][1:])) for k in ('gram', 'fac', 'admm'))
    scale_n = n_full / cols
    r = n_full / n_nnls
    est = t_int * scale_n + t_gram * r ** 2 + t_fac * r ** 3 + t_admm * r ** 2
    sample = (f"oracle (NumPy FP64) on {cols} of {n_full} columns x4 realizations "
              f"x{sched['N_max']} steps in {workers} processes ({t_int:.1f} s, x{scale_n:.1f}) and "
              f"NNLS at n={n_nnls}, m={m}, K={K}, T={T} on {workers} BLAS threads (Gram "
              f"{t_gram:.2f} s x{r:.0f}^2, Cholesky {t_fac:.2f} s x{r:.0f}^3, ADMM {t_admm:.1f} s "
              f"x{r:.0f}^2); median of {repeats} after 1 warm-up; extrapolated {est:.0f} s/eval. "
              f"Reference's own CPU figure: ~60-100 s/eval at 7,500 columns, float32 "
              f"(README.md:92-93)")
    return {'value': 1.0 / est, 'unit': 'model evals/s', 'cores': workers, 'kind': 'port',
            'sample': sample, 'nproc': os.cpu_count(), 'cpu_model': _cpu_model(),
            's_per_eval': est}


def traffic_per_launch(config, full_frac):
    """HBM bytes per ADMM-product launch from the committed rocprofv3 --pmc
    summary for this workload (FETCH_SIZE + WRITE_SIZE per full and per delta
    launch, each scaled by the same counter's reading on a copy of known
    size), weighted by this run's share of full launches; or None."""
    path = os.path.join(ROOT, 'profiles', f'admm_gemm_kernel.{config}.traffic.json')
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return full_frac * t['bytes_per_full_launch'] + (1.0 - full_frac) * t['bytes_per_delta_launch']


def bench_jackknife(args):
    """calculate_delta_chi2 (likelihood.py:248-335): one group per aperture."""
    import torch
    from schwarmax_amd import _lib, jackknife as J
    theta, dd, sched = make_workload(args.config, args.seed)
    B = int(dd['total_bins'])
    P = int(dd['Rzphi_n_tot'])

    def step():
        return J.calculate_delta_chi2(theta, dd, B, P, schedule=sched, nnls_maxiter=200,
                                      return_aux=True)
    for i in range(args.warmup):
        t0 = time.time()
        step()
        torch.cuda.synchronize()
        log(f'warmup {i}: {time.time() - t0:.1f}s')
    _lib.timing_enable(True)
    torch.cuda.synchronize()
    t0 = time.time()
    for i in range(args.steps):
        v, aux = step()
    torch.cuda.synchronize()
    dt = time.time() - t0
    stages = _lib.timing_read()
    _lib.timing_enable(False)
    log(f'delta chi2 {float(v):.6g}; {dt / args.steps:.2f} s; Neumann terms {aux["neumann_terms"]}; '
        f'ADMM operator products {aux["products"]} ({aux["full_products"]} full)')
    out = {
        'metric': 'jackknife delta-chi2 evals/s (one group per aperture, 1000 replicates)',
        'value': args.steps / dt, 'unit': 'evals/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': 1e3 * dt / args.steps, 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64/f32', 'data': 'synthetic',
        'config': {'workload': args.config, 'orbits': int(dd['w0'].shape[0]), 'replicates': B,
                   'apertures': B, 'rzphi_cells': P, 'admm_iters': 200,
                   'neumann_terms': aux['neumann_terms'],
                   'operator_products': aux['products'], 'full_products': aux['full_products'],
                   'max_shift_over_rho': float(aux['delta'].max()) / aux['rho']},
        'stages_ms': {k: round(v[0] / args.steps, 3) for k, v in stages.items()},
    }
    print(json.dumps(out))


def bench_batched(args):
    """Batched evaluation: M thetas with uneven orbit counts in one call."""
    import torch
    from schwarmax_amd import _lib, synth, batch, likelihood as L
    cfg = synth.CONFIGS[args.config]
    lo, hi = cfg['n']
    theta, dd, sched = make_workload(args.config, args.seed, n=hi)
    rng = np.random.default_rng(args.seed + 1)
    M = cfg['n_models']
    thetas = np.stack(synth.batched_thetas(M, rng))
    n_orbits = [int(v) for v in rng.integers(lo, hi + 1, size=M)]
    B = int(dd['total_bins'])
    P = int(dd['Rzphi_n_tot'])
    T = cfg['T']

    def step():
        return batch.model_likelihood_batch(thetas, dd, B, P, n_orbits=n_orbits, schedule=sched,
                                            nnls_maxiter=T, nnls_tol=args.nnls_tol,
                                            grouped_nnls=not args.per_model_nnls,
                                            nnls_budget_bytes=int(args.nnls_budget_gib * 2**30))
    # the launch-shape knobs for the batch's median shape, once per process
    # (the batch call would otherwise tune inside the first timed step)
    from schwarmax_amd import tune
    t0 = time.time()
    tune.ensure_tuned(int(np.median(n_orbits)), P + 5 * B, dd['XY_standard_normal'].shape[0])
    log(f'autotuned in {time.time() - t0:.1f}s')
    for i in range(args.warmup):
        t0 = time.time()
        step()
        torch.cuda.synchronize()
        log(f'warmup {i}: {time.time() - t0:.1f}s')
    _lib.timing_enable(True)
    torch.cuda.synchronize()
    t0 = time.time()
    for i in range(args.steps):
        v = step()
    torch.cuda.synchronize()
    dt = time.time() - t0
    stages = _lib.timing_read()
    _lib.timing_enable(False)
    log(f'batch of {M}: {dt / args.steps:.2f} s; logL finite {int(torch.isfinite(v).sum())}/{M}')
    out = {
        'metric': 'batched model evals/s (64 theta, n ~ U[5k, 40k])', 'value': M * args.steps / dt,
        'unit': 'model evals/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': 1e3 * dt / args.steps, 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f64/f32', 'data': 'synthetic',
        'config': {'workload': args.config, 'models': M, 'orbits_total': int(sum(n_orbits)),
                   'orbits_min': min(n_orbits), 'orbits_max': max(n_orbits),
                   'realizations': sched['n_realizations'], 'bs3_attempts': sched['N_max'],
                   'apertures': B, 'rzphi_cells': P, 'bootstrap_rhs': dd['XY_standard_normal'].shape[0],
                   'admm_iters': T},
        'stages_ms': {k: round(v[0] / args.steps, 3) for k, v in stages.items()},
    }
    out['nnls_grouping'] = ('model by model' if args.per_model_nnls else
                            f'grouped solves of <= {args.nnls_budget_gib:g} GiB of state')
    from schwarmax_amd import tune
    out['autotune'] = [{'shape': list(r['shape']), 'occupancy': r['occupancy'],
                        'default': r['default'],
                        'ms_full_delta_weighted': {str(o): [round(x, 4) for x in v]
                                                   for o, v in r['ms'].items()},
                        'admm_loads': r['loads'], 'gram_waves': r['factor']['gram_waves'],
                        'inv_syrkx_min': r['factor']['syrkx_min'],
                        'gemm64_kernel': r['factor']['gemm64_kernel'],
                        'gemm64_split_target': r['factor']['split_target']}
                       for r in tune._RESULTS.values()]
    if args.compare:
        t0 = time.time()
        for m in range(M):
            dm = dict(dd)
            for k in ('w0', 'jeans_normals', 'jitter_uniforms'):
                dm[k] = np.asarray(dd[k])[:n_orbits[m]]
            L.model_likelihood(thetas[m], dm, B, P, schedule=sched, nnls_maxiter=T,
                               nnls_tol=args.nnls_tol)
            torch.cuda.synchronize()
            log(f'plain model {m}: n={n_orbits[m]} {time.time() - t0:.1f}s')
        seq = time.time() - t0
        out['plain_sequential_evals_per_s'] = M / seq
        out['batched_over_plain'] = (M * args.steps / dt) / (M / seq)
    print(json.dumps(out))


def bench_batched_grad(args):
    """Forward + adjoint of M thetas, each a full-size model (gradients/s)."""
    import torch
    from schwarmax_amd import _lib, synth, grad as GR
    cfg = synth.CONFIGS[args.config]
    theta, dd, sched = make_workload(args.config, args.seed)
    rng = np.random.default_rng(args.seed + 1)
    M = cfg['n_models']
    thetas = np.stack(synth.batched_thetas(M, rng))
    B = int(dd['total_bins'])
    P = int(dd['Rzphi_n_tot'])
    T = cfg['T']

    def step():
        return GR.model_likelihood_and_grad_batch(thetas, dd, B, P, schedule=sched,
                                                  nnls_maxiter=T, nnls_tol=args.nnls_tol)
    from schwarmax_amd import tune
    t0 = time.time()
    tune.ensure_tuned(int(dd['w0'].shape[0]), P + 5 * B, dd['XY_standard_normal'].shape[0])
    log(f'autotuned in {time.time() - t0:.1f}s')
    for i in range(args.warmup):
        t0 = time.time()
        step()
        torch.cuda.synchronize()
        log(f'warmup {i}: {time.time() - t0:.1f}s')
    torch.cuda.reset_peak_memory_stats()
    _lib.timing_enable(True)
    torch.cuda.synchronize()
    t0 = time.time()
    for i in range(args.steps):
        v, g = step()
    torch.cuda.synchronize()
    dt = time.time() - t0
    stages = _lib.timing_read()
    _lib.timing_enable(False)
    peak = torch.cuda.max_memory_allocated()
    ok_v = int(torch.isfinite(v).sum())
    ok_g = int(torch.isfinite(g).all(dim=1).sum())
    log(f'batch of {M} gradients: {dt / args.steps:.1f} s; finite logL {ok_v}/{M}, '
        f'finite gradients {ok_g}/{M}; peak HBM {peak / 2**30:.1f} GiB')
    out = {
        'metric': f'batched gradient evals/s ({M} theta, {int(dd["w0"].shape[0])} orbits each)',
        'value': M * args.steps / dt, 'unit': 'gradient evals/s (logL and dlogL/dtheta)',
        'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': 1e3 * dt / args.steps, 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f64/f32', 'data': 'synthetic',
        'config': {'workload': args.config, 'models': M, 'orbits': int(dd['w0'].shape[0]),
                   'realizations': sched['n_realizations'], 'bs3_attempts': sched['N_max'],
                   'apertures': B, 'rzphi_cells': P,
                   'bootstrap_rhs': dd['XY_standard_normal'].shape[0], 'admm_iters': T,
                   'ckpt_every': 64},
        'peak_hbm_bytes': int(peak), 'finite_logl': ok_v, 'finite_gradients': ok_g,
        'stages_ms': {k: round(v_[0] / args.steps, 3) for k, v_ in stages.items()},
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--config', default='plain-20k')
    ap.add_argument('--seed', type=int, default=0)
    ap.add_argument('--nnls-tol', type=float, default=0.0)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--nnls-budget-gib', type=float, default=16.0,
                    help='batched config: solve state per grouped NNLS solve')
    ap.add_argument('--per-model-nnls', action='store_true',
                    help='batched config: solve the NNLS model by model')
    ap.add_argument('--no-autotune', action='store_true',
                    help='plain / grad configs: keep the default launch-shape knobs')
    ap.add_argument('--compare', action='store_true',
                    help='batched configs: also time the same models one by one (plain path)')
    ap.add_argument('--calibrate-copy', type=int, default=0,
                    help='only run a device copy of this many MiB (HBM counter calibration)')
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: only the "
              f"single-GPU path is built (it would shard by orbit columns, each GPU "
              f"integrating a share of the orbits, with one all-reduce of the Gram)",
              file=sys.stderr)
        sys.exit(2)
    cpu = None
    if not args.no_cpu and not args.calibrate_copy and args.config in ('plain-20k', 'plain-2k'):
        # before anything touches the GPU: the baseline's process pool forks
        log('cpu baseline ...')
        cpu = cpu_baseline(args.config, args.seed)
        log(f"cpu baseline: {cpu['s_per_eval']:.0f} s/eval on {cpu['cores']} cores")
    import torch
    assert torch.cuda.is_available(), 'needs a GPU'
    if args.calibrate_copy:
        nbytes = args.calibrate_copy << 20
        # one elementwise kernel per pass (b = a * 1): reads and writes nbytes;
        # no fill kernel, so the counters see only these dispatches
        a = torch.empty(nbytes // 4, dtype=torch.float32, device='cuda')
        b = torch.empty_like(a)
        for _ in range(5):
            torch.mul(a, 1.0, out=b)
        torch.cuda.synchronize()
        print(json.dumps({'copy_bytes': nbytes}))
        return
    from schwarmax_amd import _lib, likelihood as L, nnls as NN
    from oracle import opcount
    from oracle import model as OM
    if args.config.startswith('batched-grad'):
        return bench_batched_grad(args)
    if args.config.startswith('batched'):
        return bench_batched(args)
    if args.config.startswith('jackknife'):
        return bench_jackknife(args)
    theta, dd, sched = make_workload(args.config, args.seed)
    B = int(dd['total_bins'])
    P = int(dd['Rzphi_n_tot'])
    K = dd['XY_standard_normal'].shape[0]
    T = 1000 if args.config != 'plain-2k' else 500
    n = dd['w0'].shape[0]

    grad_mode = args.config.startswith('grad')
    if grad_mode:
        from schwarmax_amd import grad as GR

    def forward():
        return L.model_likelihood(theta, dd, B, P, schedule=sched, nnls_maxiter=T,
                                  nnls_tol=args.nnls_tol)

    def step():
        if grad_mode:
            v, g = GR.model_likelihood_and_grad(theta, dd, B, P, schedule=sched, nnls_maxiter=T,
                                                nnls_tol=args.nnls_tol, ckpt_every=64)
            return v
        return forward()

    tuning = None
    if not args.no_autotune:
        # the NNLS stages' launch-shape knobs for this shape, before any timing
        from schwarmax_amd import tune
        t0 = time.time()
        res = tune.ensure_tuned(n, P + 5 * B, K)
        f = res['factor']
        tuning = {'admm_occupancy': res['occupancy'], 'admm_loads': res['loads'],
                  'gram_waves': f['gram_waves'],
                  'inv_syrkx_min': f['syrkx_min'], 'gemm64_kernel': f['gemm64_kernel'],
                  'gemm64_split_target': f['split_target']}
        log(f'autotuned {tuning} in {time.time() - t0:.1f}s')
    for i in range(args.warmup):
        t0 = time.time()
        v = step()
        torch.cuda.synchronize()
        log(f'warmup {i}: logL={float(v):.6f} {time.time() - t0:.2f}s')
    _lib.timing_enable(True)
    torch.cuda.synchronize()
    t0 = time.time()
    vals = []
    for i in range(args.steps):
        vals.append(step())
    torch.cuda.synchronize()
    dt = time.time() - t0
    stages = _lib.timing_read()
    _lib.timing_enable(False)
    log('logL', [float(v) for v in vals], f'{dt / args.steps:.3f} s/eval')
    fwd_s = None
    if grad_mode:
        torch.cuda.synchronize()
        t1 = time.time()
        forward()
        torch.cuda.synchronize()
        fwd_s = time.time() - t1
        log(f'forward alone {fwd_s:.3f} s; gradient / forward = {dt / args.steps / fwd_s:.2f}')
    log('stages (ms per eval):', {k: round(v[0] / args.steps, 2) for k, v in stages.items()})

    # Roofline of the dominant kernel: whichever of the ADMM product (HBM
    # bound: both fp16 planes of M once, R in, the product out) and the
    # integrator (FP64 arithmetic) takes the larger share of the step.
    ph, pb = OM.theta_to_params(theta)
    _, f_acc, _ = opcount.force_flops(OM.baryon_params(pb), ph)
    N = sched['N_max']
    n_int = n * sched['n_realizations']
    int_ms = stages['integrate'][0] / max(stages['integrate'][1], 1)
    int_flops = n_int * N * (3 * f_acc + opcount.F_STEP)
    prod_ms = stages['admm_product'][0] / max(stages['admm_product'][1], 1)
    # algorithmic bytes per launch: both fp16 planes of M (4 n^2) on a full
    # product, the hi plane (2 n^2) on a delta product; the operand R or D in
    # fp16 hi + lo (4 K n) and the FP32 product out (4 K n)
    # launches: every ADMM iteration but the anchor refreshes (every 100th),
    # whose product is an FP64 GEMM instead
    n_prod = max(stages['admm_product'][1], 1)
    full_frac = NN.STATS['full_products'] / n_prod
    prod_bytes = (4.0 * full_frac + 2.0 * (1.0 - full_frac)) * n * n + 4.0 * K * n + 4.0 * K * n
    if stages['integrate'][0] >= stages['admm_product'][0]:
        ach = int_flops / (int_ms * 1e-3) / 1e12
        roof = {'bound': 'mfma', 'achieved': ach, 'peak': PEAK['fp64'], 'unit': 'TFLOP/s',
                'frac': ach / PEAK['fp64'], 'traffic': None,
                'kernel': 'integrate_library_kernel (FP64)', 'per_launch': int_flops,
                'ms_per_launch': int_ms}
    else:
        ach = prod_bytes / (prod_ms * 1e-3) / 1e9
        roof = {'bound': 'hbm', 'achieved': ach, 'peak': PEAK['hbm'], 'unit': 'GB/s',
                'frac': ach / PEAK['hbm'], 'traffic': traffic_per_launch(args.config, full_frac),
                'kernel': 'admm_gemm_kernel', 'per_launch': prod_bytes,
                'ms_per_launch': prod_ms, 'launches': stages['admm_product'][1],
                'full_products': NN.STATS['full_products'], 'products': NN.STATS['products']}
    out = {
        'metric': ('gradient evals/s @20k orbits' if grad_mode else
                   'model evals/s @20k orbits (plain)'), 'value': args.steps / dt,
        'unit': 'model evals/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': 1e3 * dt / args.steps, 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f64/f32', 'data': 'synthetic',
        'config': {'workload': args.config, 'orbits': n, 'realizations': sched['n_realizations'],
                   'bs3_attempts': N, 'apertures': B, 'rzphi_cells': P, 'bootstrap_rhs': K,
                   'admm_iters': T, 'nnls_tol': args.nnls_tol},
        'stages_ms': {k: round(v[0] / args.steps, 3) for k, v in stages.items()},
        'roofline': roof,
    }
    if tuning is not None:
        out['tuning'] = tuning
    if grad_mode:
        out['gradient_to_forward'] = (dt / args.steps) / fwd_s
        out['config']['ckpt_every'] = 64
    if cpu is not None:
        out['cpu_baseline'] = cpu
    print(json.dumps(out))


if __name__ == '__main__':
    main()
