"""Batched model evaluation: many parameter vectors theta, one galaxy.

The reference evaluates model_likelihood (likelihood.py:96-244) once per theta
inside its MCMC driver.  Here a batch of M thetas, each with its own orbit
count n_m (the first n_m orbits of dict_data), runs as:

* the Jeans ICs of all (model, star) pairs in one launch and the orbit
  libraries of all models in one persistent launch fed by a device work queue
  (``smx_build_library_batch``), on a side stream, group by group under a
  library-memory budget;
* on the main stream, overlapping the next group's integration (the
  integrator is FP64-VALU bound, the NNLS MFMA/HBM bound): per model the
  density target and the bootstrapped data, then the bootstrap NNLS of
  several models at once (all K right-hand sides of every model; one ADMM
  product launch and one update launch per iteration for the whole NNLS
  group, ``nnls.solve_group``, groups sized under an NNLS memory budget),
  then per model the maps and the chi^2.

Every model's value equals model_likelihood(theta_m, ...) on the first n_m
orbits.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from . import potentials as POT
from . import model as MD
from .likelihood import theta_to_params, grid_args, bootstrap_loglike


def _groups(n_orbits, row_bytes, budget):
    groups, cur, size = [], [], 0
    for m, n in enumerate(n_orbits):
        b = int(n) * row_bytes
        if cur and size + b > budget:
            groups.append(cur)
            cur, size = [], 0
        cur.append(m)
        size += b
    if cur:
        groups.append(cur)
    return groups


def _nnls_groups(models, n_orbits, m_rows, K, budget):
    """Consecutive runs of models whose NNLS solve states fit the budget
    (at most NNLS_GROUP_MAX per run, at least one; a size query, no GPU)."""
    L = _lib.load_library()
    groups, cur, size = [], [], 0
    for mi in models:
        b = int(L.smx_nnls_group_state_bytes(int(n_orbits[mi]), int(m_rows), int(K)))
        if cur and (size + b > budget or len(cur) == _lib.NNLS_GROUP_MAX):
            groups.append(cur)
            cur, size = [], 0
        cur.append(mi)
        size += b
    if cur:
        groups.append(cur)
    return groups


def model_likelihood_batch(thetas, dict_data, num_Vbin, Rzphi_n_tot, n_orbits=None,
                           sigma_amplify=1.0, schedule=None, nnls_maxiter=200, nnls_tol=0.0,
                           library_budget_bytes=48 << 30, nnls_budget_bytes=16 << 30,
                           grouped_nnls=True, autotune=True, return_aux=False):
    """logL for each row of thetas (M, 12); returns a CUDA tensor (M,).

    n_orbits: per-model orbit counts (default: all orbits for every model).
    schedule: one schedule dict, or a list of M (their N_max may differ).
    autotune: pick the NNLS stages' launch-shape knobs (ADMM product occupancy,
    Gram waves, inverse syrkx threshold, FP64 GEMM split target) for the
    batch's median shape first (``tune.ensure_tuned``, once per process and
    shape bucket).
    grouped_nnls: solve the NNLS of several models in grouped launches (the
    solve states of one NNLS group must fit nnls_budget_bytes); False, or
    nnls_tol > 0, solves model by model.  The default budget keeps groups
    small: a grouped product launch of many models spans several rounds of
    workgroups and keeps every SIMD busy, so the next library group's
    persistent integration (one full-register wave per SIMD, on the side
    stream) finds no free SIMD and stops overlapping the NNLS (20 models of
    n ~ U[5k, 40k]: 23.3 s with 80 GB groups, 17.6 s with 40 GB groups or
    model by model).
    """
    thetas = np.asarray(thetas, dtype=np.float64).reshape(-1, 12)
    M = thetas.shape[0]
    L = _lib.lib()
    w0_all = np.asarray(dict_data['w0'])[:, :3]
    n_all = w0_all.shape[0]
    n_orbits = [n_all] * M if n_orbits is None else [int(v) for v in n_orbits]
    if len(n_orbits) != M or min(n_orbits) < 1 or max(n_orbits) > n_all:
        raise ValueError("n_orbits must hold M counts in [1, number of orbits in dict_data]")
    K_rhs = np.asarray(dict_data['XY_standard_normal']).shape[0]
    if autotune:
        from . import tune
        m_rows = int(Rzphi_n_tot) + 5 * int(num_Vbin)
        tune.ensure_tuned(int(np.median(n_orbits)), m_rows, K_rhs)
    scheds = schedule if isinstance(schedule, (list, tuple)) else [schedule] * M
    scheds = [MD.schedule_of(s) for s in scheds]
    r = int(scheds[0]['n_realizations'])
    if any(int(s['n_realizations']) != r for s in scheds):
        raise ValueError("all models must use the same number of realizations")
    Rl, Rn, xyl, xyn = grid_args(dict_data)
    pots, grids, sstructs, pbs = [], [], [], []
    for m in range(M):
        ph, pb = theta_to_params(thetas[m])
        pbs.append(pb)
        pots.append(POT.pack_potential(MD.baryon_params(pb), ph))
        rot = MD.makeRotationMatrix(pb['alpha'], pb['beta'], pb['gamma'])
        grids.append(MD.make_grids(num_Vbin, Rn, Rl, xyl, xyn, rot, symmetry=True))
        sstructs.append(MD.make_schedule(scheds[m], -float(pb['Omega_bar'])))
    if grids[0].nR * grids[0].nz * grids[0].nphi != int(Rzphi_n_tot):
        raise ValueError("Rzphi_n_tot must equal prod(Rzphi_n_grid)")
    P, B = grids[0].nR * grids[0].nz * grids[0].nphi, grids[0].num_vbin
    rows = P + 6 * B
    dev = _lib.dev
    w0 = dev(w0_all)
    normals, uni = (dev(a) for a in MD.ic_draws(dict_data, r))
    r_total = uni.shape[1]
    bm = _lib.bin_map(dict_data['bin_mapping'], grids[0].nX, grids[0].nY)
    npb, v0, sd = dev(dict_data['num_per_bin']), _lib.velocities(dict_data['v0']), _lib.dispersions(dict_data['s'])

    main = torch.cuda.current_stream()
    side = torch.cuda.Stream()
    side.wait_stream(main)
    groups = _groups(n_orbits, rows * 8, library_budget_bytes)
    cols = 64 // r
    ws = None

    def launch(g):
        nonlocal ws
        k = len(g)
        items = sum((n_orbits[m] + cols - 1) // cols for m in g)
        need = int(L.smx_batch_workspace_bytes(grids[g[0]], k, items))
        with torch.cuda.stream(side):
            if ws is None or ws.numel() < need:
                ws = torch.empty(need, dtype=torch.uint8, device='cuda')
            out = []
            for m in g:
                n = n_orbits[m]
                out.append(dict(wb=torch.empty((n, r, 6), dtype=torch.float64, device='cuda'),
                                Tt=torch.empty((n, r), dtype=torch.float64, device='cuda'),
                                dt0=torch.empty((n, r), dtype=torch.float64, device='cuda'),
                                lib=torch.empty((n, rows), dtype=torch.float64, device='cuda')))
            arr = lambda T, vals: (T * k)(*vals)  # noqa: E731
            P_ = ctypes.c_void_p
            _lib.check(L.smx_build_library_batch(
                k, arr(_lib.SmxPotential, [pots[m] for m in g]),
                arr(_lib.SmxGrids, [grids[m] for m in g]),
                arr(_lib.SmxSchedule, [sstructs[m] for m in g]),
                _lib.ptr(w0), _lib.ptr(normals), _lib.ptr(uni),
                arr(ctypes.c_int64, [n_orbits[m] for m in g]), r, r_total,
                arr(ctypes.c_double, [float(scheds[m]['N_step_per_orb']) for m in g]),
                arr(ctypes.c_double, [float(scheds[m]['N_dynamical_time']) for m in g]),
                _lib.ptr(bm), _lib.ptr(npb), _lib.ptr(v0), _lib.ptr(sd),
                arr(P_, [o['wb'].data_ptr() for o in out]),
                arr(P_, [o['Tt'].data_ptr() for o in out]),
                arr(P_, [o['dt0'].data_ptr() for o in out]),
                arr(P_, [o['lib'].data_ptr() for o in out]),
                _lib.ptr(ws), ws.numel(), _lib.stream_ptr()), 'smx_build_library_batch')
            ev = torch.cuda.Event()
            ev.record(side)
        for o in out:
            for t in o.values():
                t.record_stream(main)
        return out, ev

    logl = torch.empty(M, dtype=torch.float64, device='cuda')
    aux = {}
    arena, nnls_subs = None, []
    if grouped_nnls and nnls_tol == 0.0:
        # one arena for every NNLS group of the batch, sized for the largest
        from .nnls import group_bytes
        nnls_subs = [_nnls_groups(g, n_orbits, P + 5 * B, K_rhs, nnls_budget_bytes) for g in groups]
        need = max(group_bytes([n_orbits[m] for m in sub], P + 5 * B, K_rhs)
                   for subs in nnls_subs for sub in subs)
        arena = torch.empty(need, dtype=torch.uint8, device='cuda')
    pending = launch(groups[0])
    for gi, g in enumerate(groups):
        out, ev = pending
        if gi + 1 < len(groups):
            pending = launch(groups[gi + 1])        # integrates while this group's NNLS runs
        main.wait_event(ev)
        libs = dict(zip(g, (o['lib'] for o in out)))
        cw = (np.sqrt(5.0), np.sqrt(5.0), np.sqrt(1.0))
        if grouped_nnls and nnls_tol == 0.0:
            for sub in nnls_subs[gi]:
                res = MD.models_from_libraries([libs[m] for m in sub], P, B, [pbs[m] for m in sub],
                                               dict_data, nnls_maxiter, 1.0, cw, arena=arena)
                for m, (md, info) in zip(sub, res):
                    logl[m] = bootstrap_loglike(md, dict_data, sigma_amplify)[0]
                    if return_aux:
                        aux[m] = dict(lib=libs[m], weights=md['weights'],
                                      nnls_iters=info['nnls_iters'])
                del res
        else:
            for m in g:
                md, info = MD.model_from_library(libs[m], P, B, pbs[m], dict_data, nnls_maxiter,
                                                 1.0, cw, nnls_tol)
                logl[m] = bootstrap_loglike(md, dict_data, sigma_amplify)[0]
                if return_aux:
                    aux[m] = dict(lib=libs[m], weights=md['weights'], nnls_iters=info['nnls_iters'])
        del out, libs
    del arena
    return (logl, aux) if return_aux else logl
