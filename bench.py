"""Benchmark: compressible FV step + discrete adjoint on one MI355X.

Workload: 3-D Taylor-Green vortex on a periodic 128^3 hex box of side 2*pi,
stored as an unstructured face list (seeded random cell permutation, then
reverse Cuthill-McKee), Mach 0.1, rho0 = 1.2, Re = 1600, Roe flux, SSP-RK3,
fixed dt.  One "step" is one time step forward plus its reverse (adjoint)
step: K forward steps keep their start states and step tapes (intermediate
stage states and stage gradients) in HBM, then K adjoint steps run in
reverse from them.  The rate with tapes off (each adjoint step recomputing
its stages, as the reference's reverse sweep does) is reported beside it.

Reported value: FP32 primal+adjoint cell-updates/s (nCells * K / time).
Also reported: the same for FP64, primal-only rates, the FP32-vs-FP64
relative L2 error per conserved variable after --l2-steps primal steps, the
roofline of the dominant kernel (HIP events on the library's stream) and the
CPU path (a C++/OpenMP FP64 step and adjoint step) timed on a bounded sample
of the same workload.

    python bench.py [--gpus 1] [--steps 20] [--warmup 2]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s


def kernel_bytes(name, s, nC, nF, nG, nO=None):
    """Algorithmic bytes one launch moves: each array element the kernel
    needs counted once (s = scalar bytes).  Cell slot tables: 6 x (int32
    face, int32 neighbour, uint8 flags) = 54 B per cell, or 30 B without the
    neighbour column; face tables: int32 owner + int32 neighbour + uint8 kind
    = 9 B per face; the record indices of a face's adjoint (int32) 4 B per
    record written by face.  Face adjoint records (20 scalars): with nO cells
    owning internal faces (owner sides summed per cell) one per internal
    face, one per such cell and two per cyclic face; otherwise two per face
    (nO None).  Every boundary face here is cyclic (periodic box)."""
    if name == 'grad_kernel':          # u, V, face (A, n, w) -> 15 gradients
        return nC * (5 + 1 + 15) * s + nF * 5 * s + 54 * nC
    if name == 'face_flux_kernel':     # u, grad, 16 face rows (all but the weight) -> 6 face values
        return nC * (5 + 15) * s + nF * (16 + 6) * s + 9 * nF
    if name == 'stage_kernel':         # face values, V, u_i, u0 (2 of 3 stages) -> u_{i+1}
        return nF * 6 * s + nC * (1 + 5 + 10. / 3 + 5) * s + 30 * nC
    recs = 2 * nF if nO is None else (nF - nG) + nO + 2 * nG
    idx = 8 * nF if nO is None else 4 * (nF - nG) + 8 * nG
    if name == 'face_vjp_kernel':      # u, grad, ubar, V, 16 face rows -> face adjoint records
        return nC * (5 + 15 + 5 + 1) * s + nF * 16 * s + recs * 20 * s + 9 * nF + idx
    if name == 'vjp_gather_kernel':    # face adjoint records -> wbar, gbar; slot flags
        return recs * 20 * s + nC * 20 * s + 6 * nC
    if name == 'grad_vjp_kernel':      # u, wbar, gbar, ubar, acc (2 of 3), V, 6 x 4 slot coefficients -> ubar, acc
        return nC * (5 + 5 + 15 + 5 + 10. / 3 + 1 + 5 + 10. / 3 + 24) * s + 30 * nC
    return None


def kernel_rooflines(prof, s, nC, nF, nG, nO, triad):
    """Per kernel: average launch (ms, HIP events), algorithmic bytes per
    launch, their rate (GB/s) and its fraction of the 8 TB/s peak and of the
    measured triad bandwidth (GB/s)."""
    out = {}
    for k, (ms_tot, launches) in prof.items():
        b = kernel_bytes(k, s, nC, nF, nG, nO)
        if not b or not launches:
            continue
        ms = ms_tot / launches
        gbs = b / (ms * 1e-3) / 1e9
        out[k] = {'avg_launch_ms': round(ms, 4), 'algorithmic_bytes': int(b), 'GBs': round(gbs, 1),
                  'frac_peak': round(gbs / (HBM_PEAK / 1e9), 3), 'frac_triad': round(gbs / triad, 3)}
    return out


def step_bytes_model(s, nC, nF, nG):
    """Primal and primal+adjoint bytes per cell-update, the step-level byte model:
    B_P = s[nC(5+5+10/3+15+15+1) + 40 nG + 17 nF] + 8 nF per stage,
    B_A = s[nC(5+15+5+5+20/3+15+15+1) + 40 nG + 17 nF] + 8 nF per stage."""
    BP = s * (nC * (5 + 5 + 10. / 3 + 15 + 15 + 1) + 40 * nG + 17 * nF) + 8 * nF
    BA = s * (nC * (5 + 15 + 5 + 5 + 20. / 3 + 15 + 15 + 1) + 40 * nG + 17 * nF) + 8 * nF
    return 3 * BP / nC, 3 * (BP + BA) / nC


def run_precision(case, u0, prec, K, W, dt, torch, profile, recompute=False):
    """K steps forward keeping each step's start state and tape (stage
    states and gradients) in HBM, then K adjoint steps in reverse.  With
    recompute, no tapes: each adjoint step re-evaluates its stages."""
    ctx = case.context(prec)
    nC = ctx.nIC
    u = torch.as_tensor(u0.astype(prec), device='cuda').contiguous()
    tuned = ctx.autotune(u, reps=3)          # block size per stencil kernel (results do not depend on it)
    store = torch.empty((K, 5, nC), dtype=ctx.tdtype, device='cuda')
    tapes = None if recompute else ctx.new_tapes(K)
    ua = torch.zeros_like(u)

    def pa_pass(k, prof=False):
        if prof:
            ctx.profile_start()
        for i in range(k):
            store[i].copy_(u)
            ctx.primal_step(u, dt, tape=None if tapes is None else tapes[i])
        ua.zero_()
        for i in reversed(range(k)):
            ctx.adjoint_step(store[i], dt, ua, obja=1.0, tape=None if tapes is None else tapes[i])
        return ctx.profile_stop() if prof else None

    pa_pass(min(W, K))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pa_pass(K)
    torch.cuda.synchronize()
    t_pa = time.perf_counter() - t0
    # per-kernel times from HIP events around every launch, in a pass of
    # their own: the event records lengthen a step by about 4%
    prof = pa_pass(K, prof=True) if profile else None
    ok = bool(torch.isfinite(u).all()) and bool(torch.isfinite(ua).all())
    # primal only
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(K):
        ctx.primal_step(u, dt)
    torch.cuda.synchronize()
    t_p = time.perf_counter() - t0
    streaming = ctx.streaming
    del ctx
    return dict(t_pa=t_pa, t_p=t_p, prof=prof, finite=ok, tuned={k: v[0] for k, v in tuned.items()},
                streaming=streaming)


def l2_error(case, u0, steps, dt, torch):
    out = {}
    for prec in (np.float32, np.float64):
        ctx = case.context(prec)
        u = torch.as_tensor(u0.astype(prec), device='cuda').contiguous()
        for _ in range(steps):
            ctx.primal_step(u, dt)
        out[prec] = u.double().cpu().numpy()
        del ctx
    a, b = out[np.float32], out[np.float64]
    V = case.mesh.volumes[:, 0]
    err = {}
    for name, sl in (('rho', slice(0, 1)), ('rhoU', slice(1, 4)), ('rhoE', slice(4, 5))):
        d = ((a[sl] - b[sl]) ** 2 * V).sum()
        n = (b[sl] ** 2 * V).sum()
        err[name] = float(np.sqrt(d / n))
    return err


def triad_bandwidth(torch, n=1 << 28, reps=10):
    """Device-memory bandwidth a plain stream reaches: c = a + 0.5 b on three
    n-element FP32 arrays (1 GiB each by default), reads 2n and writes n
    scalars per launch, timed with HIP events on torch's stream (GB/s)."""
    a = torch.ones(n, dtype=torch.float32, device='cuda')
    b = torch.ones_like(a)
    c = torch.empty_like(a)
    for _ in range(3):
        torch.add(a, b, alpha=0.5, out=c)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        torch.add(a, b, alpha=0.5, out=c)
    e1.record()
    torch.cuda.synchronize()
    s = e0.elapsed_time(e1) * 1e-3 / reps
    del a, b, c
    return 3 * n * 4 / s / 1e9


def cpu_model():
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(mesh, mu0, u0, dt, steps=8):
    """The CPU path timed on the host: the C++/OpenMP FP64 step and adjoint
    step (oracle/cpu_step.cpp, built -O3 -fopenmp) on the same mesh, initial
    state and dt as the GPU run, `steps` steps forward keeping their start
    states, then `steps` adjoint steps (each re-running its stages, as the
    reference's reverse sweep re-runs the primal), on every OpenMP thread
    (OMP_NUM_THREADS)."""
    from oracle import cpu_baseline as cb
    cs = cb.CpuStep(mesh, mu0)
    u = np.ascontiguousarray(u0, np.float64)
    w = u.copy()
    cs.step(w, dt)                                   # warm-up (first touch of the work arrays)
    t0 = time.perf_counter()
    states = []
    for _ in range(steps):
        states.append(u.copy())
        cs.step(u, dt)
    t1 = time.perf_counter()
    ua = np.zeros_like(u)
    for k in reversed(range(steps)):
        cs.adjoint_step(states[k], dt, ua, 1.0)
    t2 = time.perf_counter()
    nC = mesh.nInternalCells
    threads = cs.threads()
    return dict(value=nC * steps / (t2 - t0), unit='cell-updates/s primal+adjoint', cores=threads, kind='port',
                sample='C++/OpenMP FP64 step + adjoint step (oracle/cpu_step.cpp), TGV %d cells (the bench mesh), '
                       '%d steps forward + %d reverse at the bench dt, %d threads, %.1f s'
                       % (nC, steps, steps, threads, t2 - t0),
                primal_value=nC * steps / (t1 - t0), precision='fp64', cpu_model=cpu_model(),
                os_cpu_count=os.cpu_count(), omp_num_threads=os.environ.get('OMP_NUM_THREADS'),
                finite=bool(np.isfinite(ua).all() and np.isfinite(u).all()))


def ensemble_bench(torch, n_instances=64, n_steps=200, interval=20, seed=0, overlap=None):
    """Batched ensemble: n_instances Taylor-Green boxes, sizes drawn by seed
    from {32, 48, 64, 96}^3, +-1% seeded velocity perturbation, advanced
    primal + adjoint as one concatenated mesh, the whole sweep one captured
    HIP graph (one launch), FP32 state, FP64 per-instance objectives."""
    from adfvm_amd.ensemble import TaylorGreenEnsemble
    t0 = time.time()
    ens = TaylorGreenEnsemble(n_instances=n_instances, seed=seed)
    print('[bench] ensemble of %d instances, %d cells, built in %.1fs' % (n_instances, ens.n_cells,
                                                                           time.time() - t0),
          file=sys.stderr, flush=True)
    tuned = ens.ctx.autotune(ens.u0, reps=2)
    ens.run(n_steps=n_steps, interval=interval, overlap=overlap)  # warm-up: captures the sweep's graph, runs it
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    J, ua = ens.run(n_steps=n_steps, interval=interval, overlap=overlap)
    t = time.perf_counter() - t0
    ok = bool(np.isfinite(J).all()) and bool(torch.isfinite(ua).all())
    value = ens.n_cells * n_steps / t
    return {'metric': 'cell-updates/s primal+adjoint, FP32 vs FP64, % HBM roofline, L2 err vs FP64',
            'value': value, 'unit': 'cell-updates/s', 'n_gpus': 1, 'steps': n_steps, 'warmup': 1,
            'ms_per_step': t / n_steps * 1e3, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
            'dtype': 'f32', 'data': 'synthetic (Taylor-Green initial fields, seeded +-1% velocity perturbation)',
            'config': {'workload': 'batched ensemble: %d Taylor-Green instances, sizes %s^3 drawn by seed %d, '
                                   'one concatenated mesh, checkpoints every %d steps, the whole sweep one '
                                   'HIP graph launch' % (n_instances, sorted(set(ens.sizes)), seed, interval),
                       'n_cells': ens.n_cells, 'instances': n_instances,
                       'size_counts': {str(k): ens.sizes.count(k) for k in sorted(set(ens.sizes))},
                       'dt': ens.dt},
            'block_sizes': {k: v[0] for k, v in tuned.items()}, 'finite': ok,
            'overlapped_reverse': ens.last_sweep.overlap,
            'objective_range': [float(J.min()), float(J.max())]}


def adjoint64_bench(torch, n=64, n_steps=400, interval=20, overlap=None):
    """Taylor-Green n^3 adjoint: n_steps fixed-dt steps (CFL 0.5) primal
    plus the adjoint of the kinetic energy summed over the steps (the
    reference's objective hook on the stage-1 state), uniform checkpoints
    every `interval` steps, each checkpoint segment one HIP graph; FP32
    kernels with FP64 objective accumulation, and the same run in FP64 for
    the error and speed ratio.  The same sweeps with the objective on each
    step's start state u_n are reported beside them (objective_state
    'start')."""
    from adfvm_amd.graphs import GraphSweep
    from tests import casefactory as cf
    t0 = time.time()
    case, params = cf.tgv(n, seed=0)
    u0h = cf.tgv_state(case, params)
    nC = case.mesh.nInternalCells
    c0 = np.sqrt(cf.GAMMA * cf.R * params['T0'])
    dt = 0.5 * (2 * np.pi / n) / (params['U0'] + c0)
    print('[bench] TGV %d^3 adjoint, %d cells, built in %.1fs' % (n, nC, time.time() - t0), file=sys.stderr,
          flush=True)
    res, tuned = {}, None
    for state in ('stage1', 'start'):
        for prec in (np.float32, np.float64):
            ctx = case.context(prec)
            ctx.objective_state = state
            u0 = torch.as_tensor(u0h.astype(prec), device='cuda').contiguous()
            if prec == np.float32:
                t_ = ctx.autotune(u0, reps=3)
                tuned = tuned or t_
            sweep = GraphSweep(ctx, dt, interval, obja=1.0, overlap=overlap)
            sweep.run(u0, interval)                    # capture + one segment
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            J, ua = sweep.run(u0, n_steps)
            t = time.perf_counter() - t1
            res[state, prec] = (float(J[0]), ua.double().cpu().numpy(), t)
            if state == 'stage1' and prec == np.float32:
                ovl = sweep.overlap
            print('[bench] TGV %d^3 adjoint, objective on %s, %s: %.3f s, J %.10e'
                  % (n, state, np.dtype(prec).name, t, J[0]), file=sys.stderr, flush=True)
            del sweep, ctx
    j32, g32, t32 = res['stage1', np.float32]
    j64, g64, t64 = res['stage1', np.float64]
    k32, h32, s32 = res['start', np.float32]
    k64, h64, s64 = res['start', np.float64]
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    return {'metric': 'cell-updates/s primal+adjoint, FP32 vs FP64, % HBM roofline, L2 err vs FP64',
            'value': nC * n_steps / t32, 'unit': 'cell-updates/s', 'n_gpus': 1, 'steps': n_steps, 'warmup': interval,
            'ms_per_step': t32 / n_steps * 1e3, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
            'dtype': 'f32', 'data': 'synthetic (Taylor-Green initial field)',
            'config': {'workload': 'TGV %d^3 periodic hex, random permutation + RCM, %d steps primal + adjoint of '
                                   'the summed kinetic energy, checkpoints every %d steps, HIP-graph segments'
                                   % (n, n_steps, interval), 'n_cells': nC, 'dt': dt},
            'block_sizes': {k: v[0] for k, v in tuned.items()}, 'overlapped_reverse': ovl,
            'value_fp64': nC * n_steps / t64, 'fp32_over_fp64': t64 / t32,
            'objective_fp32': j32, 'objective_fp64': j64, 'objective_rel_err_fp32_vs_fp64': abs(j32 - j64) / abs(j64),
            'grad_rel_l2_fp32_vs_fp64': {'rho': rel(g32[0:1], g64[0:1]), 'rhoU': rel(g32[1:4], g64[1:4]),
                                         'rhoE': rel(g32[4:5], g64[4:5])},
            'objective_on_start_state': {
                'value': nC * n_steps / s32, 'value_fp64': nC * n_steps / s64, 'fp32_over_fp64': s64 / s32,
                'objective_fp32': k32, 'objective_fp64': k64, 'objective_rel_err_fp32_vs_fp64': abs(k32 - k64) / abs(k64),
                'grad_rel_l2_fp32_vs_fp64': {'rho': rel(h32[0:1], h64[0:1]), 'rhoU': rel(h32[1:4], h64[1:4]),
                                             'rhoE': rel(h32[4:5], h64[4:5])}}}


def cylinder_bench(torch, n_steps=5000, interval=100, n_theta=512, n_r=256, ratio=1.02, overlap=None):
    """2-D cylinder wake: O-grid of n_theta x n_r x 1 hex cells, radial
    spacing growing by `ratio` from r = 0.5 D to 20 D, Mach 0.2, Re = 100,
    isothermal no-slip wall, CBC_UPT far field, Lax-Friedrichs at the
    boundaries.  n_steps fixed-dt steps at CFL 0.5 (dt = 2 CFL/dtc of the
    initial state, solver.py:359-366) primal plus the adjoint of the mean
    drag, uniform checkpoints every `interval` steps, each checkpoint
    segment one HIP graph; FP32 kernels with FP64 objective accumulation,
    and the same run in FP64 for the error and speed ratio."""
    from adfvm_amd import cases
    from adfvm_amd.graphs import GraphSweep
    t0 = time.time()
    s = cases.cylinder(n_theta, n_r, ratio=ratio)
    nC = s.mesh.nInternalCells
    print('[bench] cylinder %dx%d, %d cells, built in %.1fs' % (n_theta, n_r, nC, time.time() - t0),
          file=sys.stderr, flush=True)
    res, dt, tuned = {}, None, None
    for prec in (np.float32, np.float64):
        ctx = s.context(prec, boundaryRiemann='eulerLaxFriedrichs', objective=('drag', 'cylinder'))
        u0 = torch.as_tensor(s.state(), dtype=ctx.tdtype, device='cuda').contiguous()
        if dt is None:
            dtc = torch.zeros(1, dtype=torch.float64, device='cuda')
            ctx.primal_step(u0.clone(), 1e-12, dtc_out=dtc)
            dt = 2 * 0.5 / dtc.item()
        if prec == np.float32:
            tuned = ctx.autotune(u0, reps=3)
        sweep = GraphSweep(ctx, dt, interval, obja=1.0 / n_steps, overlap=overlap)
        sweep.run(u0, interval)                    # capture + one segment
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        J, ua = sweep.run(u0, n_steps)
        t = time.perf_counter() - t1
        res[prec] = (J[0] / n_steps, ua.double().cpu().numpy(), t)
        if prec == np.float32:
            ovl = sweep.overlap
        print('[bench] cylinder %s: %.3f s, mean drag %.10e' % (np.dtype(prec).name, t, J[0] / n_steps),
              file=sys.stderr, flush=True)
    d32, g32, t32 = res[np.float32]
    d64, g64, t64 = res[np.float64]
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    return {'metric': 'cell-updates/s primal+adjoint, FP32 vs FP64, % HBM roofline, L2 err vs FP64',
            'value': nC * n_steps / t32, 'unit': 'cell-updates/s', 'n_gpus': 1, 'steps': n_steps, 'warmup': interval,
            'ms_per_step': t32 / n_steps * 1e3, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
            'dtype': 'f32', 'data': 'synthetic (uniform free stream, impulsive start)',
            'config': {'workload': '2D cylinder wake, O-grid %d x %d x 1, radial stretch %g from r = 0.5 D to 20 D, '
                                   'Mach 0.2, Re 100, isothermal no-slip wall, CBC_UPT far field, LF boundary '
                                   'flux; %d steps primal + adjoint of the mean drag, checkpoints every %d steps, '
                                   'HIP-graph segments' % (n_theta, n_r, ratio, n_steps, interval),
                       'n_cells': nC, 'dt': dt},
            'block_sizes': {k: v[0] for k, v in tuned.items()}, 'overlapped_reverse': ovl,
            'value_fp64': nC * n_steps / t64, 'fp32_over_fp64': t64 / t32,
            'mean_drag_fp32': d32, 'mean_drag_fp64': d64, 'mean_drag_rel_err_fp32_vs_fp64': abs(d32 - d64) / abs(d64),
            'grad_rel_l2_fp32_vs_fp64': {'rho': rel(g32[0:1], g64[0:1]), 'rhoU': rel(g32[1:4], g64[1:4]),
                                         'rhoE': rel(g32[4:5], g64[4:5])},
            'finite': bool(np.isfinite(g32).all() and np.isfinite(g64).all())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--n', type=int, default=128)
    ap.add_argument('--l2-steps', type=int, default=100)
    ap.add_argument('--no-fp64', action='store_true')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-recompute', action='store_true', help='skip the tape-less variant')
    ap.add_argument('--ensemble', action='store_true',
                    help='instead: the batched ensemble (64 Taylor-Green instances of 32^3..96^3, 200 steps '
                         'primal+adjoint, checkpoints every 20, HIP-graph segments)')
    ap.add_argument('--adjoint64', action='store_true',
                    help='instead: the Taylor-Green 64^3 adjoint (400 steps primal + adjoint of the summed kinetic '
                         'energy, checkpoints every 20, HIP-graph segments), FP32 and FP64')
    ap.add_argument('--cylinder', action='store_true',
                    help='instead: the 2D cylinder wake (512 x 256 O-grid, 5000 steps primal + adjoint of the mean '
                         'drag, checkpoints every 100, HIP-graph segments), FP32 and FP64')
    ap.add_argument('--no-overlap', action='store_true',
                    help='with --ensemble, --adjoint64, --cylinder: reverse segments in sequence (no re-run on a '
                         'second stream alongside the adjoint steps)')
    ap.add_argument('--calibrate-copy', type=int, default=0, metavar='MIB',
                    help='only run a device-to-device copy of MIB MiB five times (counter calibration)')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py runs on one GPU only: the library has no multi-GPU path (--gpus %d requested)' % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import torch
    from tests import casefactory as cf
    assert torch.cuda.is_available(), 'bench.py needs the GPU'
    torch.cuda.set_device(0)
    overlap = False if args.no_overlap else None
    if args.ensemble:
        print(json.dumps(ensemble_bench(torch, overlap=overlap)))
        return
    if args.cylinder:
        print(json.dumps(cylinder_bench(torch, overlap=overlap)))
        return
    if args.adjoint64:
        print(json.dumps(adjoint64_bench(torch, overlap=overlap)))
        return
    if args.calibrate_copy:
        n = args.calibrate_copy * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device='cuda')
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({'calibrate_copy_bytes_read': n * 4, 'bytes_written': n * 4, 'launches': 5}))
        return

    t0 = time.time()
    case, params = cf.tgv(args.n, seed=0)
    m = case.mesh
    nC, nF = m.nInternalCells, m.nFaces
    nG = m.nFaces - m.nInternalFaces
    own = np.asarray(m.owner[:m.nInternalFaces])
    # the face adjoint sums owner sides per cell when internal faces are in owner order
    nO = len(np.unique(own)) if np.all(np.diff(own) >= 0) else None
    u0 = cf.tgv_state(case, params)
    dx = 2 * np.pi / args.n
    c0 = np.sqrt(cf.GAMMA * cf.R * params['T0'])
    dt = 0.5 * dx / (params['U0'] + c0)       # fixed dt at CFL 0.5 (convective + acoustic)
    print('[bench] mesh %d^3: %d cells, %d faces built in %.1fs; dt=%.3e' % (args.n, nC, nF, time.time() - t0, dt),
          file=sys.stderr, flush=True)

    r32 = run_precision(case, u0, np.float32, args.steps, args.warmup, dt, torch, profile=True)
    print('[bench] fp32 P+A %.3fs primal %.3fs' % (r32['t_pa'], r32['t_p']), file=sys.stderr, flush=True)
    r32r = None
    if not args.no_recompute:
        r32r = run_precision(case, u0, np.float32, args.steps, args.warmup, dt, torch, profile=False,
                             recompute=True)
        print('[bench] fp32 P+A recompute %.3fs' % r32r['t_pa'], file=sys.stderr, flush=True)
    r64 = None
    if not args.no_fp64:
        r64 = run_precision(case, u0, np.float64, args.steps, args.warmup, dt, torch, profile=True)
        print('[bench] fp64 P+A %.3fs primal %.3fs' % (r64['t_pa'], r64['t_p']), file=sys.stderr, flush=True)
    l2 = l2_error(case, u0, args.l2_steps, dt, torch) if args.l2_steps > 0 else None
    print('[bench] L2 %s' % l2, file=sys.stderr, flush=True)

    K = args.steps
    value = nC * K / r32['t_pa']
    # dominant kernel roofline (FP32)
    prof = r32['prof']
    dom = max(prof, key=lambda k: prof[k][0])
    ms_tot, launches = prof[dom]
    avg_s = ms_tot / launches * 1e-3
    kb = kernel_bytes(dom, 4, nC, nF, nG, nO)
    achieved = kb / avg_s / 1e9 if kb else None
    traffic = None
    tf = os.path.join(ROOT, 'profiles', 'traffic.json')
    if os.path.exists(tf):
        traffic = json.load(open(tf)).get(dom)
    triad = triad_bandwidth(torch)
    bpc_p32, bpc_pa32 = step_bytes_model(4, nC, nF, nG)
    bpc_p64, bpc_pa64 = step_bytes_model(8, nC, nF, nG)
    out = {
        'metric': 'cell-updates/s primal+adjoint, FP32 vs FP64, % HBM roofline, L2 err vs FP64',
        'value': value,
        'unit': 'cell-updates/s',
        'n_gpus': 1,
        'steps': K,
        'warmup': args.warmup,
        'ms_per_step': r32['t_pa'] / K * 1e3,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32',
        'data': 'synthetic (Taylor-Green initial field)',
        'config': {'workload': 'TGV %d^3 periodic hex, random permutation + RCM, primal+adjoint, SSP-RK3, Roe, '
                               'fixed dt, start states and step tapes (stage states + gradients) of all K steps '
                               'kept in HBM' % args.n,
                   'n_cells': nC, 'n_faces': nF, 'n_owner_records': nO, 'dt': dt,
                   'objective': 'kinetic energy on the stage-1 state'},
        'roofline': {'bound': 'hbm', 'kernel': dom, 'achieved': achieved, 'peak': HBM_PEAK / 1e9, 'unit': 'GB/s',
                     'frac': (achieved / (HBM_PEAK / 1e9)) if achieved else None, 'traffic': traffic,
                     'algorithmic_bytes_per_launch': kb, 'avg_launch_ms': avg_s * 1e3,
                     'triad_GBs': triad, 'frac_of_triad': (achieved / triad) if achieved else None},
        'hbm_frac_step_model': value * bpc_pa32 / HBM_PEAK,
        'triad_frac_step_model': value * bpc_pa32 / (triad * 1e9),
        'bytes_per_cell_update_model': {'fp32_primal': bpc_p32, 'fp32_pa': bpc_pa32, 'fp64_pa': bpc_pa64},
        'primal_cell_updates_per_s_fp32': nC * K / r32['t_p'],
        'value_fp32_recompute_stages': nC * K / r32r['t_pa'] if r32r else None,
        'kernel_ms_fp32': {k: round(v[0] / K, 4) for k, v in prof.items()},
        'kernel_roofline_fp32': kernel_rooflines(prof, 4, nC, nF, nG, nO, triad),
        'block_sizes_fp32': r32['tuned'],
        'streaming_fp32': r32['streaming'],
        'finite': r32['finite'],
    }
    if r64 is not None:
        v64 = nC * K / r64['t_pa']
        out.update({'value_fp64': v64, 'fp32_over_fp64': value / v64,
                    'hbm_frac_step_model_fp64': v64 * bpc_pa64 / HBM_PEAK,
                    'primal_cell_updates_per_s_fp64': nC * K / r64['t_p'],
                    'kernel_ms_fp64': {k: round(v[0] / K, 4) for k, v in r64['prof'].items()}})
    if l2 is not None:
        out.update({'l2_rel_err_fp32_vs_fp64': l2, 'l2_steps': args.l2_steps})
    if not args.no_cpu:
        out['cpu_baseline'] = cpu_baseline(m, case.mu0, u0, dt)
        print('[bench] cpu baseline %s' % out['cpu_baseline'], file=sys.stderr, flush=True)
    print(json.dumps(out))


if __name__ == '__main__':
    main()
