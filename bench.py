#!/usr/bin/env python3
"""Benchmark: Q(f,f) DOF-evals/s on one MI355X, mixed (FP32 transforms, FP64
accumulation and FP64 conservation projection) vs plain FP64, at the workload

    vss-gll, 32^3 velocity grid on [-6,6]^3, M = 12 (full sphere),
    Nrho = 16, omega = 0.74, alpha = 1.4, Mpre = 32, 4096 DOFs of seeded
    two-Maxwellian f (single species).

A step is one evaluation of Q for all DOFs (inputs resident in HBM).  Prints
one JSON line.  ``--gpus N`` with N > 1 exits: this build runs on one GPU.
Other workloads (``--workload``): vhs16 (plain FP64, 16^3, 64 DOFs), vss32x2
(two species, four pairs), rhs24 (SSP-RK2 step with FR advection) and fno16 /
fno32
(the learned stand-in against the spectral operator on 16^3, 8192 DOFs).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {'fp64': 78.6, 'fp32': 157.3}     # MI355X vector TFLOP/s (published)
HBM_PEAK = 8000.0                          # GB/s
# HBM bytes per DOF of the dominant (product) kernel at the 32^3 workload
# (N = 32, M = 12, Nrho = 16, Mpre = 32, 4096 DOFs with 2048-DOF scratch and
# four DOF slots per accumulation workgroup: batches of 1440 DOFs; the
# per-DOF bytes do not depend on the batch), from rocprofv3
# --pmc FETCH_SIZE and WRITE_SIZE runs of this script, scaled by what the same
# counters read for a 1 GiB device copy (FETCH x2.0, WRITE x1.0).  The bench
# line carries this figure with its source, and only for that shape.
K_PROD_TRAFFIC = {
    'source': 'profiles/k_prod_vss32_traffic.txt',
    'shape': {'Nv': 32, 'M': 12, 'Nrho': 16, 'Mpre': 32},
    'per_dof': {'mixed': 17.66e6, 'fp64': 55.86e6},
    # the accumulation kernel (HBM-bound), same runs
    'accum_per_dof': {'mixed': 14.41e6, 'fp64': 53.29e6},
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--workload', default='vss32',
                    choices=['vss32', 'vhs16', 'vss32x2', 'rhs24', 'fno16', 'fno32'],
                    help='vss32: the Q(f,f) headline; vhs16: plain FP64 16^3, 64 DOFs; '
                         'vss32x2: two species, four pairs; rhs24: full SSP-RK2 step (FR advection + Q); '
                         'fno16 / fno32: the learned stand-in on the 16^3 / 32^3 grid, 8192 DOFs')
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--ndof', type=int, default=4096)
    ap.add_argument('--Nv', type=int, default=32)
    ap.add_argument('--Nrho', type=int, default=16)
    ap.add_argument('--M', type=int, default=12)
    ap.add_argument('--Mpre', type=int, default=32)
    ap.add_argument('--seed', type=int, default=2024)
    ap.add_argument('--cpu-dofs-per-proc', type=int, default=8)
    ap.add_argument('--cpu-procs', type=int, default=0, help='0 = min(cpus, 16)')
    ap.add_argument('--cpu-runs', type=int, default=5, help='CPU baseline: median of this many runs')
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--autotune', type=int, default=1,
                    help='1: run the schedule autotuner (untimed) before the warmup steps')
    ap.add_argument('--no-reference-proj', action='store_true',
                    help='skip the untimed FP64-with-projection reference evaluation')
    ap.add_argument('--modes', default='fp64,mixed,mixedproj')
    ap.add_argument('--sm-opt', action='append', default=[], metavar='KEY=VALUE',
                    help='extra [scattering-model] key, e.g. pair-pack=0 (repeatable)')
    ap.add_argument('--fno-weights', default='', help='fno16/32: trained weights (.npz); default: train first')
    ap.add_argument('--fno-save', default='', help='fno16/32: write the trained weights here')
    ap.add_argument('--fno-pairs', type=int, default=0, help='training pairs; 0 = 200000 (fno16), 50000 (fno32)')
    ap.add_argument('--fno-epochs', type=int, default=20)
    ap.add_argument('--fno-train-seconds', type=float, default=150.0)
    ap.add_argument('--fno-tol', type=float, default=1e-4)
    ap.add_argument('--fno-train-acts', default='fp32', choices=['fp32', 'fp16', 'bf16'],
                    help='fno16/32: activation storage while training')
    ap.add_argument('--fno-dofs', type=int, default=8192)
    ap.add_argument('--fno-cons-weight', type=float, default=10.0,
                    help='fno16/32: weight of the conservation penalty in training')
    ap.add_argument('--fno-batch', type=int, default=64, help='training batch (DOFs)')
    ap.add_argument('--library-steps', type=int, default=2,
                    help='timed steps of the library-FFT comparison (the same plan with every 3D '
                         'transform in hipFFT); 0 skips it')
    ap.add_argument('--calib-copy-mib', type=int, default=0,
                    help='run one device-to-device copy of this size first (counter calibration)')
    return ap.parse_args()


def sm_opts(a, sm):
    """[scattering-model] section sm with the --sm-opt overrides applied."""
    for kv in a.sm_opt:
        k, v = kv.split('=', 1)
        sm[k.strip()] = v.strip()
    return sm


def vss_cfg(a, precision, project=0, extra=None):
    from frfs_amd.inifile import Inifile
    return Inifile.from_dict({
        'non-dim': dict(T0=273.0, H0=1e-3, rho0=0.00011140244750324529, molarMass0=39.96e-3),
        'constants': {'Nv': a.Nv, 'NvBatchSize': a.Nv ** 3, 'Nrho': a.Nrho},
        'velocity-mesh': {'dev': 6.0, 'cmax': 0, 'Tmax': 273, 'molarMass1': 39.96e-3},
        'spherical-design-rule': {'ssrule': 'womersley', 'M': a.M},
        'scattering-model': sm_opts(a, {'type': 'vss-gll', 'ssrulepre': 'womersley', 'Mpre': a.Mpre,
                                        'alpha00': 1.4, 'omega00': 0.74, 'dRef00': 4.17e-10,
                                        'Tref00': 273, 'precision': precision,
                                        'conservation-projection': project, **(extra or {})}),
    })


# --------------------------------------------------------------------------
# CPU baseline: the oracle (NumPy complex128) on the host cores, run before
# the GPU is initialised (forked workers, no exec).
# --------------------------------------------------------------------------
_ORC = {}


def _cpu_worker(dofs):
    from oracle import fastspectral as fsp
    vm, prm, tab, fpar = _ORC['vm'], _ORC['prm'], _ORC['tab'], _ORC['fpar']
    out = []
    t0 = time.perf_counter()
    for d in dofs:
        f = np.zeros(vm['cv'].shape[1])
        for k in range(2):
            nk, ux, uy, uz, Tk = fpar[d, k]
            f += fsp.maxwellian(vm['cv'], nk, [ux, uy, uz], Tk)
        out.append(fsp.vss_fs([f], vm, tab, prm)[0])
    return dofs, out, time.perf_counter() - t0


def cpu_baseline(a, fpar):
    """The oracle on the host cores: ``procs`` forked processes (one thread
    each) evaluate ``cpu_dofs_per_proc`` DOFs each; the rate of a run is its
    DOF count over its wall time and the baseline is the median of
    ``cpu_runs`` runs over disjoint DOF samples of the same workload."""
    from multiprocessing import get_context
    from oracle import fastspectral as fsp
    os.environ.setdefault('OMP_NUM_THREADS', '1')
    vm = fsp.velocity_mesh(a.Nv, a.Nrho, a.M, dev=6.0, masses=[39.96e-3])
    prm = fsp.vss_parameters(vm, {'00': 1.4}, {'00': 0.74}, {'00': 4.17e-10}, {'00': 273.0},
                             'womersley', a.Mpre)
    t0 = time.perf_counter()
    tab = fsp.vss_tables(vm, prm)
    t_tab = time.perf_counter() - t0
    _ORC.update(vm=vm, prm=prm, tab=tab, fpar=fpar)
    ncpu = os.cpu_count() or 1
    procs = a.cpu_procs or min(ncpu, 16)
    per = a.cpu_dofs_per_proc
    runs = max(1, a.cpu_runs)
    total = procs * per
    rng = np.random.default_rng(7)
    sample = rng.choice(a.ndof, size=min(total * runs, a.ndof), replace=False)
    ctx = get_context('fork')
    # warm-up (one DOF, single process): first-touch of FFT plans
    _cpu_worker([int(sample[0])])
    rates, walls, ref, cpu_secs, ndone = [], [], {}, 0.0, 0
    with ctx.Pool(procs) as pool:
        for r in range(runs):
            part = np.sort(sample[r * total:(r + 1) * total]) if len(sample) >= (r + 1) * total else \
                np.sort(sample[:total])
            parts = [part[i::procs].tolist() for i in range(procs)]
            t0 = time.perf_counter()
            res = pool.map(_cpu_worker, parts)
            wall = time.perf_counter() - t0
            rates.append(len(part) / wall)
            walls.append(wall)
            for dofs, qs, secs in res:
                cpu_secs += secs
                ndone += len(dofs)
                for d, q in zip(dofs, qs):
                    ref[d] = q
    return dict(value=float(np.median(rates)), unit='DOF-evals/s', cores=procs, kind='port',
                sample=f'median of {runs} runs; each run {total} of the {a.ndof} DOFs of the same workload '
                       f'(disjoint samples), NumPy complex128 oracle, {procs} processes x {per} DOFs, '
                       f'1 thread each, host has {ncpu} CPUs; tables precomputed once ({t_tab:.1f} s, untimed)',
                run_rates=rates, single_core_dof_evals_per_s=ndone / cpu_secs,
                wall_s=walls), ref


def bench_multi(a):
    """Two species (molar masses 1:2), all four pairs (p, q) in one batched
    launch: vss-gll, 32^3 grid on [-6,6]^3, M = 12, Nrho = 16, Mpre = 32;
    omega 0.74 (00) and 0.81 (11), 0.775 across; dRef 4.17e-10 (00),
    1.3 x that (11), the mean across; alpha = 1.4.  2048 DOFs per species:
    the seeded two-Maxwellian mixture of the single-species workload, with
    species 1 taking a share x ~ U(0.2, 0.8) of the number density.  One
    DOF-eval is Q_p for one species at one DOF, all its pairs included."""
    import torch
    from frfs_amd.autotune import ONE_STREAM
    from frfs_amd.backend import HIPBackend
    from frfs_amd.inifile import Inifile
    from frfs_amd.inputs import from_aosoa, mixture_device, mixture_params, to_aosoa
    from frfs_amd.scattering import DGFSBiVSSGLLScatteringModel
    from frfs_amd.velocitymesh import DGFSBiVelocityMesh
    be = HIPBackend()
    ndof, N = a.ndof // 2, a.Nv
    fpar = mixture_params(ndof, a.seed)
    share = np.random.default_rng(a.seed + 1).uniform(0.2, 0.8, ndof)
    omega = {'00': 0.74, '11': 0.81, '01': 0.775, '10': 0.775}
    dref = {'00': 4.17e-10, '11': 1.3 * 4.17e-10, '01': 1.15 * 4.17e-10, '10': 1.15 * 4.17e-10}
    res, qs = {}, {}
    modes = [m for m in a.modes.split(',') if m in ('fp64', 'mixed')]
    runs = [(m, m == 'mixed', True) for m in modes]
    if 'mixed' in modes:
        runs.append(('fp64proj', True, False))     # accuracy reference only, untimed
    for mode, proj, timed in runs:
        sm = {'type': 'vss-gll', 'ssrulepre': 'womersley', 'Mpre': a.Mpre,
              'precision': 'mixed' if mode == 'mixed' else 'double',
              'conservation-projection': 1 if proj else 0}
        for pq in omega:
            sm.update({'alpha' + pq: 1.4, 'omega' + pq: omega[pq], 'dRef' + pq: dref[pq], 'Tref' + pq: 273})
        cfg = Inifile.from_dict({
            'non-dim': dict(T0=273.0, H0=1e-3, rho0=0.00011140244750324529, molarMass0=39.96e-3),
            'constants': {'Nv': N, 'NvBatchSize': N ** 3, 'Nrho': a.Nrho},
            'velocity-mesh': {'dev': 6.0, 'cmax': 0, 'Tmax': 273, 'molarMass1': 39.96e-3,
                              'molarMass2': 2 * 39.96e-3},
            'spherical-design-rule': {'ssrule': 'womersley', 'M': a.M},
            'scattering-model': sm_opts(a, sm)})
        vm = DGFSBiVelocityMesh(be, cfg, 2)
        model = DGFSBiVSSGLLScatteringModel(be, cfg, vm)
        n = vm.vsize()
        fin = [be.matrix((1, n, ndof)) for _ in range(2)]
        qout = [be.matrix((1, n, ndof)) for _ in range(2)]
        sh = torch.as_tensor(share, device='cuda')[:, None]
        for s, m in enumerate(vm.masses()):
            f = mixture_device(vm.cv(), fpar, mass=m) * (sh if s == 1 else 1 - sh)
            to_aosoa(f, fin[s])
        torch.cuda.synchronize()
        # the schedule autotuner (untimed) before the warmup steps
        sched = model.autotune(*fin, *qout, candidates=ONE_STREAM)['choice'] if (timed and a.autotune) else None
        for _ in range(a.warmup if timed else 1):
            model.fs_batch(*fin, *qout)
        torch.cuda.synchronize()
        if timed:
            t0 = time.perf_counter()
            for _ in range(a.steps):
                model.fs_batch(*fin, *qout)
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t0) / a.steps
            res[mode] = dict(ms=dt * 1e3, rate=2 * ndof / dt, info=model.plan_info(), schedule=sched,
                             chunk=model._h.chunk_dofs(), scratch=model._h.scratch_bytes())
        qs[mode] = [from_aosoa(q).cpu().numpy() for q in qout]
        del model, fin, qout
        torch.cuda.empty_cache()
    head = 'mixed' if 'mixed' in res else list(res)[0]

    def rel(x, y):
        return [float(np.linalg.norm(x[s] - y[s]) / np.linalg.norm(y[s])) for s in range(2)]
    err = None
    if 'mixed' in qs and 'fp64' in qs:
        err = {'mixed_proj_vs_fp64_proj': rel(qs['mixed'], qs['fp64proj']),
               'mixed_proj_vs_fp64': rel(qs['mixed'], qs['fp64']),
               'fp64_proj_vs_fp64 (projection correction)': rel(qs['fp64proj'], qs['fp64'])}
    out = {'metric': 'two-species Q_p DOF-evals/s (4 pairs in one launch), 32^3 grid',
           'value': res[head]['rate'], 'unit': 'DOF-evals/s', 'n_gpus': 1, 'steps': a.steps,
           'warmup': a.warmup, 'ms_per_step': res[head]['ms'], 'higher_is_better': True,
           'scaling': 'weak', 'vs_baseline': None,
           'dtype': 'f32+f64acc' if head == 'mixed' else 'f64', 'data': 'synthetic',
           'config': {'workload': f'vss-gll two species, masses 1:2, {N}^3 grid on [-6,6]^3, M={a.M}, '
                                  f'Nrho={a.Nrho}, Mpre={a.Mpre}, {ndof} DOFs per species, species-1 '
                                  f'density share U(0.2,0.8)', 'mode': head, 'ndof_per_species': ndof,
                      'schedule': res[head]['schedule'], 'chunk_dofs': res[head]['chunk'],
                      'scratch_GiB': res[head]['scratch'] / 2 ** 30},
           'plan': res[head]['info'],
           'fp64': ({'value': res['fp64']['rate'], 'ms_per_step': res['fp64']['ms'],
                     'schedule': res['fp64']['schedule']} if 'fp64' in res else None),
           'speedup_mixed_vs_fp64': (res['mixed']['rate'] / res['fp64']['rate']) if len(res) == 2 else None,
           'rel_l2_per_species': err}
    print(json.dumps(out))


def bench_vhs16(a):
    """The CPU-runnable configuration: plain FP64 vhs-gll (omega = 0.5),
    16^3 grid on [-5,5]^3, M = 6 half-sphere directions, Nrho = 8, 64 DOFs of
    the seeded two-Maxwellian mixture, checked against the oracle on every
    DOF (the oracle also gives the CPU rate on the same DOFs)."""
    import torch
    from frfs_amd.backend import HIPBackend
    from frfs_amd.inifile import Inifile
    from frfs_amd.inputs import from_aosoa, mixture_device, mixture_params, to_aosoa
    from frfs_amd.scattering import DGFSVHSGLLScatteringModel
    from frfs_amd.velocitymesh import DGFSVelocityMesh
    from oracle import fastspectral as fsp
    ndof, N, Nrho, M = 64, 16, 8, 6
    fpar = mixture_params(ndof, a.seed)
    # oracle first (CPU, before any GPU work)
    vm_o = fsp.velocity_mesh(N, Nrho, M, dev=5.0)
    prm = fsp.vhs_parameters(vm_o, 0.5, 4.17e-10, 273.0)
    tab = fsp.vhs_tables(vm_o, prm['gamma'])
    fo = np.zeros((ndof, N ** 3))
    for d in range(ndof):
        for k in range(2):
            nk, ux, uy, uz, Tk = fpar[d, k]
            fo[d] += fsp.maxwellian(vm_o['cv'], nk, [ux, uy, uz], Tk)
    t0 = time.perf_counter()
    ref = np.stack([fsp.vhs_fs(fo[d], vm_o, tab, prm['prefac']) for d in range(ndof)])
    cpu_s = time.perf_counter() - t0
    be = HIPBackend()
    cfg = Inifile.from_dict({
        'non-dim': dict(T0=273.0, H0=1e-3, rho0=0.00011140244750324529, molarMass0=39.96e-3),
        'constants': {'Nv': N, 'NvBatchSize': N ** 3, 'Nrho': Nrho},
        'velocity-mesh': {'dev': 5.0, 'cmax': 0, 'Tmax': 273},
        'spherical-design-rule': {'ssrule': 'womersley', 'M': M},
        'scattering-model': {'type': 'vhs-gll', 'omega': 0.5, 'dRef': 4.17e-10, 'Tref': 273,
                             'precision': 'double'}})
    vm = DGFSVelocityMesh(be, cfg)
    sm = DGFSVHSGLLScatteringModel(be, cfg, vm)
    n = vm.vsize()
    fin, qout = be.matrix((1, n, ndof)), be.matrix((1, n, ndof))
    to_aosoa(mixture_device(vm.cv(), fpar), fin)
    torch.cuda.synchronize()
    for _ in range(a.warmup):
        sm.fs_batch(fin, qout)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        sm.fs_batch(fin, qout)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / a.steps
    q = from_aosoa(qout).cpu().numpy()
    err = float(np.max(np.linalg.norm(q - ref, axis=1) / np.linalg.norm(ref, axis=1)))
    print(json.dumps({
        'metric': 'plain FP64 Q(f,f) DOF-evals/s, vhs-gll 16^3, 64 DOFs', 'value': ndof / dt,
        'unit': 'DOF-evals/s', 'n_gpus': 1, 'steps': a.steps, 'warmup': a.warmup, 'ms_per_step': dt * 1e3,
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64', 'data': 'synthetic',
        'config': {'workload': 'vhs-gll omega=0.5, 16^3 grid on [-5,5]^3, M=6 half sphere, Nrho=8, 64 DOFs of '
                               'seeded two-Maxwellian f'},
        'max_dof_rel_l2_vs_oracle': err, 'plan': sm.plan_info(),
        'cpu_baseline': {'value': ndof / cpu_s, 'unit': 'DOF-evals/s', 'cores': 1, 'kind': 'port',
                         'sample': f'all {ndof} DOFs, NumPy complex128 oracle, one process'}}))


def fno_flop_per_dof(N):
    """Network forward per DOF (N^3 grid, width 32, 8^3 modes, 4 layers): per
    layer the truncated forward transform (32 N^3 + 512 N^2 + 4096 N flop per
    channel: 8 modes of N-point lines along z, y, x; 0.33 MFLOP at 16^3, 1.70
    MFLOP at 32^3), the per-mode channel mixing (512 x 32 x 32 complex
    multiply-adds, 4.19 MFLOP), the inverse transform (as the forward) and the
    pointwise 32 x 32 layer (2048 N^3); lift and projection 64 N^3 each."""
    dft = 32 * N ** 3 + 512 * N ** 2 + 4096 * N
    return 4 * (32 * dft + 4194304 + 32 * dft + 2048 * N ** 3) + 2 * 64 * N ** 3


def bench_fno(a):
    """The learned stand-in, 4 Fourier layers of width 32 with 8^3 modes,
    trained on the GPU (unless --fno-weights) on --fno-pairs pairs of the
    conservation-projected FP64 spectral operator, then timed on 8192
    held-out DOFs of seeded sheared mixtures: the whole plugin call (gather,
    network, conservation check and projection, spectral fallback on the
    flagged DOFs) against the spectral operator on the same DOFs in plain
    FP64 and in mixed precision.
      fno16: vhs-gll-fno on the 16^3 grid (vhs-gll omega = 0.5, M = 6, Nrho = 8);
      fno32: vss-gll-fno on the 32^3 grid (vss-gll omega = 0.74, alpha = 1.4,
             M = 12, Nrho = 16, Mpre = 32, [-6, 6]^3)."""
    import torch
    from frfs_amd import train_surrogate as ts
    from frfs_amd.backend import HIPBackend
    from frfs_amd.inputs import from_aosoa, shear_mixture_device, shear_mixture_params, to_aosoa
    be = HIPBackend()
    grid = 'vss32' if a.workload == 'fno32' else 'vhs16'
    ndof = a.fno_dofs
    if a.fno_weights:
        cfg = ts.config(tol=a.fno_tol, weights=a.fno_weights, grid=grid)
        sm = ts.model(be, cfg)
        vm = sm.vm
        rep = {'weights': os.path.basename(a.fno_weights)}
    else:
        pairs = a.fno_pairs or (50000 if grid == 'vss32' else 200000)
        sm, rep = ts.train(be, pairs=pairs, epochs=a.fno_epochs, time_limit=a.fno_train_seconds,
                           seed=a.seed, train_acts=a.fno_train_acts, grid=grid, batch=a.fno_batch,
                           cons_weight=a.fno_cons_weight,
                           log=lambda m: print(m, file=sys.stderr, flush=True))
        sm.tol = a.fno_tol
        vm = sm.vm
        if a.fno_save:
            sm.net.save(a.fno_save)
    NV = vm.Nv()
    n = vm.vsize()
    f = shear_mixture_device(vm.cv(), shear_mixture_params(ndof, a.seed + 1), device=be.device)
    fin = be.matrix((1, n, ndof))
    to_aosoa(f, fin)
    del f
    qref, qs = be.matrix(fin.ioshape), be.matrix(fin.ioshape)

    def timed(fn):
        for _ in range(a.warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / a.steps

    # the teacher (FP64 with the projection) gives the reference Q; the plain
    # FP64 operator (no projection) gives the speed reference
    sm.fs_spectral(fin, qref)

    def spectral(precision, project):
        c = ts.config(grid=grid, project=project)
        c.set('scattering-model', 'type', 'vss-gll' if grid == 'vss32' else 'vhs-gll')
        c.set('scattering-model', 'precision', precision)
        from frfs_amd.scattering import DGFSBiVSSGLLScatteringModel, DGFSVHSGLLScatteringModel
        from frfs_amd.velocitymesh import DGFSBiVelocityMesh, DGFSVelocityMesh
        if grid == 'vss32':
            return DGFSBiVSSGLLScatteringModel(be, c, DGFSBiVelocityMesh(be, c, 1))
        return DGFSVHSGLLScatteringModel(be, c, DGFSVelocityMesh(be, c))
    smp = spectral('double', 0)
    qplain = be.matrix(fin.ioshape)
    dt_fp64 = timed(lambda: smp.fs_batch(fin, qplain))
    del smp
    smm = spectral('mixed', 1)
    qm = be.matrix(fin.ioshape)
    dt_mixed = timed(lambda: smm.fs_batch(fin, qm))
    del smm
    torch.cuda.empty_cache()
    stats = {}

    def run():
        stats.update(sm.fs_batch(fin, qs))
    dt = timed(run)
    # the same call with a defect tolerance the operator itself meets on most
    # DOFs, so that the network serves them (the sampled check still applies)
    loose, tol0, stats_l = 3e-2, sm.tol, {}
    sm.tol = loose
    ql = be.matrix(fin.ioshape)

    def run_loose():
        stats_l.update(sm.fs_batch(fin, ql))
    dt_l = timed(run_loose)
    sm.tol = tol0
    # the network forward alone, HIP events on the stream it runs on
    g = torch.empty((ndof, n), dtype=torch.float32, device=be.device)
    out = torch.empty_like(g)
    from frfs_amd import _lib
    from frfs_amd.surrogate import F_REF
    lib, bst = _lib.load(), be.stream_handle()
    rho = torch.empty(ndof, dtype=torch.float64, device=be.device)
    _lib.check(lib.fq_sur_gather(NV, vm.cw(), F_REF, 1, fin._as_parameter_, 1, fin.leaddim, ndof, 0, ndof, None,
                                 rho.data_ptr(), g.data_ptr(), bst))
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    sm.net.predict(g, out=out)
    e0.record(st)
    for _ in range(a.steps):
        sm.net.predict(g, out=out)
    e1.record(st)
    torch.cuda.synchronize()
    fwd_ms = e0.elapsed_time(e1) / a.steps
    # the same with 16-bit activation storage
    acts16 = {}
    for kind in ('fp16', 'bf16'):
        out16 = torch.empty_like(g)
        sm.net.predict(g, out=out16, acts=kind)
        e0.record(st)
        for _ in range(a.steps):
            sm.net.predict(g, out=out16, acts=kind)
        e1.record(st)
        torch.cuda.synchronize()
        ms16 = e0.elapsed_time(e1) / a.steps
        r16 = ((out16 - out).pow(2).sum(1) / out.pow(2).sum(1)).sqrt().cpu().numpy()
        acts16[kind] = {'network_forward_ms': ms16, 'speedup_vs_fp64': dt_fp64 / (ms16 * 1e-3),
                        'rel_l2_vs_fp32_activations': {'median': float(np.median(r16)), 'max': float(r16.max())}}
    qa, qb = from_aosoa(qs), from_aosoa(qref)
    r = ((qa - qb).pow(2).sum(1) / qb.pow(2).sum(1)).sqrt().cpu().numpy()
    fl = stats.get('defect', np.zeros(ndof)) > sm.tol
    net_r = r[~fl] if (~fl).any() else np.zeros(1)
    rm = from_aosoa(qm)
    rmix = float(((rm - qb).pow(2).sum(1) / qb.pow(2).sum(1)).sqrt().median())
    # untimed: the network (conservation-projected) on every DOF, no fallback
    # and no sampled check
    tol, every = sm.tol, sm.check_every
    sm.tol, sm.check_every = 2.0, 0
    qn = be.matrix(fin.ioshape)
    sm.fs_batch(fin, qn)
    sm.tol, sm.check_every = tol, every
    rn = ((from_aosoa(qn) - qb).pow(2).sum(1) / qb.pow(2).sum(1)).sqrt().cpu().numpy()
    dfc = stats['defect']
    # the teacher's own conservation defect, and its projection, through the
    # same check kernel: Q_fp64 / (rho^2 q_ref) -> defect and projected Q
    gq = torch.empty((ndof, n), dtype=torch.float32, device=be.device)
    _lib.check(lib.fq_sur_gather(NV, vm.cw(), sm.net.q_ref, 2, qref._as_parameter_, 1, qref.leaddim, ndof, 0, ndof,
                                 rho.data_ptr(), None, gq.data_ptr(), bst))
    qrp = be.matrix(fin.ioshape)
    tdef = torch.empty(ndof, dtype=torch.float64, device=be.device)
    tflag = torch.empty(ndof, dtype=torch.int32, device=be.device)
    _lib.check(lib.fq_sur_finish(NV, vm.L(), vm.cw(), sm.net.q_ref, 2.0, gq.data_ptr(), rho.data_ptr(),
                                 sm._ginv.data_ptr(), sm._gdiag.data_ptr(), qrp._as_parameter_, 1, qrp.leaddim, ndof,
                                 0, ndof, tdef.data_ptr(), tflag.data_ptr(), bst))
    qbp = from_aosoa(qrp)
    rnp = ((from_aosoa(qn) - qbp).pow(2).sum(1) / qbp.pow(2).sum(1)).sqrt().cpu().numpy()
    tdef = tdef.cpu().numpy()
    cpu = None
    if not a.no_cpu_baseline:
        from oracle import surrogate as osur
        th = sm.net.theta.double().cpu().numpy()
        gs = g[:4].double().cpu().numpy()
        t0 = time.perf_counter()
        osur.forward(th, gs, NV, 32, 8, sm.net.layers)
        cs = time.perf_counter() - t0
        cpu = {'value': 4 / cs, 'unit': 'DOF-evals/s', 'cores': 1, 'kind': 'port',
               'sample': '4 DOFs of the same workload, network forward only, NumPy float64 oracle, one process'}
    flop = fno_flop_per_dof(NV)
    tflops = flop * ndof / (fwd_ms * 1e-3) / 1e12
    rplain = ((from_aosoa(qs) - from_aosoa(qplain)).pow(2).sum(1)
              / from_aosoa(qplain).pow(2).sum(1)).sqrt().cpu().numpy()
    teacher = ('vss-gll omega=0.74 alpha=1.4, 32^3 grid on [-6,6]^3, M=12, Nrho=16, Mpre=32' if grid == 'vss32'
               else 'vhs-gll omega=0.5, 16^3 grid on [-5,5]^3, M=6, Nrho=8')
    print(json.dumps({
        'metric': 'learned stand-in Q(f,f) DOF-evals/s (with conservation check and spectral fallback), '
                  f'{sm.scattering_model} {NV}^3, {ndof} DOFs', 'value': ndof / dt,
        'unit': 'DOF-evals/s', 'n_gpus': 1, 'steps': a.steps, 'warmup': a.warmup, 'ms_per_step': dt * 1e3,
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32', 'data': 'synthetic',
        'config': {'workload': 'Fourier-layer stand-in (4 layers, width 32, 8^3 modes) for ' + teacher +
                               f'; {ndof} DOFs of seeded mixtures of up to 3 sheared Maxwellians; trained on '
                               'the conservation-projected FP64 operator; fallback tolerance ' + repr(sm.tol)},
        'rel_l2_vs_plain_fp64': {'median': float(np.median(rplain)), 'p90': float(np.quantile(rplain, 0.9))},
        'spectral_fp64_dof_evals_per_s': ndof / dt_fp64, 'spectral_mixed_dof_evals_per_s': ndof / dt_mixed,
        'speedup_vs_fp64': dt_fp64 / dt, 'network_forward_ms': fwd_ms,
        'network_only_speedup_vs_fp64': dt_fp64 / (fwd_ms * 1e-3),
        'network_16bit_activations': acts16,
        'fallback_rate': float(fl.mean()),
        'at_defect_tol_3e-2': {'dof_evals_per_s': ndof / dt_l, 'speedup_vs_fp64': dt_fp64 / dt_l,
                               'fallback_rate': stats_l['fallback'] / ndof,
                               'sampled_check_rel_l2_median': stats_l['check_rel_l2_median'],
                               'rel_l2_vs_fp64_median': float(((from_aosoa(ql) - from_aosoa(qref)).pow(2).sum(1)
                                                               / from_aosoa(qref).pow(2).sum(1)).sqrt().median())},
        'fallback_rate_at_tol': {t: float((dfc > t).mean()) for t in (1e-3, 3e-3, 1e-2, 3e-2)},
        'defect': {'median': float(np.median(dfc)), 'p90': float(np.quantile(dfc, 0.9))},
        'network_projected_rel_l2_vs_fp64_all_dofs': {'median': float(np.median(rn)),
                                                      'p90': float(np.quantile(rn, 0.9)),
                                                      'max': float(rn.max())},
        'network_projected_rel_l2_vs_fp64_projected_all_dofs': {'median': float(np.median(rnp)),
                                                                'p90': float(np.quantile(rnp, 0.9)),
                                                                'max': float(rnp.max())},
        'spectral_fp64_defect': {'median': float(np.median(tdef)), 'p90': float(np.quantile(tdef, 0.9)),
                                 'fraction_above_tol': float((tdef > sm.tol).mean())},
        'rel_l2_vs_fp64': {'network_dofs_median': float(np.median(net_r)),
                           'network_dofs_p90': float(np.quantile(net_r, 0.9)),
                           'network_dofs_max': float(net_r.max()), 'all_dofs_median': float(np.median(r))},
        'spectral_mixed_rel_l2_median': rmix,
        'training': rep,
        'roofline': {'bound': 'fp32-valu', 'achieved': tflops, 'peak': PEAK['fp32'], 'unit': 'TFLOP/s',
                     'frac': tflops / PEAK['fp32'], 'traffic': None,
                     'kernel': 'network forward (all its kernels), ' + str(flop) + ' flop per DOF'},
        'cpu_baseline': cpu}))


def bench_rhs(a):
    """Full explicit step: 32x32 periodic quads, p = 2 (9216 DOFs), 24^3 velocity
    grid on [-5,5]^3, vhs-gll M = 6 (the 12-point design as 6 antipodal pairs),
    Nrho = 12; SSP-RK2 with FR upwind advection and Q, each step captured once
    in a HIP graph and replayed.  Initial f: Maxwellian with a seeded
    sinusoidal shear u_x = 0.05 sin(2 pi y + phase)."""
    import torch
    from frfs_amd.backend import HIPBackend
    from frfs_amd.fr import PeriodicQuadAdvection
    from frfs_amd.inifile import Inifile
    from frfs_amd.system import DGFSCollisionSystem, STEPPERS
    be = HIPBackend()
    nel, N = 32, 24
    res = {}
    for mode in [m for m in a.modes.split(',') if m in ('fp64', 'mixed')]:
        cfg = Inifile.from_dict({
            'non-dim': dict(T0=273.0, H0=1e-3, rho0=0.00011140244750324529, molarMass0=39.96e-3),
            'constants': {'Nv': N, 'NvBatchSize': N ** 3, 'Nrho': 12},
            'velocity-mesh': {'dev': 5.0, 'cmax': 0, 'Tmax': 273},
            'spherical-design-rule': {'ssrule': 'womersley', 'M': 6},
            'scattering-model': sm_opts(a, {'type': 'vhs-gll', 'omega': 0.5, 'dRef': 4.17e-10, 'Tref': 273,
                                            'precision': 'mixed' if mode == 'mixed' else 'double',
                                            'conservation-projection': 1 if mode == 'mixed' else 0,
                                            'autotune': a.autotune})})
        sysm = DGFSCollisionSystem(be, cfg, 9, nel * nel)
        adv = PeriodicQuadAdvection(be, sysm.vm, nel, nel, 1.0 / nel, 1.0 / nel, order=2)
        sysm.rhs = lambda t, uin, fout: adv(sysm.banks[uin], sysm.banks[fout])
        stp = STEPPERS['dgfs-tvd-rk2'](sysm, graph=True)
        rng = np.random.default_rng(a.seed)
        phase = rng.uniform(0, 2 * np.pi)
        ploc = adv.ploc()
        cv = torch.as_tensor(sysm.vm.cv(), device='cuda')
        ux = torch.as_tensor(0.05 * np.sin(2 * np.pi * ploc[:, 1, :] + phase), device='cuda')   # (9, neles)
        f0 = (np.pi ** -1.5) * torch.exp(-((cv[0][None, :, None] - ux[:, None, :]) ** 2
                                           + cv[1][None, :, None] ** 2 + cv[2][None, :, None] ** 2))
        sysm.banks[0].set(f0.cpu().numpy())
        dt = 1e-3
        for k in range(a.warmup):
            stp.step(k * dt, dt)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(a.steps):
            stp.step(k * dt, dt)
        torch.cuda.synchronize()
        res[mode] = (time.perf_counter() - t0) / a.steps
        del sysm, adv, stp
        torch.cuda.empty_cache()
    ndof = nel * nel * 9
    head = 'mixed' if 'mixed' in res else list(res)[0]
    out = {'metric': 'full SSP-RK2 step DOF-steps/s (FR advection + Q(f,f)), 24^3 grid, 9216 DOFs',
           'value': ndof / res[head], 'unit': 'DOF-steps/s', 'n_gpus': 1, 'steps': a.steps,
           'warmup': a.warmup, 'ms_per_step': res[head] * 1e3, 'higher_is_better': True,
           'scaling': 'weak', 'vs_baseline': None,
           'dtype': 'f32+f64acc' if head == 'mixed' else 'f64', 'data': 'synthetic',
           'config': {'workload': 'dgfs-tvd-rk2 step, 32x32 periodic quads p=2 (9216 DOFs), 24^3 grid on '
                                  '[-5,5]^3, vhs-gll M=6 (12-point design) Nrho=12, HIP graph per step',
                      'mode': head},
           'collision_dof_evals_per_s': 2 * ndof / res[head],
           'fp64': ({'value': ndof / res['fp64'], 'ms_per_step': res['fp64'] * 1e3} if 'fp64' in res else None)}
    print(json.dumps(out))


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({'error': 'this build runs on one GPU; --gpus must be 1',
                          'n_gpus_requested': a.gpus}))
        sys.exit(2)
    if a.workload == 'rhs24':
        bench_rhs(a)
        return
    if a.workload == 'vhs16':
        bench_vhs16(a)
        return
    if a.workload == 'vss32x2':
        bench_multi(a)
        return
    if a.workload in ('fno16', 'fno32'):
        bench_fno(a)
        return
    from frfs_amd.inputs import mixture_params
    fpar = mixture_params(a.ndof, a.seed)
    cpu, ref = (None, {}) if a.no_cpu_baseline else cpu_baseline(a, fpar)

    import torch
    from frfs_amd.autotune import ONE_STREAM
    from frfs_amd.backend import HIPBackend
    from frfs_amd.inputs import from_aosoa, mixture_device, to_aosoa
    from frfs_amd.scattering import DGFSBiVSSGLLScatteringModel
    from frfs_amd.velocitymesh import DGFSBiVelocityMesh

    be = HIPBackend()
    if a.calib_copy_mib:
        src = torch.ones(a.calib_copy_mib << 17, dtype=torch.float64, device='cuda')
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    modes = a.modes.split(',')
    res = {}
    qs = {}
    # fp64: plain FP64; mixed: FP32 transforms and products, FP64 weight
    # accumulation; mixedproj: mixed plus the FP64 conservation projection.
    # FP64 with the projection is evaluated once (untimed) so that rounding
    # and the projection's own correction are reported separately.
    runs = [(m, m == 'mixedproj', True) for m in modes]
    if 'mixedproj' in modes and not a.no_reference_proj:
        runs.append(('fp64proj', True, False))
    for mode, proj, timed in runs:
        cfg = vss_cfg(a, 'mixed' if mode.startswith('mixed') else 'double', int(proj))
        vm = DGFSBiVelocityMesh(be, cfg, 1)
        sm = DGFSBiVSSGLLScatteringModel(be, cfg, vm)
        n = vm.vsize()
        vm_cv = vm.cv()
        fin = be.matrix((1, n, a.ndof))
        qout = be.matrix((1, n, a.ndof))
        to_aosoa(mixture_device(vm.cv(), fpar), fin)
        torch.cuda.synchronize()
        if not timed:
            sm.fs_batch(fin, qout)
            qs[mode] = from_aosoa(qout).cpu().numpy()
            del sm, fin, qout
            torch.cuda.empty_cache()
            continue
        # the schedule autotuner (untimed): candidates timed on these arrays,
        # the fastest kept for the timed steps
        sched = sm.autotune(fin, qout, candidates=ONE_STREAM)['choice'] if a.autotune else None
        for _ in range(a.warmup):
            sm.fs_batch(fin, qout)
        torch.cuda.synchronize()
        sm._h.profile(True)
        # per-step HIP events on the stream the operator runs on (torch's
        # current stream), and the wall clock around all K steps
        st = torch.cuda.current_stream()
        evs = [torch.cuda.Event(enable_timing=True) for _ in range(a.steps + 1)]
        t0 = time.perf_counter()
        evs[0].record(st)
        for k in range(a.steps):
            sm.fs_batch(fin, qout)
            evs[k + 1].record(st)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / a.steps
        step_ms = [evs[k].elapsed_time(evs[k + 1]) for k in range(a.steps)]
        prof = sm._h.profile_read()
        sm._h.profile(False)
        info = sm.plan_info()
        res[mode] = dict(ms=dt * 1e3, rate=a.ndof / dt, prof=prof, info=info, step_ms=step_ms,
                         median_ms=float(np.median(step_ms)), chunk=sm._h.chunk_dofs(),
                         scratch=sm._h.scratch_bytes(), schedule=sched, paths=sm._h.paths())
        qs[mode] = from_aosoa(qout).cpu().numpy()
        del sm, fin, qout
        torch.cuda.empty_cache()

    # The same plan with every 3D transform in the ROCm FFT library (hipFFT)
    # and pointwise kernels between them: the reference's structure (cuFFT
    # plans plus pointwise kernels) on this GPU.  Untuned, wall clock over
    # a few steps after one warmup step.
    lib = {}
    for mode in (m for m in ('fp64', 'mixedproj') if m in res and a.library_steps > 0):
        cfg = vss_cfg(a, 'mixed' if mode.startswith('mixed') else 'double', int(mode == 'mixedproj'),
                      {'library-fft': 1, 'max-chunk': 512})
        vm = DGFSBiVelocityMesh(be, cfg, 1)
        sm = DGFSBiVSSGLLScatteringModel(be, cfg, vm)
        fin = be.matrix((1, vm.vsize(), a.ndof))
        qout = be.matrix((1, vm.vsize(), a.ndof))
        to_aosoa(mixture_device(vm.cv(), fpar), fin)
        sm.fs_batch(fin, qout)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.library_steps):
            sm.fs_batch(fin, qout)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / a.library_steps
        q = from_aosoa(qout).cpu().numpy()
        lib[mode] = {'value': a.ndof / dt, 'ms_per_step': dt * 1e3, 'paths': sm._h.paths(),
                     'native_speedup': res[mode]['rate'] * dt / a.ndof,
                     'rel_l2_vs_native': float(np.linalg.norm(q - qs[mode]) / np.linalg.norm(qs[mode]))}
        del sm, fin, qout
        torch.cuda.empty_cache()

    N = a.Nv
    n = N ** 3
    info = res[modes[0]]['info']
    f_alg = info['n_fft_ref'] * 5 * n * math.log2(n)          # nominal flops per DOF-eval
    # the headline is the mixed operator with the FP64 conservation projection
    # (the workload's mixed mode); without projection when that is not run
    head = 'mixedproj' if 'mixedproj' in res else ('mixed' if 'mixed' in res else modes[0])
    r = res[head]
    prec_peak = PEAK['fp32'] if head.startswith('mixed') else PEAK['fp64']
    tot = {k: v for k, v in r['prof'].items() if v[1] > 0}
    dom = max(tot, key=lambda k: tot[k][0])
    dom_ms, dom_n = tot[dom]
    dom_avg_s = dom_ms / dom_n * 1e-3
    dofs_per_launch = a.ndof * a.steps / dom_n
    # Share of F_alg per kernel.  Per VSS pair the reference runs 3 M Nrho + 4
    # complex 3D transforms, 3 N^2 length-N passes each: 2 M Nrho inverse
    # item transforms and M Nrho forward product transforms, plus the forward
    # transforms of f and the inverse gain and loss transforms.  The product
    # kernel performs all item passes and the x, y passes of the forward
    # product transforms (8 M Nrho of the 3 (3 M Nrho + 4) passes); the
    # accumulation kernel the z passes of the forward product transforms and
    # of the inverse gain and loss transforms (M Nrho + 2).
    MN = a.M * a.Nrho
    passes_ref = 3 * (3 * MN + 4)
    share = {'k_prod': 8 * MN / passes_ref, 'k_accum': (MN + 2) / passes_ref,
             'k_fwd_z': 1 / passes_ref, 'k_fwd_xy': 2 / passes_ref, 'k_out': 4 / passes_ref,
             'k_nyq': 0.0, 'k_project': 0.0}
    achieved = share.get(dom, 0.0) * f_alg * dofs_per_launch / dom_avg_s / 1e12
    # what the dominant kernel executes: with pair packing one complex 3D
    # inverse transform per product (3 N^2 passes) and the x, y passes of one
    # forward transform per group (2 N^2), or per group pair with group
    # pairs (N^2 per group); 5 N log2 N flops per pass
    executed = None
    if dom == 'k_prod':
        npr = info['n_groups']
        fwd = 1 if 'group_pairs' in r['paths'] else 2
        executed = npr * (3 + fwd) * N * N * 5 * N * math.log2(N) * dofs_per_launch / dom_avg_s / 1e12
    op_eff = a.ndof / (r['median_ms'] * 1e-3) * f_alg / 1e12
    traffic, traffic_src = None, None
    if (dom == 'k_prod' and {'Nv': N, 'M': a.M, 'Nrho': a.Nrho, 'Mpre': a.Mpre} == K_PROD_TRAFFIC['shape']):
        traffic = K_PROD_TRAFFIC['per_dof']['mixed' if head.startswith('mixed') else 'fp64'] * dofs_per_launch
        traffic_src = (K_PROD_TRAFFIC['source'] + ': FETCH_SIZE + WRITE_SIZE per dispatch of this kernel at '
                       'this shape (counter runs with 1440-DOF batches), scaled by a '
                       '1 GiB copy, per DOF times the DOFs of one launch')
    # live HBM rate of the accumulation kernel (the operator's HBM-bound kernel):
    # its counter bytes per DOF x DOFs per launch / its HIP-event launch time
    accum_hbm = None
    if {'Nv': N, 'M': a.M, 'Nrho': a.Nrho, 'Mpre': a.Mpre} == K_PROD_TRAFFIC['shape'] and 'k_accum' in tot:
        acc_ms, acc_n = tot['k_accum']
        if acc_n:
            per = K_PROD_TRAFFIC['accum_per_dof']['mixed' if head.startswith('mixed') else 'fp64']
            gbps = per * (a.ndof * a.steps / acc_n) / (acc_ms / acc_n * 1e-3) / 1e9
            accum_hbm = {'kernel': 'k_accum', 'GBps': gbps, 'hbm_frac': gbps / HBM_PEAK,
                         'bytes_per_dof': per, 'source': K_PROD_TRAFFIC['source']}
    err = None

    def relerr(x, y):
        d = x - y
        return dict(rel_l2=float(np.linalg.norm(d) / np.linalg.norm(y)),
                    max_dof_rel_l2=float(np.max(np.linalg.norm(d, axis=1) / np.linalg.norm(y, axis=1))))
    proj = None
    if 'mixed' in qs and 'fp64' in qs:
        err = relerr(qs['mixed'], qs['fp64'])
    # collision invariants per DOF: max |sum_v Q phi| / sum_v |Q phi| over the batch
    cv = np.asarray(vm_cv)
    phi = np.stack([np.ones(cv.shape[1]), cv[0], cv[1], cv[2], (cv ** 2).sum(0)])

    def defect(q):
        return float(np.max(np.abs(q @ phi.T) / (np.abs(q) @ np.abs(phi.T))))
    if 'mixedproj' in res:
        proj = {'value': res['mixedproj']['rate'], 'ms_per_step': res['mixedproj']['ms'],
                'moment_defect': defect(qs['mixedproj'])}
        if 'fp64' in qs:
            proj['moment_defect_fp64'] = defect(qs['fp64'])
            proj['vs_fp64'] = relerr(qs['mixedproj'], qs['fp64'])
        if 'fp64proj' in qs:
            proj['vs_fp64_with_projection'] = relerr(qs['mixedproj'], qs['fp64proj'])
            if 'fp64' in qs:
                proj['projection_correction'] = relerr(qs['fp64proj'], qs['fp64'])
    oracle_err = None
    vs_oracle = {}
    if ref:
        ds = sorted(ref)
        for m in ('fp64', 'mixed'):
            if m in qs:
                vs_oracle[m] = float(max(np.linalg.norm(qs[m][d] - ref[d]) / np.linalg.norm(ref[d]) for d in ds))
        oracle_err = vs_oracle.get('fp64')
    # the mixed mode's accuracy target (rel-L2 <= 1e-5) and what it is measured
    # against: the projected mixed operator against the FP64 operator with the
    # same projection (the two differ by FP32 rounding only); the distance to
    # the plain FP64 operator is dominated by the projection's own correction,
    # which is reported beside it
    accuracy = {'target_rel_l2': 1e-5}
    if proj and 'vs_fp64_with_projection' in proj:
        accuracy['mixedproj_vs_fp64proj'] = proj['vs_fp64_with_projection']
        accuracy['target_met_vs_fp64proj'] = proj['vs_fp64_with_projection']['rel_l2'] <= 1e-5
    if proj and 'vs_fp64' in proj:
        accuracy['mixedproj_vs_plain_fp64'] = proj['vs_fp64']
        accuracy['target_met_vs_plain_fp64'] = proj['vs_fp64']['rel_l2'] <= 1e-5
    if proj and 'projection_correction' in proj:
        accuracy['fp64proj_vs_plain_fp64'] = proj['projection_correction']
    if err:
        accuracy['mixed_unprojected_vs_plain_fp64'] = err
    if vs_oracle:
        accuracy['max_dof_rel_l2_vs_oracle'] = dict(vs_oracle, dofs=len(ref), Nrho=a.Nrho)
    out = {
        'metric': 'Q(f,f) DOF-evals/s, 32^3 grid, mixed/surrogate vs plain FP64 at stated L2 err',
        'value': r['rate'], 'unit': 'DOF-evals/s', 'n_gpus': 1, 'steps': a.steps, 'warmup': a.warmup,
        'ms_per_step': r['ms'], 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f32+f64acc' if head.startswith('mixed') else 'f64', 'data': 'synthetic',
        'median_of_steps': {'steps': a.steps, 'median_ms': r['median_ms'],
                            'value': a.ndof / (r['median_ms'] * 1e-3),
                            'fp64_median_ms': res['fp64']['median_ms'] if 'fp64' in res else None,
                            'speedup_vs_fp64': (res['fp64']['median_ms'] / r['median_ms'])
                            if 'fp64' in res else None},
        'accuracy': accuracy,
        'config': {'workload': f'vss-gll Q(f,f), {N}^3 grid on [-6,6]^3, M={a.M} full sphere, '
                               f'Nrho={a.Nrho}, omega=0.74, alpha=1.4, Mpre={a.Mpre}, single species, '
                               f'{a.ndof} DOFs of seeded two-Maxwellian f'
                               + ('; FP32 transforms and products, FP64 weight accumulation and FP64 '
                                  'conservation projection' if head == 'mixedproj' else ''),
                   'mode': head, 'ndof': a.ndof, 'Nv': N, 'M': a.M, 'Nrho': a.Nrho, 'Mpre': a.Mpre,
                   'seed': a.seed, 'chunk_dofs': r['chunk'], 'scratch_GiB': r['scratch'] / 2 ** 30,
                   'schedule': r['schedule'], 'paths': r['paths']},
        'fp64': {'value': res['fp64']['rate'], 'ms_per_step': res['fp64']['ms'], 'chunk_dofs': res['fp64']['chunk'],
                 'scratch_GiB': res['fp64']['scratch'] / 2 ** 30,
                 'schedule': res['fp64']['schedule']} if 'fp64' in res else None,
        'speedup_vs_fp64': (r['rate'] / res['fp64']['rate']) if 'fp64' in res else None,
        'mixed_without_projection': {'value': res['mixed']['rate'], 'ms_per_step': res['mixed']['ms'],
                                     'speedup_vs_fp64': (res['mixed']['rate'] / res['fp64']['rate'])
                                     if 'fp64' in res else None} if 'mixed' in res else None,
        'mixed_vs_fp64_error': err,
        'mixed_with_projection': proj,
        'fp64_vs_oracle_max_rel_l2': oracle_err,
        'plan': info,
        'roofline': {'bound': 'valu', 'kernel': dom, 'achieved': achieved, 'peak': prec_peak,
                     'unit': 'TFLOP/s', 'frac': achieved / prec_peak, 'traffic': traffic,
                     'traffic_source': traffic_src,
                     'traffic_GBps': traffic / dom_avg_s / 1e9 if traffic else None,
                     'hbm_frac': traffic / dom_avg_s / 1e9 / HBM_PEAK if traffic else None,
                     'method': 'share of F_alg done by the kernel x F_alg per DOF x DOFs per launch / '
                               'average launch time (HIP events on the launch stream) / peak',
                     'f_alg_per_dof': f_alg, 'kernel_share_of_f_alg': share.get(dom, 0.0),
                     'dofs_per_launch': dofs_per_launch,
                     'avg_launch_ms': dom_avg_s * 1e3, 'launches': dom_n,
                     'executed_tflops': executed,
                     'operator_effective_tflops': op_eff, 'operator_frac': op_eff / prec_peak,
                     'accum_hbm': accum_hbm,
                     'mfma_util': 0.0,
                     'mfma_note': 'the transforms run on the vector ALU: the Q path issues no MFMA '
                                  'instructions (dense DFT stages as GEMMs cost more than the radix form)',
                     'kernel_ms_per_step': {k: v[0] / a.steps for k, v in tot.items()}},
        'cpu_baseline': cpu,
        'library_fft_path': ({**lib, 'steps': a.library_steps, 'chunk_dofs': 512,
                              'note': 'the same plan with every 3D transform in hipFFT (batched plans) and '
                                      'pointwise kernels between them, as for grids without native kernels; '
                                      'native_speedup = native rate / library rate of the same mode'}
                             if lib else None),
    }
    if 'fp64' in res and head != 'fp64':
        rp = {k: v for k, v in res['fp64']['prof'].items() if v[1] > 0}
        out['fp64']['kernel_ms_per_step'] = {k: v[0] / a.steps for k, v in rp.items()}
    print(json.dumps(out))


if __name__ == '__main__':
    main()
