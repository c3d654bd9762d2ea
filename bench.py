#!/usr/bin/env python3
"""Throughput of the D3Q19 MRT time step on one MI355X.

Workload: single-phase D3Q19 MRT permeability lattice on a 256^3 periodic
overlapping-sphere pack (radius uniform in [4,10], porosity 0.20, seed 2),
body force 1e-6 along x, tau 0.8, rho0 1 (synthetic geometry from
lbflow_amd.geometry).  A "step" is one full lattice update (stream,
bounce-back, macro, force, MRT collision) of the whole domain with the
populations resident in HBM.

value = site MLUPS (lattice sites updated per second / 1e6, every voxel of
the 256^3 box counted, as the reference streams every voxel); fluid MLUPS
(the reference's own metric) is reported beside it.  The dominant kernel's
roofline uses its algorithmic HBM bytes per launch and its average launch
time measured with HIP events on the engine's stream.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, public specification
METRIC = 'D3Q19 MRT MLUPS + % HBM roofline; time-to-permeability warm vs cold'


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(size):
    from lbflow_amd import geometry as G
    labels = G.sphere_pack((size, size, size), 0.20, 4.0, 10.0, seed=2)
    return pack_case(f'c2-{size}', labels), labels


def pack_case(name, labels):
    """Single-phase D3Q19 MRT, tau 0.8, F 1e-6 along x on pore label 75,
    halfway bounce-back on label 255."""
    from lbflow_amd import cases as CS
    return CS.make(name, labels, [CS.label_line(75, 'GV', 1e-6, 1, 0), CS.label_line(255, 'SD', 90)],
                   model='scsp', eq='MRT', tau=(0.8,))


PMC_PROFILE = ROOT / 'profiles' / 'scsp_step_f32_sparse_256cubed_pmc.md'


def measured_traffic(kernel, size):
    """HBM bytes per launch of the sparse FP32 step at 256^3, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes (scaled by the copy
    kernel's counters), or None for another kernel or size."""
    if size != 256 or 'sparse' not in kernel or not PMC_PROFILE.exists():
        return None
    for line in PMC_PROFILE.read_text().splitlines():
        if line.startswith('traffic_bytes_per_launch:'):
            return int(line.split(':', 1)[1])
    return None


def measure(spec, labels, precision, stream_solids, steps, warmup, dense=False):
    from lbflow_amd import cases as CS
    from lbflow_amd.engine import Engine
    case = CS.lattice(spec)
    eng = Engine(case, labels.ravel(), precision, stream_solids, dense).init()
    eng.advance(warmup)
    eng.synchronize()
    t0 = time.perf_counter()
    ms = eng.timed_advance(steps)
    eng.synchronize()
    wall = time.perf_counter() - t0
    eng.check_status()
    out = dict(kernel=eng.step_kernel, bytes_per_launch=eng.step_bytes, ms_per_step=ms / steps,
               wall_s=wall, device_bytes=eng.device_bytes, sites=eng.N)
    eng.close()
    return out


def cpu_baseline(spec, labels, budget_s):
    """The CPU oracle (single-threaded C restatement of the reference) timed on
    the same 256^3 lattice for as many whole steps as fit the budget."""
    from lbflow_amd import cases as CS
    from oracle.pyoracle import Oracle
    case = CS.lattice(spec)
    o = Oracle(case, labels.ravel()).init()
    t0 = time.perf_counter()
    o.step(1)
    one = time.perf_counter() - t0
    n = max(1, min(10, int(budget_s / max(one, 1e-3)) - 1))
    t0 = time.perf_counter()
    o.step(n)
    dt = time.perf_counter() - t0
    steps = n
    return dict(value=labels.size * steps / dt / 1e6, unit='site MLUPS', cores=1, kind='port',
                sample=f'oracle/lbf_oracle.c, same 256^3 D3Q19 MRT lattice, {steps} steps after 1 warm-up '
                       f'step, {dt:.1f} s, 1 thread')


def colour_bench(steps, warmup, fp64=False):
    """Two-phase colour-gradient D3Q19 MRT on a 128x128x256 periodic sphere pack
    (z y x = 256 128 128, radius 6-12, porosity 0.30, seed 3), random 8^3 blocks
    of the two phases (50 %), tau 1/1, densities 1/1, alpha0 1/3, sigma 1e-3,
    beta 0.7, contact angle 60 deg, force 1e-6 along z (direction 5)."""
    from lbflow_amd import geometry as G
    from lbflow_amd import cases as CS
    from lbflow_amd.engine import FP32, FP64
    labels = G.sphere_pack((256, 128, 128), 0.30, 6.0, 12.0, seed=3)
    labels = G.random_blocks(labels, 8, 0.5, seed=103, phase0=75, phase1=76, pore=75)
    spec = CS.make('c3', labels, [CS.label_line(75, 'GV', 1e-6, 5, 0), CS.label_line(76, 'GV', 1e-6, 5, 1),
                                  CS.label_line(255, 'SD', 60)],
                   model='color', eq='MRT', tau=(1.0, 1.0), sigma=1e-3, beta=0.7, alpha=1 / 3)
    out = {'workload': 'colour-gradient D3Q19 MRT, 128x128x256 sphere pack (r 6-12, porosity '
                       f'{float(np.isin(labels, (75, 76)).mean()):.4f}, seed 3), 8^3 phase blocks 50 %, '
                       'tau 1/1, sigma 1e-3, beta 0.7, 60 deg, F 1e-6 z'}
    for name, prec, solids in (('fp32', FP32, False), ('fp64', FP64, True)):
        if name == 'fp64' and not fp64:
            continue
        r = measure(spec, labels, prec, solids, steps if name == 'fp32' else max(5, steps // 4),
                    warmup if name == 'fp32' else 2)
        out[name] = dict(kernel=r['kernel'], ms_per_step=round(r['ms_per_step'], 5),
                         site_mlups=round(labels.size / r['ms_per_step'] / 1e3, 1),
                         bytes_per_launch=r['bytes_per_launch'],
                         achieved_gbs=round(r['bytes_per_launch'] / r['ms_per_step'] / 1e6, 1))
        log('colour', name, out[name])
    return out


def time_to_permeability(spec, labels, interval, max_steps):
    """Permeability of the 256^3 pack in FP64 and FP32: wall clock from init to the
    reference's stop rule (residual 1e-6, three stable reports) and the
    Darcy K of each, with |K32 - K64| / K64."""
    from lbflow_amd import cases as CS
    from lbflow_amd.engine import FP32, FP64, Engine
    from lbflow_amd.permeability import converge
    case = CS.lattice(spec)
    n = labels.shape[2]
    out = {}
    for name, prec in (('fp64', FP64), ('fp32', FP32)):
        t0 = time.perf_counter()
        eng = Engine(case, labels.ravel(), prec, False).init()
        r = converge(eng, axis=2, first=1, last=n - 2, tol=1e-6, interval=interval, max_steps=max_steps,
                     force=1e-6, tau=0.8)
        wall = time.perf_counter() - t0
        eng.close()
        out[name] = dict(steps=r['steps'], converged=r['converged'], K_lattice=r['K'], seconds=round(wall, 3))
        log('time-to-permeability', name, out[name])
    out['k_rel_diff_fp32_vs_fp64'] = abs(out['fp32']['K_lattice'] - out['fp64']['K_lattice']) / out['fp64']['K_lattice']
    out['report_interval'] = interval
    return out


def ttp_warm_vs_cold(labels, interval, max_steps):
    """Time to permeability of the 256^3 pack from rest and from the warm-start
    network (trained on 64^3 crops; fully convolutional and periodic, so it
    runs on the whole 256^3 lattice): the batched FP32 engine with one crop,
    same stop rule (residual 1e-6, reports every `interval` steps), wall
    clock from init to the stop, inference and set-up included."""
    from lbflow_amd import surrogate as S
    weights = ROOT / 'lbflow_amd' / 'weights' / 'warm64.npz'
    crops = [np.ascontiguousarray(labels)]
    cold, t_cold, _ = S.run_crops(crops, 0.8, 1e-6, 1e-6, interval, max_steps)
    net, scaling = S.load(weights, capacity=1, side=labels.shape[0])
    warm, t_warm, t_inf = S.run_crops(crops, 0.8, 1e-6, 1e-6, interval, max_steps, net, scaling, relax=5)
    net.close()
    out = dict(cold=dict(steps=cold[0]['steps'], converged=bool(cold[0]['converged']), K_lattice=cold[0]['K'],
                         seconds=round(t_cold, 3)),
               warm=dict(steps=warm[0]['steps'], converged=bool(warm[0]['converged']), K_lattice=warm[0]['K'],
                         seconds=round(t_warm, 3), inference_and_relax_seconds=round(t_inf, 3)),
               report_interval=interval)
    out['speedup_iterations'] = round(cold[0]['steps'] / max(1, warm[0]['steps']), 3)
    out['speedup_wall'] = round(t_cold / t_warm, 3)
    out['k_rel_diff_warm_vs_cold'] = abs(warm[0]['K'] - cold[0]['K']) / cold[0]['K']
    log('time-to-permeability warm vs cold', out)
    return out


def slab_bench(labels, ways, steps, warmup):
    """The FP32 sparse step on one GPU's share of an R-way z-slab sharding of
    the lattice: the first nz / R planes (periodic in z here; in a sharded run
    the two boundary planes pull from the neighbouring slabs instead).  Per
    step each slab would exchange the 5 populations crossing each z-face."""
    from lbflow_amd import cases as CS
    from lbflow_amd.engine import FP32
    out = {}
    nz, ny, nx = labels.shape
    for r in ways:
        sub = np.ascontiguousarray(labels[:nz // r])
        spec = pack_case(f'slab-{r}', sub)
        m = measure(spec, sub, FP32, False, steps, warmup)
        fluid_face = int((sub[0] == 75).sum())
        out[f'{r}-way'] = dict(shape=f'{nx}x{ny}x{nz // r}', kernel=m['kernel'],
                               ms_per_step=round(m['ms_per_step'], 5),
                               site_mlups_per_gpu=round(sub.size / m['ms_per_step'] / 1e3, 1),
                               porosity=round(float((sub == 75).mean()), 4),
                               exchange_bytes_per_face=5 * 4 * fluid_face,
                               achieved_gbs=round(m['bytes_per_launch'] / m['ms_per_step'] / 1e6, 1))
    return out


def batch_crops(ncrops):
    """The batched workload's crops (before the GPU is touched: a process pool forks)."""
    from lbflow_amd import geometry as G
    t0 = time.perf_counter()
    crops, _ = G.batch_crops(ncrops, workers=16)
    log(f'batch: {ncrops} crops generated in {time.perf_counter() - t0:.1f} s')
    return crops


def batch_bench(crops, interval, max_steps, tune_only=False):
    """Batched permeability: independent cubic sphere-pack crops (side 32-96,
    porosity 0.15-0.35, seeds 1000+i), D3Q19 MRT FP32, tau 0.8, F 1e-6 along x,
    every crop run to residual 1e-5 in one persistent hipGraph-captured launch
    sequence.  With tune_only, only the autotuner's eager steps of all crops
    run (no captured graph, no convergence run)."""
    from lbflow_amd.engine import Batch
    ncrops = len(crops)
    labels = {75: ('GV', 1e-6, 1, 0), 255: ('SD', 90.0, 0, 0)}
    b = Batch(3, 19, 'MRT', 0.8, 1.0, labels, crops)
    per_cu, tune_ms = b.tune(steps=10)
    log(f'batch: {b.sites} sites, {b.fluid} fluid, tuned {per_cu} workgroups/CU, {tune_ms:.3f} ms/step')
    if tune_only:
        out = dict(crops=ncrops, sites=b.sites, fluid_sites=b.fluid, tuned_workgroups_per_cu=per_cu,
                   all_crops_ms_per_step=round(tune_ms, 4))
        b.close()
        return out
    t0 = time.perf_counter()
    b.init()
    res, sec = b.run(axis=2, tol=1e-5, interval=interval, max_steps=max_steps)
    wall = time.perf_counter() - t0
    site_updates = sum(int(g.size) * r['steps'] for g, r in zip(crops, res))
    fluid_updates = sum(int((g == 75).sum()) * r['steps'] for g, r in zip(crops, res))
    steps = [r['steps'] for r in res]
    out = dict(crops=ncrops, converged=sum(r['converged'] for r in res), seconds=round(wall, 3),
               crops_per_s=round(ncrops / wall, 3), site_mlups=round(site_updates / wall / 1e6, 1),
               fluid_mlups=round(fluid_updates / wall / 1e6, 1), sites=b.sites, fluid_sites=b.fluid,
               steps_min=min(steps), steps_max=max(steps), steps_mean=round(sum(steps) / len(steps), 1),
               report_interval=interval, tuned_workgroups_per_cu=per_cu,
               all_crops_ms_per_step=round(tune_ms, 4),
               workload='512 cubic sphere-pack crops, side U{32..96}, porosity U[0.15,0.35], seeds 1000+i, '
                        'D3Q19 MRT FP32, tau 0.8, F 1e-6 x, residual 1e-5')
    b.close()
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=200)
    ap.add_argument('--warmup', type=int, default=20)
    ap.add_argument('--size', type=int, default=256)
    ap.add_argument('--cpu-budget', type=float, default=20.0, help='seconds of CPU oracle timing (0: skip)')
    ap.add_argument('--fp64', action='store_true', help='also time the FP64 reference-order kernel')
    ap.add_argument('--ttp', action='store_true', help='also run FP64 and FP32 to convergence (time-to-K)')
    ap.add_argument('--ttp-interval', type=int, default=500)
    ap.add_argument('--ttp-max-steps', type=int, default=100000)
    ap.add_argument('--ttp-warm', action='store_true',
                    help='also run the 256^3 pack to convergence from rest and from the warm-start network')
    ap.add_argument('--colour', action='store_true', help='also time the two-phase colour model (128x128x256)')
    ap.add_argument('--batch', type=int, default=0, help='also run N batched crops to convergence (512)')
    ap.add_argument('--slabs', default='', help='also time one GPU\'s z-slab of the lattice for R-way '
                    'sharding, e.g. 2,4,8 (the step at the shape each GPU would hold)')
    ap.add_argument('--batch-interval', type=int, default=200)
    ap.add_argument('--batch-max-steps', type=int, default=60000)
    ap.add_argument('--batch-tune-only', action='store_true',
                    help='with --batch: only tune and time one eager step of all crops')
    args = ap.parse_args()
    if args.gpus != 1:
        print(f'bench.py: the engine runs on one GPU; --gpus {args.gpus} is not supported '
              '(sharding across GPUs is not built)', file=sys.stderr)
        sys.exit(2)

    crops = batch_crops(args.batch) if args.batch else None
    if args.ttp_warm:
        # the network's tensors live in torch's HIP runtime, which must start
        # before the engine's (two runtimes share the process)
        from lbflow_amd import surrogate as S
        S._torch()
    from lbflow_amd.engine import FP32, FP64, copy_bandwidth
    spec, labels = workload(args.size)
    porosity = float((labels == 75).mean())
    log(f'workload {args.size}^3 porosity {porosity:.4f}')
    copy_gbs = copy_bandwidth(1 << 30, 20)
    log(f'copy kernel {copy_gbs:.0f} GB/s')
    main_run = measure(spec, labels, FP32, False, args.steps, args.warmup)
    log('fp32', main_run)
    extra = {}

    def record(name, r):
        extra[name] = dict(kernel=r['kernel'], ms_per_step=round(r['ms_per_step'], 5),
                           site_mlups=round(labels.size / r['ms_per_step'] / 1e3, 1),
                           bytes_per_launch=r['bytes_per_launch'],
                           achieved_gbs=round(r['bytes_per_launch'] / r['ms_per_step'] / 1e6, 1))
    record('fp32_dense_all_sites', measure(spec, labels, FP32, True, args.steps, args.warmup))
    record('fp32_dense_fluid_sites', measure(spec, labels, FP32, False, args.steps, args.warmup, dense=True))
    if args.fp64:
        record('fp64_sparse', measure(spec, labels, FP64, False, max(10, args.steps // 2), 5))
        record('fp64_dense_all_sites', measure(spec, labels, FP64, True, max(10, args.steps // 4), 5))
    if args.ttp:
        extra['time_to_permeability'] = time_to_permeability(spec, labels, args.ttp_interval, args.ttp_max_steps)
    if args.ttp_warm:
        extra['time_to_permeability_warm_vs_cold'] = ttp_warm_vs_cold(labels, args.ttp_interval, args.ttp_max_steps)
    if args.colour:
        extra['colour'] = colour_bench(max(10, args.steps // 2), 5, args.fp64)
    if args.batch:
        extra['batch'] = batch_bench(crops, args.batch_interval, args.batch_max_steps, args.batch_tune_only)
    if args.slabs:
        extra['slabs'] = slab_bench(labels, [int(r) for r in args.slabs.split(',')], args.steps, args.warmup)
    for k, v in extra.items():
        log(k, v)

    ms = main_run['ms_per_step']
    site_mlups = labels.size / ms / 1e3
    achieved = main_run['bytes_per_launch'] / ms / 1e6    # GB/s
    cpu = None
    if args.cpu_budget > 0:
        cpu = cpu_baseline(spec, labels, args.cpu_budget)
        log('cpu', cpu)
    line = {
        'metric': METRIC, 'value': round(site_mlups, 1), 'unit': 'MLUPS', 'n_gpus': 1,
        'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': round(ms, 5),
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'fp32',
        'data': 'synthetic',
        'config': {'workload': f'D3Q19 MRT single-phase, {args.size}^3 periodic sphere pack '
                               f'(r 4-10, porosity {porosity:.4f}, seed 2), tau 0.8, F 1e-6 x, '
                               'FP32 shifted populations, fluid sites stored and updated (sparse)',
                   'fluid_mlups': round(site_mlups * porosity, 1), 'porosity': porosity,
                   'kernel': main_run['kernel'], 'parallelism': 'single GPU'},
        'roofline': {'bound': 'hbm', 'achieved': round(achieved, 1), 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': round(achieved / HBM_PEAK_GBS, 4),
                     'traffic': measured_traffic(main_run['kernel'], args.size),
                     'bytes_per_launch': main_run['bytes_per_launch'], 'copy_kernel_gbs': round(copy_gbs, 1)},
        'cpu_baseline': cpu,
        'extra': extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
