#!/usr/bin/env python3
"""D3Q19 MRT throughput, roofline and time to permeability on one MI355X.

Headline workload: single-phase D3Q19 MRT permeability lattice on a 256^3
periodic overlapping-sphere pack (radius uniform in [4, 10], porosity 0.20,
seed 2), body force 1e-6 along x, tau 0.8, rho0 1 (synthetic geometry from
lbflow_amd.geometry).  A "step" is one full lattice update (stream,
bounce-back, macro, force, MRT collision) of the whole domain with the
populations resident in HBM, FP32 shifted populations on the fluid sites.

value = fluid MLUPS, the reference's own throughput metric (fluid-site
updates per second / 1e6; scripts/performance.py:80-85).  The dominant
kernel's roofline uses its algorithmic HBM bytes per launch and its average
launch time, measured with HIP events on the engine's stream.

The same run also measures, in `extra`: the FP32 layouts that stream every
voxel, the FP64 kernels, time to permeability of the 256^3 pack in FP64 and
FP32 (residual 1e-6) and from the learned warm start, the two-phase colour
model on a 128x128x256 pack, 512 batched crops run to convergence, and the
warm start against the cold start on 256 unseen 64^3 crops.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick] [--legs dense,fp64,...]
"""
import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, public specification
METRIC = 'D3Q19 MRT MLUPS + % HBM roofline; time-to-permeability warm vs cold'
WEIGHTS = ROOT / 'lbflow_amd' / 'weights' / 'warm64.npz'
PMC_PROFILE = ROOT / 'profiles' / 'scsp_step_f32_sparse_256cubed_pmc.md'


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pack_case(name, labels):
    """Single-phase D3Q19 MRT, tau 0.8, F 1e-6 along x on pore label 75,
    halfway bounce-back on label 255."""
    from lbflow_amd import cases as CS
    return CS.make(name, labels, [CS.label_line(75, 'GV', 1e-6, 1, 0), CS.label_line(255, 'SD', 90)],
                   model='scsp', eq='MRT', tau=(0.8,))


def workload(size):
    from lbflow_amd import geometry as G
    labels = G.sphere_pack((size, size, size), 0.20, 4.0, 10.0, seed=2)
    return pack_case(f'c2-{size}', labels), labels


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
    eng = Engine(CS.lattice(spec), labels.ravel(), precision, stream_solids, dense).init()
    eng.advance(warmup)
    eng.synchronize()
    ms = eng.timed_advance(steps)
    eng.check_status()
    out = dict(kernel=eng.step_kernel, bytes_per_launch=eng.step_bytes, ms_per_step=ms / steps,
               device_bytes=eng.device_bytes, sites=eng.N)
    eng.close()
    return out


def summary(r, labels, fluid):
    return dict(kernel=r['kernel'], ms_per_step=round(r['ms_per_step'], 5),
                fluid_mlups=round(fluid / r['ms_per_step'] / 1e3, 1),
                site_mlups=round(labels.size / r['ms_per_step'] / 1e3, 1),
                bytes_per_launch=r['bytes_per_launch'],
                achieved_gbs=round(r['bytes_per_launch'] / r['ms_per_step'] / 1e6, 1))


def time_to_permeability(spec, labels, interval, max_steps):
    """The 256^3 pack in FP64 and FP32 from rest to the reference's stop rule
    (residual 1e-6, three stable reports): wall clock from init, steps, Darcy
    K, and |K32 - K64| / K64."""
    from lbflow_amd import cases as CS
    from lbflow_amd.engine import FP32, FP64, Engine
    from lbflow_amd.permeability import converge
    n = labels.shape[2]
    out = {}
    for name, prec in (('fp64', FP64), ('fp32', FP32)):
        t0 = time.perf_counter()
        eng = Engine(CS.lattice(spec), labels.ravel(), prec, False).init()
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
    crops = [np.ascontiguousarray(labels)]
    cold, t_cold, _ = S.run_crops(crops, 0.8, 1e-6, 1e-6, interval, max_steps)
    net, scaling = S.load(WEIGHTS, capacity=1, side=labels.shape[0])
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
    log('time-to-permeability warm vs cold (256^3)', out)
    return out


def colour_bench(steps, warmup, fp64=True):
    """Two-phase colour-gradient D3Q19 MRT on a 128x128x256 periodic sphere pack
    (z y x = 256 128 128, radius 6-12, porosity 0.30, seed 3), random 8^3 blocks
    of the two phases (50 %), tau 1/1, densities 1/1, alpha0 1/3, sigma 1e-3,
    beta 0.7, contact angle 60 deg, force 1e-6 along z (direction 5)."""
    from lbflow_amd import cases as CS
    from lbflow_amd import geometry as G
    from lbflow_amd.engine import FP32, FP64
    labels = G.sphere_pack((256, 128, 128), 0.30, 6.0, 12.0, seed=3)
    labels = G.random_blocks(labels, 8, 0.5, seed=103, phase0=75, phase1=76, pore=75)
    spec = CS.make('c3', labels, [CS.label_line(75, 'GV', 1e-6, 5, 0), CS.label_line(76, 'GV', 1e-6, 5, 1),
                                  CS.label_line(255, 'SD', 60)],
                   model='color', eq='MRT', tau=(1.0, 1.0), sigma=1e-3, beta=0.7, alpha=1 / 3)
    fluid = int(np.isin(labels, (75, 76)).sum())
    out = {'workload': 'colour-gradient D3Q19 MRT, 128x128x256 sphere pack (r 6-12, porosity '
                       f'{fluid / labels.size:.4f}, seed 3), 8^3 phase blocks 50 %, '
                       'tau 1/1, sigma 1e-3, beta 0.7, 60 deg, F 1e-6 z'}
    for name, prec, solids in (('fp32', FP32, False), ('fp64', FP64, True)):
        if name == 'fp64' and not fp64:
            continue
        r = measure(spec, labels, prec, solids, steps if name == 'fp32' else max(5, steps // 4),
                    warmup if name == 'fp32' else 2)
        out[name] = summary(r, labels, fluid)
        log('colour', name, out[name])
    return out


def batch_bench(crops, interval, max_steps):
    """Batched permeability: independent cubic sphere-pack crops (side 32-96,
    porosity 0.15-0.35, seeds 1000+i), D3Q19 MRT FP32, tau 0.8, F 1e-6 along x,
    every crop run to residual 1e-5 in one persistent hipGraph-captured launch
    sequence; the autotuner first picks the launch shape (threads per
    workgroup, slots per work unit) and the persistent grid."""
    from lbflow_amd.engine import Batch
    labels = {75: ('GV', 1e-6, 1, 0), 255: ('SD', 90.0, 0, 0)}
    b = Batch(3, 19, 'MRT', 0.8, 1.0, labels, crops)
    per_cu, tune_ms = b.tune(steps=10)
    threads, unit, grid = b.shape()
    log(f'batch: {b.sites} sites, {b.fluid} fluid, tuned {threads} threads x {unit}-slot units, '
        f'{per_cu} workgroups/CU, {tune_ms:.3f} ms/step')
    t0 = time.perf_counter()
    b.init()
    res, _ = b.run(axis=2, tol=1e-5, interval=interval, max_steps=max_steps)
    wall = time.perf_counter() - t0
    site_updates = sum(int(g.size) * r['steps'] for g, r in zip(crops, res))
    fluid_updates = sum(int((g == 75).sum()) * r['steps'] for g, r in zip(crops, res))
    steps = [r['steps'] for r in res]
    out = dict(crops=len(crops), converged=sum(r['converged'] for r in res), seconds=round(wall, 3),
               crops_per_s=round(len(crops) / wall, 3), fluid_mlups=round(fluid_updates / wall / 1e6, 1),
               site_mlups=round(site_updates / wall / 1e6, 1), sites=b.sites, fluid_sites=b.fluid,
               steps_min=min(steps), steps_max=max(steps), steps_mean=round(sum(steps) / len(steps), 1),
               report_interval=interval, tuned_workgroups_per_cu=per_cu, tuned_threads_per_workgroup=threads,
               tuned_slots_per_unit=unit, all_crops_ms_per_step=round(tune_ms, 4),
               workload=f'{len(crops)} cubic sphere-pack crops, side U{{32..96}}, porosity U[0.15,0.35], '
                        'seeds 1000+i, D3Q19 MRT FP32, tau 0.8, F 1e-6 x, residual 1e-5')
    b.close()
    log('batch', out)
    return out


def slab_bench(labels, steps, warmup, ways=(2, 4, 8)):
    """One GPU's share of a z-slab sharding of the pack: the FP32 sparse step
    on the first 1/k of the z-layers (periodic in z for the timing), with
    the fluid sites on a z-face (what a slab would exchange: 5 FP32
    populations per face site and direction)."""
    from lbflow_amd.engine import FP32
    out = {}
    for k in ways:
        sub = np.ascontiguousarray(labels[:labels.shape[0] // k])
        r = measure(pack_case(f'c2-slab{k}', sub), sub, FP32, False, steps, warmup)
        face = int((sub[0] == 75).sum())
        out[f'{k}way'] = dict(slab=list(sub.shape), ms_per_step=round(r['ms_per_step'], 5),
                              site_mlups=round(sub.size / r['ms_per_step'] / 1e3, 1),
                              fluid_mlups=round(int((sub == 75).sum()) / r['ms_per_step'] / 1e3, 1),
                              face_fluid_sites=face, exchange_bytes_per_face_and_direction=face * 5 * 4)
        log('slab', k, out[f'{k}way'])
    return out


def batch_tune_bench(crops, steps=20):
    """The batched step alone: the autotuner's sweep of launch shapes and
    grids, then `steps` eager launches of the picked shape (no graph, no
    convergence bookkeeping), for kernel traces of the step itself."""
    from lbflow_amd.engine import Batch
    labels = {75: ('GV', 1e-6, 1, 0), 255: ('SD', 90.0, 0, 0)}
    b = Batch(3, 19, 'MRT', 0.8, 1.0, labels, crops)
    per_cu, tune_ms = b.tune(steps=10)
    threads, unit, grid = b.shape()
    b.init()
    t0 = time.perf_counter()
    b.advance(steps)
    wall = time.perf_counter() - t0
    out = dict(crops=len(crops), fluid_sites=b.fluid, tuned_threads_per_workgroup=threads, tuned_slots_per_unit=unit,
               tuned_workgroups_per_cu=per_cu, tuned_ms_per_step=round(tune_ms, 4),
               eager_ms_per_step=round(wall / steps * 1e3, 4))
    b.close()
    log('batch tune', out)
    return out


def warm_start_bench(crops, interval=50, max_steps=60000):
    """Learned warm start on unseen 64^3 crops (seeds 1000000+i, porosity
    0.15-0.35): the batched FP32 engine from rest and from the network's
    prediction (+5 consistent-initialisation sweeps) to the unchanged stop
    rule at tolerance 1e-3; wall clock from init to the last stop, the
    network's inference included."""
    from lbflow_amd import surrogate as S
    net, scaling = S.load(WEIGHTS, capacity=8)
    cold, t_cold, _ = S.run_crops(crops, 0.8, 1e-6, 1e-3, interval, max_steps)
    warm, t_warm, t_inf = S.run_crops(crops, 0.8, 1e-6, 1e-3, interval, max_steps, net, scaling, relax=5)
    net.close()
    sc = np.array([r['steps'] for r in cold])
    sw = np.array([r['steps'] for r in warm])
    kc = np.array([r['K'] for r in cold])
    kw = np.array([r['K'] for r in warm])
    out = dict(crops=len(crops), tolerance=1e-3, report_interval=interval,
               cold=dict(seconds=round(t_cold, 3), steps_mean=round(float(sc.mean()), 1),
                         converged=int(sum(r['converged'] for r in cold))),
               warm=dict(seconds=round(t_warm, 3), steps_mean=round(float(sw.mean()), 1),
                         inference_and_relax_seconds=round(t_inf, 3),
                         converged=int(sum(r['converged'] for r in warm))),
               speedup_iterations=round(float(sc.sum() / sw.sum()), 3), speedup_wall=round(t_cold / t_warm, 3),
               k_rel_warm_vs_cold_mean=float(np.mean(np.abs(kw - kc) / kc)),
               k_rel_warm_vs_cold_max=float(np.max(np.abs(kw - kc) / kc)))
    log('warm start (64^3 crops)', out)
    return out


def cpu_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


CPU_CHILD = r'''
import json, sys, time
sys.path.insert(0, sys.argv[1])
from bench import workload
from lbflow_amd import cases as CS
from oracle.pyoracle import Oracle
spec, labels = workload(int(sys.argv[2]))
o = Oracle(CS.lattice(spec), labels.ravel()).init()
steps, reps = int(sys.argv[3]), int(sys.argv[4])
o.step(1)                                    # discarded warm-up step
times = []
for _ in range(reps):
    t0 = time.perf_counter()
    o.step(steps)
    times.append(time.perf_counter() - t0)
print(json.dumps(dict(times=times, fluid=int((labels == 75).sum()), sites=int(labels.size))))
'''


def cpu_baseline(size, cores, steps=2, reps=3):
    """The CPU restatement of the reference's step (oracle/lbf_oracle.c, the
    single-phase step loops on OpenMP threads, one pinned per core) on the
    same 256^3 lattice: one discarded warm-up step, then `reps` timed runs of
    `steps` steps (scripts/performance.py:61-118 protocol), median reported."""
    env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND='close', OMP_PLACES='cores')
    p = subprocess.run([sys.executable, '-c', CPU_CHILD, str(ROOT), str(size), str(steps), str(reps)],
                       env=env, capture_output=True, text=True, timeout=600)
    if p.returncode:
        raise RuntimeError(f'cpu baseline failed: {p.stderr[-2000:]}')
    r = json.loads(p.stdout.strip().splitlines()[-1])
    t = float(np.median(r['times']))
    return dict(value=round(r['fluid'] * steps / t / 1e6, 3), unit='fluid MLUPS', cores=cores, kind='port',
                site_mlups=round(r['sites'] * steps / t / 1e6, 3), seconds_per_step=round(t / steps, 4),
                sample=f'oracle/lbf_oracle.c (CPU restatement of the reference step, OpenMP, {cores} threads '
                       f'pinned one per core), same {size}^3 D3Q19 MRT lattice: 1 discarded warm-up step, then '
                       f'{reps} runs of {steps} steps, median {t:.2f} s per run')


# legs of the default run after the headline FP32 sparse step
LEGS = ('dense', 'fp64', 'ttp', 'ttpwarm', 'colour', 'batch', 'warm')
OPTIONAL_LEGS = ('batchtune', 'slabs')   # run only when named in --legs


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=200)
    ap.add_argument('--warmup', type=int, default=20)
    ap.add_argument('--size', type=int, default=256)
    ap.add_argument('--quick', action='store_true', help='only the headline step and its roofline')
    ap.add_argument('--cpu-cores', type=int, default=0, help='threads of the CPU baseline (0: every core '
                    'this process may use; -1: skip)')
    ap.add_argument('--ttp-interval', type=int, default=500)
    ap.add_argument('--max-steps', type=int, default=100000)
    ap.add_argument('--batch', type=int, default=512, help='crops of the batched leg (0: skip)')
    ap.add_argument('--warm-crops', type=int, default=256, help='unseen crops of the warm-start leg (0: skip)')
    ap.add_argument('--legs', default=','.join(LEGS),
                    help='comma-separated legs after the headline step (default: all of ' + ','.join(LEGS) +
                    '; also ' + ','.join(OPTIONAL_LEGS) + ')')
    args = ap.parse_args()
    legs = set(x for x in args.legs.split(',') if x)
    if legs - set(LEGS) - set(OPTIONAL_LEGS):
        ap.error(f'unknown legs {sorted(legs - set(LEGS) - set(OPTIONAL_LEGS))}')
    if args.gpus != 1:
        print(f'bench.py: the engine runs on one GPU; --gpus {args.gpus} is not supported '
              '(sharding across GPUs is not built)', file=sys.stderr)
        sys.exit(2)
    if args.quick:
        legs = set()

    # CPU-side generation first: the process pools fork before the GPU is used
    crops = warm_crops = None
    if legs & {'batch', 'batchtune'} and args.batch:
        from lbflow_amd import geometry as G
        t0 = time.perf_counter()
        crops, _ = G.batch_crops(args.batch, workers=16)
        log(f'batch: {args.batch} crops generated in {time.perf_counter() - t0:.1f} s')
    if 'warm' in legs and args.warm_crops and WEIGHTS.exists():
        from lbflow_amd import surrogate as S
        warm_crops = S.make_crops(args.warm_crops, S.EVAL_SEED0, 64)
    spec, labels = workload(args.size)
    fluid = int((labels == 75).sum())
    porosity = fluid / labels.size
    log(f'workload {args.size}^3 porosity {porosity:.4f}')

    from lbflow_amd.engine import FP32, FP64, copy_bandwidth
    copy_gbs = copy_bandwidth(1 << 30, 20)
    log(f'copy kernel {copy_gbs:.0f} GB/s')
    main_run = measure(spec, labels, FP32, False, args.steps, args.warmup)
    log('fp32 sparse', main_run)
    extra = {}
    if 'dense' in legs:
        extra['c2_fp32_dense_all_sites'] = summary(measure(spec, labels, FP32, True, args.steps, args.warmup),
                                                   labels, fluid)
        extra['c2_fp32_dense_fluid_sites'] = summary(
            measure(spec, labels, FP32, False, args.steps, args.warmup, dense=True), labels, fluid)
    if 'fp64' in legs:
        extra['c2_fp64_sparse'] = summary(measure(spec, labels, FP64, False, max(10, args.steps // 2), 5),
                                          labels, fluid)
        extra['c2_fp64_dense_all_sites'] = summary(measure(spec, labels, FP64, True, max(10, args.steps // 4), 5),
                                                   labels, fluid)
    if 'ttp' in legs:
        extra['c2_time_to_permeability'] = time_to_permeability(spec, labels, args.ttp_interval, args.max_steps)
    if 'ttpwarm' in legs and WEIGHTS.exists():
        extra['c2_time_to_permeability_warm_vs_cold'] = ttp_warm_vs_cold(labels, args.ttp_interval, args.max_steps)
    if 'colour' in legs:
        extra['c3_colour'] = colour_bench(max(10, args.steps // 2), 5)
    if crops and 'batch' in legs:
        extra['c4_batch'] = batch_bench(crops, 200, 60000)
    if 'slabs' in legs:
        extra['c2_slabs'] = slab_bench(labels, args.steps, args.warmup)
    if crops and 'batchtune' in legs:
        extra['c4_batch_tune'] = batch_tune_bench(crops)
    if warm_crops:
        extra['c5_warm_vs_cold'] = warm_start_bench(warm_crops)
    cpu = None
    cores = cpu_cores() if args.cpu_cores == 0 else args.cpu_cores
    if cores > 0:
        cpu = cpu_baseline(args.size, cores)
        log('cpu', cpu)

    ms = main_run['ms_per_step']
    fluid_mlups = fluid / ms / 1e3
    achieved = main_run['bytes_per_launch'] / ms / 1e6    # GB/s
    line = {
        'metric': METRIC, 'value': round(fluid_mlups, 1), 'unit': 'fluid MLUPS', 'n_gpus': 1,
        'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': round(ms, 5),
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'fp32',
        'data': 'synthetic',
        'config': {'workload': f'D3Q19 MRT single-phase, {args.size}^3 periodic sphere pack '
                               f'(r 4-10, porosity {porosity:.4f}, seed 2), tau 0.8, F 1e-6 x, '
                               'FP32 shifted populations, fluid sites stored and updated (sparse)',
                   'fluid_sites': fluid, 'porosity': porosity,
                   'site_equivalent_mlups': round(labels.size / ms / 1e3, 1),
                   'kernel': main_run['kernel'], 'parallelism': 'single GPU'},
        'roofline': {'bound': 'hbm', 'achieved': round(achieved, 1), 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': round(achieved / HBM_PEAK_GBS, 4),
                     'traffic': measured_traffic(main_run['kernel'], args.size),
                     'bytes_per_launch': main_run['bytes_per_launch'],
                     'frac_at_153_bytes_per_fluid_site': round(153.0 * fluid / ms / 1e6 / HBM_PEAK_GBS, 4),
                     'copy_kernel_gbs': round(copy_gbs, 1)},
        'cpu_baseline': cpu,
        'extra': extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
