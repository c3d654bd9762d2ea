"""The reference's recorded Bentheimer48 runs, replayed on the GPU.

The reference ran ten cases on the 48^3 Bentheimer sandstone crop with two
MPI ranks (research/drlab-publication/paper/data/porous-media.json,
rock_single and rock_two; cases as scripts/rock_workflow.py:18-39, :92-102):
single-phase D3Q19 MRT at three forces and two relaxation times, the two
pure-phase colour reductions, and three two-phase colour runs at 60 and 120
degrees on curved grain walls.  Its fields do not depend on the rank count
(scripts/ci_smoke.py:55-56), so a one-GPU run must reproduce every recorded
history value: Darcy velocity, porosity, total mass, density range, maximum
speed and Mach number, and for two phases the phase masses, saturations,
advective phase fluxes and the diffuse pore fraction, all computed from the
fields as scripts/porous_media.py:20-44 computes them from the VTI output.

These two-phase runs are the reference's only recorded output of the 3D
contact-angle rotation on curved walls.  FP64 runs must match every value
exactly; the FP32 engine's error against each recorded value is checked
against the tolerances in FP32_TOLERANCE.
"""
import math
import sys

import numpy as np
import pytest

from lbflow_amd import cases as CS
from lbflow_amd.engine import FP32, FP64, STEP_FIELDS, Engine
from tests import cases as TC
from tests.helpers import velocity_grid

pytestmark = pytest.mark.gpu
ROCK = TC.reference('rock')


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def pairwise_sum(a):
    """Sum of a contiguous float64 array in the order of numpy 1.x's reduction
    (pairwise: blocks of at most 128 elements summed in 8 interleaved
    partial sums, halves split at multiples of 8).  The recorded sums were
    taken in that order; numpy 2.x sums in another one and differs in the
    last bits."""
    def rec(lo, n):
        if n < 8:
            s = 0.0
            for v in vals[lo:lo + n]:
                s += v
            return s
        if n <= 128:
            r = vals[lo:lo + 8]
            i = 8
            while i < n - n % 8:
                r = [r[j] + vals[lo + i + j] for j in range(8)]
                i += 8
            s = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]))
            for v in vals[lo + i:lo + n]:
                s += v
            return s
        h = n // 2
        h -= h % 8
        return rec(lo, h) + rec(lo + h, n - h)
    vals = np.ascontiguousarray(a, dtype=np.float64).ravel().tolist()
    return 0.0 + rec(0, len(vals))


def site_mean(v):
    """Mean over sites of per-site vectors v (sites, k), summed site after
    site, as numpy reduces the leading axes of a (z, y, x, k) array."""
    return (np.cumsum(v, axis=0)[-1] / v.shape[0]).tolist()


def pore_measurement(r0, r1, u, case):
    """The history values of one output (porous_media.py:20-44 on the VTI
    fields: density / r0 + r1, velocity / v with components x, y, z)."""
    rho = r0 if r1 is None else r0 + r1
    velocity = velocity_grid(u, case).reshape(rho.size, -1)
    pore = rho > 0
    speed = np.sqrt((velocity[:, 0] * velocity[:, 0] + velocity[:, 1] * velocity[:, 1])
                    + velocity[:, 2] * velocity[:, 2])
    out = dict(darcy_velocity=site_mean(velocity), porosity=float(pore.mean()),
               total_mass=pairwise_sum(rho),
               density_range=[float(rho[pore].min()), float(rho[pore].max())],
               max_speed=float(speed.max()), max_mach=float(speed.max() * math.sqrt(3)))
    if r1 is not None:
        chi = np.divide(r0 - r1, rho, out=np.zeros_like(rho), where=rho != 0)
        phi = np.divide(r0, rho, out=np.zeros_like(rho), where=pore)
        fractions = (phi, pore.astype(float) - phi)
        count = int(pore.sum())
        out.update(phase_masses=[pairwise_sum(r0), pairwise_sum(r1)],
                   phase_saturations=[pairwise_sum(f[pore]) / count for f in fractions],
                   phase_advective_darcy_flux=[site_mean(f[:, None] * velocity) for f in fractions],
                   diffuse_pore_fraction=float((np.abs(chi[pore]) < .9).mean()))
    return out


def k_lattice(spec, darcy_x):
    """k = nu <u_x> / F with the viscosity of the flowing phase (rock_workflow.py:63-69)."""
    phase = CS.lattice(spec)['labels'][75][3] if spec['model'] == 'color' else 0
    return (spec['tau'][phase] - .5) / 3 * darcy_x / spec['force']


def replay(spec, history, precision):
    """Run the case through the recorded output steps; yield (record, measured, engine)."""
    case = CS.lattice(spec)
    eng = Engine(case, CS.geometry(spec), precision, stream_solids=precision == FP64).init()
    for h in history:
        while eng.t < h['step']:
            eng.step(STEP_FIELDS if eng.t + 1 == h['step'] else 0)
        eng.check_status()
        yield h, pore_measurement(*eng.fields(), case), eng
    eng.close()


FIELDS = {}


@pytest.mark.parametrize('spec', ROCK, ids=TC.names(ROCK))
def test_rock_history_fp64(spec, records):
    rec = records['rock'][spec['name'] + '-2rank']
    say('rock fp64', spec['name'], 'steps', rec['history'][-1]['step'])
    assert CS.lattice(spec)['labels'][255][1] == rec['contact_angle']
    for h, got, eng in replay(spec, rec['history'], FP64):
        for key, want in h.items():
            if key != 'step':
                assert got[key] == want, (h['step'], key, got[key], want)
        if h is rec['history'][-1]:
            r0, r1, u = eng.fields()
            FIELDS[spec['name']] = (r0 if r1 is None else r0 + r1, u)
    if rec['K_lattice'] is not None:
        assert k_lattice(spec, rec['history'][-1]['darcy_velocity'][0]) == rec['K_lattice']


@pytest.mark.parametrize('phase,tau', [(0, 0.6), (1, 1.4)])
def test_pure_phase_colour_reduces_to_single_phase(phase, tau, records):
    """On the rock, pure-phase colour flow equals single-phase flow (rock_workflow.py:124-142):
    the recorded differences of density and of SCSP velocity / density are reproduced."""
    want = next(c for c in records['rock_pure_phase'] if c['phase'] == phase)
    a, b = FIELDS.get(f'rock-scsp-tau{tau}'), FIELDS.get(f'rock-color-pure{phase}-tau{tau}')
    if a is None or b is None:
        pytest.skip('needs both rock runs of this module')
    rho_s, u_s = a
    rho_c, u_c = b
    assert float(np.max(np.abs(rho_s - rho_c))) == want['max_density_difference']
    normalized = u_s / np.where(rho_s > 0, rho_s, 1.0)[None, :]
    assert float(np.max(np.abs(normalized - u_c))) == want['max_velocity_difference']
    assert float(np.max(np.abs(u_s - u_c))) == want['raw_velocity_difference']
    assert max(want['max_density_difference'], want['max_velocity_difference']) <= 2e-11


# FP32 engine against the recorded FP64 values: the largest error of each
# quantity over the recorded steps, relative to its largest recorded
# magnitude (vectors: over all components).
# Single-phase and pure-phase runs track FP64 closely; the two-phase runs
# are transient interface motion, whose Darcy velocity and peak speed follow
# FP64 to a few per cent over 16 000 steps while the phase masses and
# saturations stay within a few 1e-6.
FP32_TOLERANCE = {'single': dict(darcy_velocity=5e-6, total_mass=1e-9, density_range=1e-9, max_speed=5e-6),
                  'two': dict(darcy_velocity=0.1, total_mass=2e-7, density_range=2e-4, max_speed=0.1,
                              phase_masses=5e-6, phase_saturations=5e-6, diffuse_pore_fraction=1e-2)}


@pytest.mark.parametrize('spec', ROCK, ids=TC.names(ROCK))
def test_rock_history_fp32(spec, records):
    rec = records['rock'][spec['name'] + '-2rank']
    tol = FP32_TOLERANCE['two' if spec['two_phase'] else 'single']
    diff = {k: 0.0 for k in tol}
    scale = {k: max(float(np.max(np.abs(h[k]))) for h in rec['history']) for k in tol}
    for h, got, _ in replay(spec, rec['history'], FP32):
        for key in tol:
            diff[key] = max(diff[key], float(np.max(np.abs(np.subtract(got[key], h[key])))))
    worst = {k: diff[k] / scale[k] for k in tol}
    say('rock fp32', spec['name'], {k: f'{v:.2e}' for k, v in worst.items()})
    for key, limit in tol.items():
        assert worst[key] <= limit, (key, worst[key], limit)
