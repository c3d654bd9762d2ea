#!/usr/bin/env python3
"""Extract the reference's own recorded results, and the cases they were
recorded on, into small fixtures.

Run in a container that has the reference mounted (default /root/reference):

    python tests/golden/make_golden.py [--reference /root/reference]

Writes, next to this script:

- reference_records.json: recorded measurements (paper/data/*.json), as data;
- reference_cases.json: every recorded study case as a lattice case
  (lbflow_amd.cases layout: sizes, relaxation times, steps, label lines and
  the label grid).  The study parameters are restated below from the
  reference's study scripts, which are read as text and never imported or
  run; the file:line each family follows is given with it;
- bentheimer48_single_phase.raw, bentheimer48_two_phase.raw: the 48^3
  Bentheimer crops (examples/bentheimer48, CC BY 4.0, see
  bentheimer48_DATA-LICENSE.txt), checked against the manifest's SHA-256.
"""
import argparse
import hashlib
import json
import math
import shutil
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))
from lbflow_amd import cases as CS  # noqa: E402


def gv(key, force, phase=0, direction=1):
    return CS.label_line(key, 'GV', force, direction, phase)


def sd(key, angle):
    return CS.label_line(key, 'SD', angle)


def yz(n, ny, nx):
    """Index grids (z, y, x) of a (n, ny, nx) lattice."""
    return np.indices((n, ny, nx))


def round_up(v, quantum):
    return int(math.ceil(v / quantum) * quantum)


def force_family():
    """Uniform periodic body force, 8 x 6 (x 4), 100 steps (benchmarks.py:176-188)."""
    out = []
    for model in ('scsp', 'color'):
        phase = 1 if model == 'color' else 0
        for dims in (2, 3):
            for eq in ('SRT', 'MRT'):
                for tau in (.6, 1.4):
                    labels = np.full((1 if dims == 2 else 4, 6, 8), 75, np.uint8)
                    out.append(CS.make(f'force-{model}-d{dims}-{eq}-tau{tau}', labels,
                                       [gv(75, 1e-6, phase)], model=model, eq=eq, dimensions=dims,
                                       tau=(.8, tau) if model == 'color' else (tau,), steps=100, write=50))
    return out


def channel_family():
    """Poiseuille channels between halfway bounce-back walls at y = 0 and
    y = H + 1, run for 18 viscous decay times H^2/(nu pi^2) rounded up to
    500 steps, force 1e-6 (8/H)^3 (benchmarks.py:191-208)."""
    out = []
    for model, dims in (('scsp', 3), ('color', 2)):
        phase = 1 if model == 'color' else 0
        for eq in ('SRT', 'MRT'):
            for tau in (.6, 1.4):
                for h in (8, 16, 32):
                    steps = round_up(math.ceil(18 * h * h / ((tau - .5) / 3 * math.pi ** 2)), 500)
                    force = 1e-6 * (8 / h) ** 3
                    z, y, x = yz(1 if dims == 2 else 4, h + 2, 4)
                    labels = np.where((y == 0) | (y == h + 1), 255, 75).astype(np.uint8)
                    out.append(CS.make(f'channel-{model}-d{dims}-{eq}-tau{tau}-H{h}', labels,
                                       [gv(75, force, phase), sd(255, 90)], model=model, eq=eq,
                                       dimensions=dims, tau=(.8, tau) if model == 'color' else (tau,),
                                       steps=steps, write=steps // 2, force=force, height=h))
    return out


def layered_family():
    """Two-phase D2Q9 channels, phase 0 in the lower half, arithmetic and
    harmonic viscosity, 20 decay times of the thinner fluid rounded up to
    1000 steps, force 1e-6 (16/H)^3 (layered_flow.py:10-23)."""
    out = []
    for interp in ('arithmetic', 'harmonic'):
        for tau0, tau1 in ((.6, 1.4), (1.4, .6)):
            for h in (16, 32, 64):
                force = 1e-6 * (16 / h) ** 3
                steps = round_up(math.ceil(20 * h * h / ((min(tau0, tau1) - .5) / 3 * math.pi ** 2)), 1000)
                z, y, x = yz(1, h + 2, 4)
                labels = np.where((y == 0) | (y == h + 1), 255, np.where(y <= h // 2, 75, 76)).astype(np.uint8)
                name = f'layered-tau{tau0}-{tau1}-H{h}' + ('-harmonic' if interp == 'harmonic' else '')
                out.append(CS.make(name, labels, [gv(75, force, 0), gv(76, force, 1), sd(255, 90)],
                                   model='color', tau=(tau0, tau1), steps=steps, write=steps // 2,
                                   height=h, force=force, viscosity_interpolation=interp))
    return out


def bundle_family():
    """Two isolated slit pores of widths h and 2h (h = 8 s) in a 36 s high
    box, single phase or one phase per slit, 18 slowest decay times rounded
    up to 1000 steps, force 1e-6 / s^3 (porous_media.py:67-100)."""
    out = []
    for dims in (2, 3):
        for s in (1, 2, 3):
            h, ny = 8 * s, 36 * s
            force = 1e-6 / s ** 3
            for taus in ((1.,), (.6, 1.4), (1.4, .6)):
                model = 'scsp' if len(taus) == 1 else 'color'
                nu = [(t - .5) / 3 for t in taus]
                decay = max(w * w / (nu[i % len(taus)] * math.pi ** 2) for i, w in enumerate((h, 2 * h)))
                steps = round_up(18 * decay, 1000)
                z, y, x = yz(1 if dims == 2 else 4 * s, ny, 4 * s)
                first = (y >= 4 * s) & (y < 4 * s + h)
                second = (y >= 8 * s + h) & (y < 8 * s + 3 * h)
                labels = np.where(first, 75, np.where(second, 75 if model == 'scsp' else 76, 255)).astype(np.uint8)
                indices = [gv(75, force, 0)] + ([gv(76, force, 1)] if model == 'color' else []) + [sd(255, 90)]
                out.append(CS.make(f'bundle-d{dims}-s{s}-{model}-tau' + '-'.join(map(str, taus)), labels,
                                   indices, model=model, dimensions=dims, tau=taus, steps=steps,
                                   write=steps // 2, force=force, scale=s, widths=[h, 2 * h]))
    return out


def entry_family():
    """Capillary entry: phase 0 pushed by a Zou-He pressure inlet into a
    60-degree slit (2D) or square tube (3D) of width h = 16 held by phase 1,
    at 0.5 and 1.5 times the sharp entry pressure, legacy and renormalized
    wall gradients (porous_media.py:153-182; scale 1)."""
    out = []
    h, nx, sigma = 16, 65, .005
    for dims in (2, 3):
        pc = (2 if dims == 2 else 4) * sigma * math.cos(math.radians(60)) / h
        for wall in ('legacy', 'renormalized'):
            for ratio in (.5, 1.5):
                dp = ratio * pc
                z, y, x = yz(1 if dims == 2 else h + 2, h + 2, nx)
                walls = (y == 0) | (y == h + 1)
                if dims == 3:
                    walls |= (z == 0) | (z == h + 1)
                fluid = np.where(x == 0, 50, np.where(x == nx - 1, 200, np.where(x < h, 0, 100)))
                labels = np.where(walls, 255, fluid).astype(np.uint8)
                indices = [CS.label_line(50, 'PB', 1 + 3 * dp, 1, 0),
                           CS.label_line(200, 'PB', 1, 3 if dims == 2 else 2, 1),
                           CS.label_line(0, 'VD', phase=0), CS.label_line(100, 'VD', phase=1), sd(255, 60)]
                out.append(CS.make('entry-' + ('d3-' if dims == 3 else '') + f's1-{wall}-ratio{ratio}', labels,
                                   indices, model='color', dimensions=dims, tau=(1., 1.), sigma=sigma,
                                   steps=8000, write=2000, height=h, pressure_drop=dp, wall_gradient=wall))
    return out


def pressure_family():
    """Pressure-driven slit between Zou-He pressure boundaries, inlet x = 0 at
    rho 1 + 3 dp, outlet x = nx - 1 at rho 1, dp 1e-5, h = 16, report along x
    (porous_media.py:257-264)."""
    h, nx, nz, dp = 16, 65, 4, 1e-5
    z, y, x = yz(nz, h + 2, nx)
    labels = np.where((y == 0) | (y == h + 1), 255,
                      np.where(x == 0, 50, np.where(x == nx - 1, 200, 75))).astype(np.uint8)
    indices = [CS.label_line(50, 'PB', 1 + 3 * dp, 1, 0), CS.label_line(200, 'PB', 1, 2, 0),
               CS.label_line(75, 'VD', phase=0), sd(255, 90)]
    return [CS.make('pressure-channel-report', labels, indices, tau=(1.,), steps=6000, write=2000,
                    report_dir=1, stop_error=1e-12, overrides={'latticeLen': 1e-6}, height=h,
                    pressure_drop=dp)]


def rock_family(dataset):
    """Bentheimer48 runs (rock_workflow.py:18-39 for the case, :92-102 for the
    study): sigma 0.002, outputs every quarter of the run, periodic crop with
    halfway bounce-back on the grain (label 255, contact angle 60 or 120)."""
    manifest = json.loads((dataset / 'manifest.json').read_text())
    if manifest['refinement'] != 1:
        raise ValueError('only the unrefined crop is recorded')

    def grid(file):
        raw = (dataset / file).read_bytes()
        if hashlib.sha256(raw).hexdigest() != manifest['files'][file]['sha256']:
            raise ValueError(f'{file} differs from its manifest')
        nx, ny, nz = manifest['simulation_size_xyz']
        return np.frombuffer(raw, np.uint8).reshape(nz, ny, nx)

    single, two = grid('single-phase.raw'), grid('two-phase.raw')

    def rock(name, model='scsp', tau=(1.,), force=2e-6, phase=0, steps=4000, angle=60, two_phase=False):
        indices = [gv(75, force, phase)] + ([gv(76, force, 1)] if two_phase else []) + [sd(255, angle)]
        return CS.make(name, two if two_phase else single, indices, model=model, tau=tau, sigma=.002,
                       steps=steps, write=steps // 4, restart=steps // 2, force=force,
                       contact_angle=angle, two_phase=two_phase)

    out = [rock(f'rock-scsp-F{f}', force=f) for f in (1e-6, 2e-6, 4e-6)]
    for phase, tau in ((0, .6), (1, 1.4)):
        steps = 12000 if tau == .6 else 4000
        out.append(rock(f'rock-scsp-tau{tau}', tau=(tau,), steps=steps))
        out.append(rock(f'rock-color-pure{phase}-tau{tau}', model='color', tau=(.6, 1.4), phase=phase, steps=steps))
    out += [rock(f'rock-two-theta{a}-F{f}', model='color', tau=(1., 1.), force=f, angle=a, two_phase=True,
                 steps=16000) for a, f in ((60, 0.), (60, 2e-6), (120, 2e-6))]
    return out, single, two


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument('--reference', default='/root/reference')
    ref = Path(ap.parse_args().reference)
    data = ref / 'research/drlab-publication/paper/data'
    meas = json.loads((data / 'measurements.json').read_text())
    porous = json.loads((data / 'porous-media.json').read_text())

    out = {'source': 'zhou1800/lbflow research/drlab-publication/paper/data '
                     '(measurements.json, porous-media.json); recorded by the reference solver'}
    ver = meas['verification']['measurements']
    out['force'] = {m['name']: {k: m[k] for k in ('tau', 'absolute_error', 'transverse_max')}
                    for m in ver if m['kind'] == 'force'}
    out['channel'] = {m['name']: {k: m[k] for k in ('height', 'tau', 'relative_l2_error',
                                                    'relative_flow_error', 'measured_profile')}
                      for m in ver if m['kind'] == 'channel'}
    out['layered'] = {m['name']: {k: m[k] for k in ('tau', 'height', 'interpolation', 'profile',
                                                    'relative_l2_error')}
                      for m in meas['layered']}
    out['bundle'] = {m['name']: {'history': m['history'], 'measured_K': m.get('measured_K'),
                                 'phase_flux': m.get('phase_flux')}
                     for m in porous['bundle']['measurements']}
    out['rock_scsp'] = {m['name']: {'history': m['history'], 'K_lattice': m['K_lattice'],
                                    'force': m['force'], 'tau': m['tau'],
                                    'geometry_sha256': m['geometry_sha256']}
                        for m in porous['rock_single']['measurements'] if m['name'].startswith('rock-scsp-F')}
    # every Bentheimer48 record: single-phase, pure-phase colour and two-phase runs
    out['rock'] = {m['name']: {k: m.get(k) for k in ('model', 'tau', 'force', 'sigma', 'contact_angle',
                                                      'two_phase', 'history', 'K_lattice', 'geometry_sha256',
                                                      'max_relative_mass_drift')}
                   for key in ('rock_single', 'rock_two') for m in porous[key]['measurements']}
    out['rock_pure_phase'] = porous['rock_single']['pure_phase_comparisons']
    out['entry'] = {m['name']: {'history': [{k: h[k] for k in ('step', 'phase0_equivalent_length', 'pore_mean_ux',
                                                               'total_mass', 'phase_masses')}
                                            for h in m['history']],
                                'height': m['height'], 'pressure_drop': m['pressure_drop']}
                    for key in ('entry', 'entry3d') for m in porous[key]['measurements']
                    if m['wall_gradient'] in ('legacy', 'renormalized') and m['scale'] == 1}
    pr = porous['pressure_report']
    out['pressure_report'] = {k: pr[k] for k in ('name', 'prescribed_pressure_drop', 'report_slice_pressure_drop',
                                                 'expected_report_slice_pressure', 'reported_K_um2',
                                                 'expected_K_um2')}
    (HERE / 'reference_records.json').write_text(json.dumps(out, indent=1) + '\n')

    rocks, single, two = rock_family(ref / 'examples/bentheimer48')
    families = {'force': force_family(), 'channel': channel_family(), 'layered': layered_family(),
                'bundle': bundle_family(), 'entry': entry_family(), 'pressure': pressure_family()}
    for c in rocks:   # the big grids travel as their own .raw files
        c['geometry_file'] = 'bentheimer48_two_phase.raw' if c['two_phase'] else 'bentheimer48_single_phase.raw'
    cases = {k: [CS.to_json(c) for c in v] for k, v in families.items()}
    cases['rock'] = [{k: (list(v) if isinstance(v, tuple) else v) for k, v in c.items() if k != 'geometry'}
                     for c in rocks]
    (HERE / 'reference_cases.json').write_text(json.dumps(cases, indent=1) + '\n')
    for name, g in (('bentheimer48_single_phase.raw', single), ('bentheimer48_two_phase.raw', two)):
        (HERE / name).write_bytes(g.tobytes())
    shutil.copyfile(ref / 'examples/bentheimer48/DATA-LICENSE.txt', HERE / 'bentheimer48_DATA-LICENSE.txt')
    print('wrote', ', '.join(p.name for p in sorted(HERE.glob('*')) if p.suffix in ('.json', '.raw')))


if __name__ == '__main__':
    main()
