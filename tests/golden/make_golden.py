#!/usr/bin/env python3
"""Extract the reference's own recorded results into small JSON fixtures.

Run in a container that has the reference mounted (default /root/reference):

    python tests/golden/make_golden.py [--reference /root/reference]

Only data is copied: recorded measurements (paper/data/*.json) and the
Bentheimer48 voxel crop (examples/bentheimer48/single-phase.raw, CC BY 4.0,
see examples/bentheimer48/DATA-LICENSE.txt in the reference).  No reference
source is imported or run.
"""
import argparse
import hashlib
import json
import shutil
from pathlib import Path

HERE = Path(__file__).resolve().parent


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
    rock = {}
    for m in porous['rock_single']['measurements']:
        if m['name'].startswith('rock-scsp-F'):
            rock[m['name']] = {'history': m['history'], 'K_lattice': m['K_lattice'],
                               'force': m['force'], 'tau': m['tau'],
                               'geometry_sha256': m['geometry_sha256']}
    out['rock_scsp'] = rock
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

    raw = ref / 'examples/bentheimer48/single-phase.raw'
    dst = HERE / 'bentheimer48_single_phase.raw'
    shutil.copyfile(raw, dst)
    print('wrote', HERE / 'reference_records.json', dst,
          hashlib.sha256(dst.read_bytes()).hexdigest())


if __name__ == '__main__':
    main()
