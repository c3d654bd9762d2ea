"""The drop-in LBFlow executable end to end on the GPU.

Checks the process contract of the reference (solver/apps/LBSim/lbsim.cpp,
scsp.cpp:46-114, colorfluid.cpp:109-212): outputs, checkpoints, restart,
failure contracts, report lines and the convergence decision, against the
reference's recorded digest and values and against the CPU oracle.
"""
import hashlib
import json
import math
import re
import struct
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from lbflow_amd import BIN_DIR
from lbflow_amd import cases as CS
from lbflow_amd.vti import read_vti
from tests import cases as TC
from tests.helpers import C1_SHA256, c1_spec, oracle_run, velocity_grid

pytestmark = pytest.mark.gpu
LBFLOW = BIN_DIR / 'LBFlow'


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def run(spec, folder, start=0, checkpoint=None, extra=()):
    prm = CS.write_input(spec, folder, start)
    if checkpoint:
        for suffix in ('.rst', '.rst.meta'):
            Path(folder / f'field.{start:010d}{suffix}').write_bytes(Path(str(checkpoint) + suffix).read_bytes())
    say('LBFlow', spec['name'], folder)
    p = subprocess.run([str(LBFLOW), str(prm), *extra], cwd=folder, capture_output=True, text=True, timeout=300)
    timing = [json.loads(l[len('LBFlowTiming '):]) for l in p.stdout.splitlines() if l.startswith('LBFlowTiming ')]
    return p.returncode, p.stdout, timing[0] if timing else None


def test_c1_process_and_checkpoint_digest(tmp_path):
    spec = c1_spec()
    code, out, timing = run(spec, tmp_path)
    assert code == 0, out[-2000:]
    for marker in ('Model constructing...Done!', 'Simulation started!', 'Initializing...Done!',
                   'Step 20000 - Time elapsed: ', 'Simulation finished!', 'Total Time elapsed: '):
        assert marker in out
    assert timing['completed_steps'] == 20000 and timing['ranks'] == 1
    assert timing['fluid_nodes'] == 256 and timing['total_voxels'] == 320
    rst = tmp_path / 'field.0000020000.rst'
    assert hashlib.sha256(rst.read_bytes()).hexdigest() == C1_SHA256
    meta = (tmp_path / 'field.0000020000.rst.meta').read_text()
    assert meta == ('DRLAB_OUTGOING_V1\n20000 1 3 19\n4 10 8 \nNEXT_STREAM 20001\n'
                    'INITIAL_COLLISION_INCLUDED 20001\nCONVERGENCE_HISTORY_RESET_ON_LOAD\n')
    assert len(sorted(tmp_path.glob('field.*.vti'))) == 20000 // 250 + 1
    # VTI fields equal the oracle's macro fields
    (nx, ny, nz), f = read_vti(tmp_path / 'field.0000020000.vti')
    o, case = oracle_run(spec, 20000)
    r0, _, u = o.fields()
    assert np.array_equal(f['density'][..., 0].ravel(), r0)
    assert np.array_equal(f['velocity'], velocity_grid(u, case))
    assert (tmp_path / 'field.0000020000.vti').stat().st_size == 10770


@pytest.mark.parametrize('spec', [s for s in TC.reference('force') if s['dimensions'] == 3], ids=lambda s: s['name'])
def test_force_cases_through_process(spec, tmp_path, records):
    code, out, timing = run(spec, tmp_path)
    assert code == 0, out[-2000:]
    assert timing['completed_steps'] == spec['steps']
    _, f = read_vti(tmp_path / f'field.{spec["steps"]:010d}.vti')
    v = f['v' if spec['model'] == 'color' else 'velocity']
    rec = records['force'][spec['name'] + '-1rank']
    assert float(np.max(np.abs(v[..., 0] - 1e-4))) == rec['absolute_error']


def drop_spec():
    """12^3 two-phase drop (scripts/ci_smoke.py:42-56)."""
    z, y, x = np.indices((12, 12, 12))
    inside = (x - 5.5) ** 2 + (y - 5.5) ** 2 + (z - 5.5) ** 2 <= 9
    return CS.make('drop-continuous', np.where(inside, 100, 0), [CS.label_line(0, 'VD', phase=0), CS.label_line(100, 'VD', phase=1)],
                   model='color', steps=40, write=20, restart=20)


def test_restart_is_bitwise_and_failure_contracts(tmp_path):
    spec = drop_spec()
    cont = tmp_path / 'continuous'
    code, out, timing = run(spec, cont)
    assert code == 0, out[-2000:]
    checkpoint = cont / 'field.0000000020'
    resumed = tmp_path / 'resumed'
    code, out, timing = run(spec, resumed, 20, checkpoint)
    assert code == 0, out[-2000:]
    assert timing['completed_steps'] == 20
    a = (cont / 'field.0000000040.vti').read_bytes()
    b = (resumed / 'field.0000000040.vti').read_bytes()
    assert a == b
    assert (cont / 'field.0000000040.rst').read_bytes() == (resumed / 'field.0000000040.rst').read_bytes()
    # failure contracts (ci_smoke.py:64-78)
    pops = Path(str(checkpoint) + '.rst').read_bytes()
    meta = Path(str(checkpoint) + '.rst.meta').read_text()
    cases = {'missing-stage': (pops, 'NO_VALID_STAGE\n', 'Missing or incompatible restart metadata'),
             'short-populations': (pops[:-8], meta, 'Restart population file has an invalid length.'),
             'nonfinite-population': (struct.pack('<d', float('nan')) + pops[8:], meta,
                                      'Non-finite or non-positive total fluid density.')}
    for name, (data, m, message) in cases.items():
        folder = tmp_path / name
        folder.mkdir()
        src = folder / 'src'
        Path(str(src) + '.rst').write_bytes(data)
        Path(str(src) + '.rst.meta').write_text(m)
        code, out, timing = run(spec, folder, 20, src)
        assert code != 0 and '0 Failed to excute: ' in out and message in out, (name, out[-500:])
        assert timing is None


@pytest.mark.parametrize('name', ['film-zh-3d-MRT', 'film-block-2d-SRT'])
def test_film_case_through_process(name, tmp_path):
    """Half-permeable films through the drop-in: the VTI fields equal the
    oracle's bit for bit, and a restart from the step-200 checkpoint writes
    the same step-300 files."""
    spec = dict(next(s for s in TC.film_cases() if s['name'] == name), restart=100)
    cont = tmp_path / 'continuous'
    code, out, timing = run(spec, cont)
    assert code == 0, out[-2000:]
    assert timing['completed_steps'] == spec['steps']
    _, f = read_vti(cont / f'field.{spec["steps"]:010d}.vti')
    o, case = oracle_run(spec, spec['steps'])
    r0, r1, u = o.fields()
    assert np.array_equal(f['r0'][..., 0].ravel(), r0)
    assert np.array_equal(f['r1'][..., 0].ravel(), r1)
    assert np.array_equal(f['v'], velocity_grid(u, case))
    resumed = tmp_path / 'resumed'
    code, out, timing = run(spec, resumed, 200, cont / 'field.0000000200')
    assert code == 0, out[-2000:]
    for suffix in ('.vti', '.rst'):
        last = f'field.{spec["steps"]:010d}{suffix}'
        assert (cont / last).read_bytes() == (resumed / last).read_bytes()


@pytest.mark.parametrize('name', ['drop', 'wall'])
def test_csf_through_process(name, tmp_path):
    """The reference's CSF replay cases (scripts/capillary_checks.py:46-65):
    a 12^3 drop and a 2D contact line between tanh walls run through the
    drop-in; a restart from the middle checkpoint writes the same final
    files; the drop's fields equal the oracle's bit for bit, the tanh-wall
    case's within 1e-12 (atanh/tanh differ by an ulp between the host and
    device libraries)."""
    spec = TC.csf_drop() if name == 'drop' else dict(TC.csf_wall(), restart=50)
    half = spec['steps'] // 2
    cont = tmp_path / 'continuous'
    code, out, timing = run(spec, cont)
    assert code == 0, out[-2000:]
    last = f'field.{spec["steps"]:010d}'
    _, f = read_vti(cont / f'{last}.vti')
    o, case = oracle_run(spec, spec['steps'])
    r0, r1, u = o.fields()
    got = np.concatenate([f['r0'].ravel(), f['r1'].ravel(), f['v'].ravel()])
    want = np.concatenate([r0, r1, velocity_grid(u, case).ravel()])
    if name == 'drop':
        assert np.array_equal(got, want)
    else:
        assert np.abs(got - want).max() <= 1e-12
    resumed = tmp_path / 'resumed'
    code, out, timing = run(spec, resumed, half, cont / f'field.{half:010d}')
    assert code == 0, out[-2000:]
    for suffix in ('.vti', '.rst'):
        assert (cont / (last + suffix)).read_bytes() == (resumed / (last + suffix)).read_bytes()


# Core grids (z, y, x) of an R-rank reference run on a 12^3 grid, by hand from
# GridMc::initCoreDistribution (gridmc.cpp:55-73): the primes of R, largest
# first, each divide the axis with the largest remaining extent (first of
# equals in z, y, x order).
CORE_GRIDS_12 = {2: (2, 1, 1), 3: (3, 1, 1), 4: (2, 2, 1), 5: (5, 1, 1), 6: (3, 2, 1),
                 7: (7, 1, 1), 8: (2, 2, 2), 12: (3, 2, 2)}


def test_output_ranks_vti_layout(tmp_path):
    """VTI files laid out as R-rank reference runs write them
    (filevti.cpp:39-172, gridmc.cpp:55-95): one piece per rank in rank order
    (core coordinates with x fastest), the reference parser's rules hold
    (every block where its offset says, pieces tile the grid once), the
    fields equal the one-piece file's, pieces along a split axis differ by
    at most one site with the larger ones last, and checkpoints do not
    depend on R."""
    spec = drop_spec()
    base = tmp_path / 'r1'
    code, out, _ = run(spec, base)
    assert code == 0, out[-2000:]
    _, ref = read_vti(base / 'field.0000000040.vti')
    for R, cores in CORE_GRIDS_12.items():
        folder = tmp_path / f'r{R}'
        code, out, _ = run(spec, folder, extra=(f'--output-ranks={R}',))
        assert code == 0, out[-2000:]
        pieces = []
        _, f = read_vti(folder / 'field.0000000040.vti', pieces)
        assert f.keys() == ref.keys() and all(np.array_equal(f[k], ref[k]) for k in f)
        assert len(pieces) == R
        # rank r sits at core coordinates CalcI2P(r, cores) (grid.h:153-161)
        starts = [sorted({p[2 * a] for p in pieces}) for a in (2, 1, 0)]
        assert tuple(len(v) for v in starts) == cores
        for r, p in enumerate(pieces):
            cz, cy, cx = r // (cores[1] * cores[2]), (r // cores[2]) % cores[1], r % cores[2]
            assert (p[4], p[2], p[0]) == (starts[0][cz], starts[1][cy], starts[2][cx])
        for a, c in zip((2, 1, 0), cores):
            sizes = [p[2 * a + 1] - p[2 * a] + 1 for p in pieces
                     if all(p[2 * b] == 0 for b in (2, 1, 0) if b != a)]
            assert len(sizes) == c and sum(sizes) == 12
            assert sizes == sorted(sizes) and sizes[-1] - sizes[0] <= 1
        # data blocks of each array appear in rank order
        root = (folder / 'field.0000000040.vti').read_bytes()
        offsets = [[int(m) for m in re.findall(rb'Name="' + n + rb'"[^>]*offset="(\d+)"', root)]
                   for n in (b'r0', b'r1', b'v')]
        assert all(o == sorted(o) and len(o) == R for o in offsets)
        assert (folder / 'field.0000000040.rst').read_bytes() == (base / 'field.0000000040.rst').read_bytes()
    # 2D: an 8 x 10 plane (x, y) in 3 ranks splits y as 3, 3, 4
    y = np.indices((1, 10, 8))[1]
    spec2 = CS.make('plane-ranks', np.where(y < 4, 100, 0), [CS.label_line(0, 'VD', phase=0), CS.label_line(100, 'VD', phase=1)], model='color', steps=20, write=20)
    code, out, _ = run(spec2, tmp_path / 'p1')
    assert code == 0, out[-2000:]
    code, out, _ = run(spec2, tmp_path / 'p3', extra=('--output-ranks=3',))
    assert code == 0, out[-2000:]
    pieces = []
    _, a = read_vti(tmp_path / 'p1' / 'field.0000000020.vti')
    _, b = read_vti(tmp_path / 'p3' / 'field.0000000020.vti', pieces)
    assert all(np.array_equal(a[k], b[k]) for k in a)
    assert pieces == [[0, 7, 0, 2, 0, 0], [0, 7, 3, 5, 0, 0], [0, 7, 6, 9, 0, 0]]


def report_lines(out):
    return [l for l in out.splitlines() if l.startswith(' - ')]


def test_reports_and_convergence_match_oracle(tmp_path):
    """Report values and the stop step follow SCSP::report / Simulator::updateConvergence,
    evaluated here with the oracle's fields (sums to rounding, decisions exactly)."""
    spec = TC.reference_case('channel', 'channel-scsp-d3-MRT-tau1.4-H8')
    spec = dict(spec, steps=3000, write=3000, report_interval=50, report_dir=1, stop_error=1e-4, restart=3000,
                name='converge')
    code, out, timing = run(spec, tmp_path)
    assert code == 0, out[-2000:]
    lines = report_lines(out)
    # oracle emulation of the report loop
    o, case = oracle_run(spec, 0)
    N, D = o.N, case['D']
    nx = spec['size'][0]
    first, last = 1, nx - 2
    prev, has_prev, stable, kold, t = None, False, 0, 0.0, 0
    nu = (spec['tau'][0] - 0.5) / 3
    kpar = (last - first) * nu / (spec['size'][1] * spec['size'][2]) * 1e12
    expected = []
    adj = o.fields()[2].ravel().copy()
    while t < spec['steps']:
        o.step(1)
        t += 1
        if t % 50:
            adj = o.fields()[2].ravel().copy()
            continue
        s = o.report_sums(2, first, last)
        dP = (s[3] / s[7] - s[5] / s[8]) / 3.0
        Q = s[0] / (last - first + 1)
        drop = abs(dP) > 1e-14
        K = kpar * Q / dP if drop else 0.0
        errK = abs(K - kold) / max(abs(K), 1e-30) if drop else 0.0
        kold = K
        expected.append(Q)
        u = o.fields()[2].ravel().copy()
        if has_prev:
            c = o.convergence_sums(prev, adj)
            rms = math.sqrt(c[0] / c[2])
            res = math.sqrt(c[0] / c[1]) if c[1] > 0 else 0.0
            arms = math.sqrt(c[3] / c[2])
            ares = math.sqrt(c[3] / c[1]) if c[1] > 0 else 0.0
            ok = (not drop or errK < 1e-4) and (rms <= 1e-12 or res < 1e-4) and (arms <= 1e-12 or ares < 1e-4)
            stable = stable + 1 if ok else 0
        has_prev = True
        prev = u
        adj = u
        if stable >= 3:
            break
    assert timing['completed_steps'] == t
    q_printed = [float(m.group(1)) for m in (re.search(r'Q = ([-+0-9.e]+)\(lu\^3/ts\)', l) for l in lines) if m]
    assert len(q_printed) == len(expected)
    np.testing.assert_allclose(q_printed, expected, rtol=1e-5)
    assert (tmp_path / f'field.{t:010d}.rst').exists()
    assert any('pressure-based permeability unavailable' in l for l in lines)


def test_fp32_sparse_engine_option(tmp_path):
    """`--fp32 --sparse` runs the FP32 shifted sparse engine; the Poiseuille profile stays
    within 1e-5 relative of the reference's recorded FP64 profile (the FP32 tolerance;
    3e-6 measured at H8, tau 0.6)."""
    spec = TC.reference_case('channel', 'channel-scsp-d3-MRT-tau0.6-H8')
    code, out, timing = run(spec, tmp_path, extra=('--fp32', '--sparse'))
    assert code == 0, out[-2000:]
    _, f = read_vti(tmp_path / f'field.{spec["steps"]:010d}.vti')
    prof = f['velocity'][..., 0].mean(axis=(0, 2))[1:-1]
    import json as _j
    want = np.array(_j.loads((Path(__file__).parent / 'golden' / 'reference_records.json').read_text())
                    ['channel'][spec['name'] + '-1rank']['measured_profile'])
    assert np.max(np.abs(prof - want) / np.abs(want)) < 1e-5


def test_pressure_slit_report_line(tmp_path, records):
    """scripts/porous_media.py:257-279 through the drop-in executable: the last
    ' - dP = ' and ' - K = ' report lines carry the values the reference printed."""
    spec = TC.reference('pressure')[0]
    code, out, timing = run(spec, tmp_path)
    assert code == 0, out[-2000:]
    dP = re.findall(r' - dP = ([\deE+.-]+)', out)[-1]
    K = re.findall(r' - K = ([\deE+.-]+)', out)[-1]
    rec = records['pressure_report']
    say('dP', dP, 'K', K)
    assert float(dP) == float(f"{rec['report_slice_pressure_drop']:.6e}")
    assert float(K) == float(f"{rec['reported_K_um2']:.6e}")
    assert timing['completed_steps'] == 6000
