"""The drop-in LBFlow executable's input contract, on the CPU.

Every case here fails during set-up, before any GPU call, with the reference's
message (solver/src/file/fileprm.cpp:136-295, solver/apps/LBSim/lbsim.cpp:66-70,
scripts/ci_smoke.py:58-63, scripts/benchmarks.py:257-260).
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from lbflow_amd import BIN_DIR
from lbflow_amd import cases as CS
from tests import cases as TC

LBFLOW = BIN_DIR / 'LBFlow'


def run_case(tmp_path, spec, start=0):
    prm = CS.write_input(spec, tmp_path, start)
    p = subprocess.run([str(LBFLOW), str(prm)], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    return p.returncode, p.stdout + p.stderr


def base():
    return TC.reference('force')[-1]


@pytest.mark.parametrize('overrides,message', [
    ({'writeInterval': 0}, 'Invalid time range or zero output/report/restart interval.'),
    ({'rexlationTime': '0.8 0.5'}, 'Densities must be positive; relaxation times must exceed 0.5.'),
    ({'rexlationTime': '.5 .5'}, 'Densities must be positive; relaxation times must exceed 0.5.'),
    ({'alpha0': 1.1}, 'Invalid phase weights, surface tension, or recoloring beta.'),
    ({'latticeSpeed': 2}, 'This research implementation requires latticeSpeed=1.'),
    ({'directions': 15}, 'Only D2Q9 and D3Q19 are validated for simulation.'),
    ({'fileType': '"uint16"'}, 'Simulation geometry must use uint8 labels.'),
    ({'stopError': 0}, 'Invalid convergence tolerance or lattice length.'),
    ({'equilModel': '"TRT"'}, 'Equilibrium model must be SRT or MRT.'),
    ({'wallGradient': '"bogus"'}, 'wallGradient must be legacy, renormalized, tanh, tangent, halfway or ghost.'),
    ({'viscosityInterpolation': '"geometric"'}, 'viscosityInterpolation must be arithmetic or harmonic.'),
    ({'reportDir': 25}, 'Report direction is outside the lattice.'),
    ({'reportDir': 1, 'reportSlice': '1'}, 'At least two report slices are required.'),
    ({'reportDir': 1, 'reportSlice': '1 40'}, 'Report slice is outside the grid.'),
    ({'reportDir': 3, 'reportSlice': '1 6'}, 'Report slice is outside the selected axis.'),
    ({'fluidModel': '"shanchen"'}, 'Only scsp and color fluid models are implemented.'),
    ({'fluidModel': '"lattice-gas"'}, 'Main(): Unsupport fluid model'),
])
def test_invalid_inputs_fail_with_reference_message(tmp_path, overrides, message):
    spec = dict(base(), overrides=overrides)
    code, out = run_case(tmp_path, spec)
    assert code != 0
    assert '0 Failed to excute: ' in out and message in out
    assert 'Simulation finished!' not in out


def test_undefined_label(tmp_path):
    spec = base()
    g = CS.grid(spec).copy()
    g[..., 0] = 9
    spec = dict(spec, geometry=g.ravel())
    code, out = run_case(tmp_path, spec)
    assert code != 0 and '0 Failed to excute: map::at' in out


def test_missing_input_file(tmp_path):
    p = subprocess.run([str(LBFLOW), str(tmp_path / 'absent.prm')], capture_output=True, text=True)
    assert p.returncode != 0 and 'FilePrm::Load() - Input file does not exist.' in p.stdout


def test_missing_argument():
    p = subprocess.run([str(LBFLOW)], capture_output=True, text=True)
    assert p.returncode == 1 and 'you must specify an input filename.' in p.stderr


def test_short_geometry(tmp_path):
    spec = base()
    prm = CS.write_input(spec, tmp_path)
    raw = tmp_path / 'geometry.raw'
    raw.write_bytes(raw.read_bytes()[:-1])
    p = subprocess.run([str(LBFLOW), str(prm)], cwd=tmp_path, capture_output=True, text=True)
    assert p.returncode != 0 and 'FileRaw::Load() - Failed to read from raw file' in p.stdout


@pytest.mark.parametrize('overrides,message', [
    ({'wallGradient': '"tangent"'}, 'wallGradient "tangent" is not implemented by this engine'),
    ({'recoloringModel': '"projection"'}, 'recoloringModel "projection" is not implemented by this engine'),
])
def test_unimplemented_d2q9_options(tmp_path, overrides, message):
    """The reference's remaining experimental colour options run on D2Q9 only
    and are rejected here with a message."""
    spec = [s for s in TC.reference('force') if s['model'] == 'color' and s['dimensions'] == 2][-1]
    code, out = run_case(tmp_path, dict(spec, overrides=overrides))
    assert code != 0 and message in out, out[-300:]


def test_missing_key(tmp_path):
    spec = base()
    prm = CS.write_input(spec, tmp_path)
    text = '\n'.join(l for l in prm.read_text().splitlines() if not l.startswith('endTime'))
    prm.write_text(text + '\n')
    p = subprocess.run([str(LBFLOW), str(prm)], cwd=tmp_path, capture_output=True, text=True)
    assert p.returncode != 0 and 'the value may not be defined in parameter file.' in p.stdout


def test_unsupported_boundary_type(tmp_path):
    spec = dict(base(), indices=['75 "AB" "N/A" "N/A" "N/A"'])
    code, out = run_case(tmp_path, spec)
    assert code != 0 and 'Axisymmetric AB boundary handling is not implemented.' in out


@pytest.mark.parametrize('ranks,message', [
    ('"0"', 'outputRanks must be a rank count of at least 1.'),
    ('"two"', 'outputRanks must be a rank count of at least 1.'),
    # 8 x 6 x 4 grid (x, y, z): 11 ranks cannot split any axis
    # (GridMc::initCoreDistribution, gridmc.cpp:55-73)
    ('"11"', 'too many processors'),
])
def test_output_ranks_validation(tmp_path, ranks, message):
    spec = dict(base(), overrides={'outputRanks': ranks})
    code, out = run_case(tmp_path, spec)
    assert code != 0 and message in out, out[-300:]


CSF_GUARD = CS.make('guard', np.zeros((1, 4, 4)), [CS.label_line(0, 'VD', phase=0)], model='color', steps=2, write=1,
                    capillary_model='csf')


@pytest.mark.parametrize('overrides,indices,message', [
    ({'capillaryModel': '"missing"'}, None, 'capillaryModel must be perturbation or csf.'),
    ({'wallGradient': '"missing"'}, None, 'wallGradient must be legacy, renormalized, tanh, tangent, halfway or ghost.'),
    ({'density': '1 2'}, None, 'Experimental CSF requires equal reference densities and standard pressure weights.'),
    ({'latticeSpeed': 2}, None, 'This research implementation requires latticeSpeed=1.'),
    ({'alpha0': .5}, None, 'Experimental CSF requires equal reference densities and standard pressure weights.'),
    ({}, ['0 "PB" "1" "1" "0"'], 'Pressure CSF requires flat D2Q9 ghost walls and phaseBoundary flux.'),
    ({}, ['0 "VB" "0" "1" "0"'], 'CSF requires closed flow or guarded flat ghost/flux pressure reservoirs.'),
    ({}, ['0 "HF" "N/A" "N/A" "0"'], 'CSF requires closed flow or guarded flat ghost/flux pressure reservoirs.'),
], ids=['invalid-capillary', 'invalid-wall', 'unsupported-density', 'unsupported-speed',
        'unsupported-pressure-weights', 'pressure-boundary', 'velocity-boundary', 'selective-film'])
def test_csf_failure_contracts(tmp_path, overrides, indices, message):
    """The reference's CSF failure cases (scripts/capillary_checks.py:66-82):
    each run must fail with the reference's message."""
    spec = dict(CSF_GUARD, overrides=overrides)
    if indices:
        spec['indices'] = indices
    code, out = run_case(tmp_path, spec)
    assert code != 0 and '0 Failed to excute: ' + message in out, out[-300:]


def test_film_needs_the_colour_model(tmp_path):
    """HF labels run in the colour model only: the reference's single-phase
    model never initialises or collides HF sites (scsp.cpp:116-175)."""
    scsp = [s for s in TC.reference('force') if s['model'] == 'scsp'][-1]
    spec = dict(scsp, indices=['75 "HF" "N/A" "N/A" "0"'])
    code, out = run_case(tmp_path, spec)
    assert code != 0 and 'half-permeable film (HF)' in out


def test_create_rejects_lattices_above_the_index_range():
    """More than 2^32 - 1 sites on one device is refused before any device
    call (site indices are 32-bit)."""
    import ctypes as C
    from lbflow_amd.engine import Desc, Label, load_library
    L = load_library()
    d = Desc()
    d.D, d.Q, d.model, d.equil, d.precision = 3, 19, 0, 1, 1
    d.sz[0], d.sz[1], d.sz[2] = 1024, 2048, 2048
    d.tau[0], d.density[0] = 0.8, 1.0
    labels = (Label * 256)()
    h = C.c_void_p()
    one_byte = C.create_string_buffer(1)   # never read: the size check comes first
    assert L.lbf_create(C.byref(d), labels, one_byte, C.byref(h)) != 0
    assert 'more than 2^32-1 sites' in L.lbf_last_error(None).decode()


def test_library_exports_every_declared_symbol():
    """Every function declared in include/lbflow.h is exported by liblbflow.so."""
    import re
    from lbflow_amd.engine import load_library
    header = (Path(__file__).resolve().parents[1] / 'include' / 'lbflow.h').read_text()
    names = set(re.findall(r'^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(lbf_[a-z_0-9]+)\s*\(', header, re.M))
    assert len(names) >= 20
    L = load_library()
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing
