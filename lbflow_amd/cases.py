"""Lattice cases: one description shared by the engine, the drop-in and tests.

A case is a plain dict, so that cases can be stored as JSON fixtures:

    name         str
    size         (nx, ny, nz), nz = 1 in 2D
    dimensions   2 or 3
    model        'scsp' (single phase) or 'color' (two-phase colour gradient)
    eq           'SRT' or 'MRT'
    tau, density one entry per phase
    alpha, sigma, beta
    indices      label lines exactly as input.prm holds them,
                 '<u8> "TYPE" "value" "dir" "phase"' (fileprm.cpp:563-575)
    geometry     uint8 labels, x fastest, then y, then z (fileraw.h:270-291)
    steps, write, restart and optional run keys (report_interval,
    report_dir, stop_error, wall_gradient, viscosity_interpolation,
    capillary_model, curvature_stencil, overrides)

Any other key (height, pressure_drop, ...) is carried along for whoever
built the case.  ``make`` builds a case from a label grid; ``lattice`` turns
it into the plain lattice description the engine and the CPU oracle take;
``write_input`` writes input.prm + geometry.raw for the drop-in executable.
"""
import base64
import zlib
from pathlib import Path

import numpy as np

KINDS = ('VD', 'GV', 'PB', 'VB', 'AB', 'BB', 'HF', 'SD')
# which of value / direction / phase each label kind reads (fileprm.cpp:633-670)
_HAS_VALUE = {'GV', 'PB', 'VB', 'SD'}
_HAS_DIR = {'GV', 'PB', 'VB'}
_HAS_PHASE = {'VD', 'GV', 'PB', 'VB', 'HF'}


def label_line(key, kind, value=None, direction=None, phase=None):
    """One input.prm label line; fields the kind does not read are written N/A."""
    if kind not in KINDS:
        raise ValueError(f'unknown label kind {kind!r}')

    def field(present, v, fmt):
        return format(v, fmt) if present and v is not None else 'N/A'
    parts = (kind, field(kind in _HAS_VALUE, value, '.17g'), field(kind in _HAS_DIR, direction, 'd'),
             field(kind in _HAS_PHASE, phase, 'd'))
    return f'{int(key)} ' + ' '.join(f'"{p}"' for p in parts)


def parse_label(line):
    """(key, (TYPE, value, dir, phase)) of one label line.  Fields the kind
    does not read are 0, and so is a field that is not a number, as the
    reference's stream extraction leaves it (fileprm.cpp:633-670)."""
    head, _, rest = line.strip().partition(' ')
    quoted = rest.split('"')[1::2]
    kind = quoted[0]

    def number(i, conv, used):
        if not used or i >= len(quoted):
            return conv(0)
        try:
            return conv(quoted[i])
        except ValueError:
            return conv(0)
    return int(head), (kind, number(1, float, kind in _HAS_VALUE), number(2, int, kind in _HAS_DIR),
                       number(3, int, kind in _HAS_PHASE))


def make(name, labels, indices, *, model='scsp', eq='MRT', tau=None, density=None, dimensions=None,
         alpha=None, sigma=0.005, beta=0.7, steps=1000, write=250, restart=None, **extra):
    """A case from a uint8 label grid of shape (nz, ny, nx) (2D: (ny, nx) or
    (1, ny, nx)).  Defaults: one relaxation time of 1 per phase, unit
    densities, alpha0 4/9 (2D) or 1/3 (3D), restart at the last step."""
    g = np.asarray(labels, dtype=np.uint8)
    if g.ndim == 2:
        g = g[None]
    if g.ndim != 3:
        raise ValueError('labels must be a 2D or 3D grid')
    nz, ny, nx = g.shape
    dims = dimensions or (2 if nz == 1 else 3)
    if dims == 2 and nz != 1:
        raise ValueError('a 2D case has one z-plane')
    phases = 2 if model == 'color' else 1
    case = dict(name=name, size=(nx, ny, nz), dimensions=dims, model=model, eq=eq,
                tau=tuple(tau) if tau is not None else (1.0,) * phases,
                density=tuple(density) if density is not None else (1.0,) * phases,
                alpha=alpha if alpha is not None else (4 / 9 if dims == 2 else 1 / 3),
                sigma=sigma, beta=beta, indices=list(indices), geometry=g.ravel(),
                steps=steps, write=write, restart=steps if restart is None else restart)
    case.update(extra)
    return case


def geometry(case):
    """uint8 labels, x fastest."""
    return np.ascontiguousarray(case['geometry'], dtype=np.uint8).ravel()


def grid(case):
    """The label grid as (nz, ny, nx)."""
    nx, ny, nz = case['size']
    return geometry(case).reshape(nz, ny, nx)


def lattice(case):
    """The plain lattice description the engine (engine.Engine) and the CPU
    oracle take: sizes slowest axis first, labels decoded."""
    nx, ny, nz = case['size']
    d = case['dimensions']
    return dict(D=d, Q=9 if d == 2 else 19, size=(ny, nx) if d == 2 else (nz, ny, nx),
                model=case['model'], eq=case['eq'], tau=tuple(case['tau']), density=tuple(case['density']),
                alpha=case['alpha'], sigma=case['sigma'], beta=case['beta'],
                harmonic=case.get('viscosity_interpolation', 'arithmetic') == 'harmonic',
                wall_gradient=case.get('wall_gradient', 'legacy'),
                capillary=case.get('capillary_model', 'perturbation'),
                curvature=case.get('curvature_stencil', 'normal'),
                labels=dict(parse_label(line) for line in case['indices']))


def _g17(v):
    return format(v, '.17g')


def input_prm(case, folder, start=0):
    """input.prm text for the case, its files under ``folder``.  Keys are the
    reference's input names (fileprm.h:55-90); ``overrides`` replaces or adds
    raw key/value pairs."""
    folder = Path(folder)
    nx, ny, nz = case['size']
    two_d = case['dimensions'] == 2
    keys = [
        ('gridFile', f'"{folder / "geometry.raw"}"'),
        ('dimension', f'{ny} {nx}' if two_d else f'{nz} {ny} {nx}'),
        ('directions', 9 if two_d else 19),
        ('fileType', '"uint8"'),
        ('latticeLen', 1),
        ('startTime', start),
        ('endTime', case['steps']),
        ('writeInterval', case['write']),
        ('reportInterval', case.get('report_interval', case['write'])),
        ('restartInterval', case['restart']),
        ('stopError', case.get('stop_error', 1e-8)),
        ('outPath', f'"{folder / "field"}"'),
        ('logPath', f'"{folder}"'),
        ('reportDir', case.get('report_dir', 0)),
        ('reportSlice', f'1 {nx - 2}'),
        ('fluidModel', f'"{case["model"]}"'),
        ('equilModel', f'"{case["eq"]}"'),
        ('rexlationTime', ' '.join(_g17(t) for t in case['tau'])),
        ('latticeSpeed', 1),
        ('alpha0', _g17(case['alpha'])),
        ('density', ' '.join(str(r) for r in case['density'])),
        ('contactAngle', 90),
        ('surfaceTension', case['sigma']),
        ('interfaceThickness', case['beta']),
        ('wallGradient', f'"{case.get("wall_gradient", "legacy")}"'),
        ('viscosityInterpolation', f'"{case.get("viscosity_interpolation", "arithmetic")}"'),
    ]
    for key, prm in (('capillary_model', 'capillaryModel'), ('curvature_stencil', 'curvatureStencil')):
        if key in case:
            keys.append((prm, f'"{case[key]}"'))
    table = dict(keys)
    table.update(case.get('overrides', {}))
    lines = [f'{k} {v};' for k, v in table.items()] + list(case['indices'])
    return '\n'.join(lines) + '\n'


def write_input(case, folder, start=0):
    """Write geometry.raw and input.prm into ``folder``; returns the prm path."""
    folder = Path(folder)
    folder.mkdir(parents=True, exist_ok=True)
    (folder / 'geometry.raw').write_bytes(geometry(case).tobytes())
    prm = folder / 'input.prm'
    prm.write_text(input_prm(case, folder, start))
    return prm


def to_json(case):
    """JSON-ready copy: the label grid as base64 of its zlib-compressed bytes."""
    out = {k: (list(v) if isinstance(v, tuple) else v) for k, v in case.items() if k != 'geometry'}
    out['geometry_z64'] = base64.b64encode(zlib.compress(geometry(case).tobytes(), 9)).decode()
    return out


def from_json(d):
    """Inverse of ``to_json``."""
    case = {k: (tuple(v) if k in ('size', 'tau', 'density') else v) for k, v in d.items()
            if k != 'geometry_z64'}
    raw = zlib.decompress(base64.b64decode(d['geometry_z64']))
    nx, ny, nz = case['size']
    if len(raw) != nx * ny * nz:
        raise ValueError(f'{case["name"]}: {len(raw)} label bytes for a {nx}x{ny}x{nz} grid')
    case['geometry'] = np.frombuffer(raw, dtype=np.uint8).copy()
    return case
