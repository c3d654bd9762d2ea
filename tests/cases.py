"""Test cases: the reference's recorded study cases (from the fixtures that
tests/golden/make_golden.py writes) and the cases this test suite designs
itself (Zou-He boundaries, CSF, films, the 8x10x4 channel, sphere packs)."""
import functools
import json
from pathlib import Path

import numpy as np

from lbflow_amd import cases as CS

GOLDEN = Path(__file__).resolve().parent / 'golden'


@functools.lru_cache(maxsize=None)
def _fixture():
    return json.loads((GOLDEN / 'reference_cases.json').read_text())


def reference(family, **match):
    """The recorded cases of one family ('force', 'channel', 'layered',
    'bundle', 'entry', 'pressure', 'rock'), optionally filtered by key values
    (a tuple of values matches any of them)."""
    out = []
    for d in _fixture()[family]:
        if 'geometry_file' in d:
            raw = (GOLDEN / d['geometry_file']).read_bytes()
            case = {k: (tuple(v) if k in ('size', 'tau', 'density') else v) for k, v in d.items()}
            case['geometry'] = np.frombuffer(raw, np.uint8).copy()
        else:
            case = CS.from_json(d)
        if all(case.get(k) in (v if isinstance(v, tuple) else (v,)) for k, v in match.items()):
            out.append(case)
    return out


def reference_case(family, name):
    return next(c for c in reference(family) if c['name'] == name)


def names(cases):
    return [c['name'] for c in cases]


def grid_indices(nz, ny, nx):
    return np.indices((nz, ny, nx))


def c1():
    """8x10x4 D3Q19 SRT channel, walls at y = 0 and 9, tau 1, F 1e-6 along x,
    20 000 steps (the reference's examples/generate.py:15-53 with SRT)."""
    z, y, x = grid_indices(4, 10, 8)
    labels = np.where((y == 0) | (y == 9), 255, 75)
    return CS.make('c1', labels, [CS.label_line(75, 'GV', 1e-6, 1, 0), CS.label_line(255, 'SD', 90)],
                   eq='SRT', steps=20000, write=250, restart=500)


def zouhe_cases(nx=20, ny=10, nz=6, u0=1e-3, rho_in=1.0003):
    """Single-phase Zou-He cases, every axial direction and both signs:
    'zh-x': velocity inlet (VB, +x) at x = 0, pressure outlet (PB, -x) at
    x = nx - 1, walls at y = 0, ny - 1 and a 2-voxel obstacle block;
    'zh-t': pressure inlet along +z (2D: +y) at the first plane, velocity
    boundary along -z (2D: -y) at the last plane, walls at x = 0, nx - 1."""
    out = []
    for dims in (3, 2):
        z, y, x = grid_indices(nz if dims == 3 else 1, ny, nx)
        block = (x >= nx // 2) & (x < nx // 2 + 2) & (y >= ny // 2 - 1) & (y < ny // 2 + 1)
        along_x = np.where((y == 0) | (y == ny - 1) | block, 255,
                           np.where(x == 0, 50, np.where(x == nx - 1, 200, 75)))
        c = z if dims == 3 else y
        last = c.max()
        across = np.where((x == 0) | (x == nx - 1), 255, np.where(c == 0, 50, np.where(c == last, 200, 75)))
        din, dout = (5, 6) if dims == 3 else (2, 4)
        for eq in ('SRT', 'MRT'):
            out.append(CS.make(f'zh-x-{dims}d-{eq}', along_x,
                               [CS.label_line(50, 'VB', u0, 1, 0), CS.label_line(200, 'PB', 1, 3 if dims == 2 else 2, 0),
                                CS.label_line(75, 'VD', phase=0), CS.label_line(255, 'SD', 90)],
                               dimensions=dims, eq=eq, tau=(0.8,), steps=300, write=100))
            out.append(CS.make(f'zh-t-{dims}d-{eq}', across,
                               [CS.label_line(50, 'PB', rho_in, din, 0), CS.label_line(200, 'VB', u0, dout, 0),
                                CS.label_line(75, 'VD', phase=0), CS.label_line(255, 'BB')],
                               dimensions=dims, eq=eq, tau=(0.8,), steps=300, write=100))
    return out


def zouhe_color_cases(nx=20, ny=10, nz=6, u0=1e-3, rho_in=1.0003):
    """Colour-model Zou-He cases: phase 0 enters through a velocity inlet (VB,
    +x) at x = 0 into a channel held by phase 1 with a pressure outlet (PB,
    -x, phase 1) at x = nx - 1; and a pressure inlet (+z / 2D +y, phase 0)
    against a velocity boundary (-z / 2D -y, phase 1).  60-degree walls."""
    out = []
    for dims in (3, 2):
        z, y, x = grid_indices(nz if dims == 3 else 1, ny, nx)
        along_x = np.where((y == 0) | (y == ny - 1), 255,
                           np.where(x == 0, 50, np.where(x == nx - 1, 200, np.where(x < 5, 0, 100))))
        c = z if dims == 3 else y
        last = c.max()
        across = np.where((x == 0) | (x == nx - 1), 255,
                          np.where(c == 0, 50, np.where(c == last, 200, np.where(c < 3, 0, 100))))
        din, dout = (5, 6) if dims == 3 else (2, 4)
        fluid = [CS.label_line(0, 'VD', phase=0), CS.label_line(100, 'VD', phase=1), CS.label_line(255, 'SD', 60)]
        for eq in ('SRT', 'MRT'):
            out.append(CS.make(f'zhc-x-{dims}d-{eq}', along_x,
                               [CS.label_line(50, 'VB', u0, 1, 0), CS.label_line(200, 'PB', 1, 3 if dims == 2 else 2, 1)]
                               + fluid, model='color', dimensions=dims, eq=eq, tau=(1.0, 0.8), steps=300, write=100))
            out.append(CS.make(f'zhc-t-{dims}d-{eq}', across,
                               [CS.label_line(50, 'PB', rho_in, din, 0), CS.label_line(200, 'VB', u0, dout, 1)]
                               + fluid, model='color', dimensions=dims, eq=eq, tau=(1.0, 1.0), steps=300, write=100))
    return out


def csf_cases():
    """Experimental CSF capillary force (the reference's capillary_checks.py:23-45):
    a bulk planar interface (phase 0 at x < 16) and a uniform body force,
    32 x 4 (x 4) sites, D2Q9/D3Q19, SRT/MRT, 600 steps."""
    out = []
    for dims in (2, 3):
        z, y, x = grid_indices(1 if dims == 2 else 4, 4, 32)
        for eq in ('SRT', 'MRT'):
            common = dict(model='color', dimensions=dims, eq=eq, capillary_model='csf', steps=600, write=300)
            out.append(CS.make(f'csf-plane-d{dims}-{eq}', np.where(x < 16, 0, 100),
                               [CS.label_line(0, 'VD', phase=0), CS.label_line(100, 'VD', phase=1)], **common))
            out.append(CS.make(f'csf-force-d{dims}-{eq}', np.full(x.shape, 75),
                               [CS.label_line(75, 'GV', 1e-6, 1, 0)], **common))
    return out


def csf_drop(wall_gradient='tanh', curvature='normal'):
    """12^3 drop of radius 3 with CSF (the reference's capillary_checks.py:46-50)."""
    z, y, x = grid_indices(12, 12, 12)
    inside = (x - 5.5) ** 2 + (y - 5.5) ** 2 + (z - 5.5) ** 2 <= 9
    return CS.make('csf-drop-continuous', np.where(inside, 100, 0),
                   [CS.label_line(0, 'VD', phase=0), CS.label_line(100, 'VD', phase=1)],
                   model='color', steps=40, write=20, restart=20, capillary_model='csf',
                   wall_gradient=wall_gradient, curvature_stencil=curvature)


def csf_wall(wall_gradient='tanh'):
    """2D contact line between 60-degree walls with CSF (capillary_checks.py:57-62)."""
    z, y, x = grid_indices(1, 10, 16)
    labels = np.where((y == 0) | (y == 9), 255, np.where((x >= 4) & (x < 12), 0, 100))
    return CS.make('csf-wall-continuous', labels,
                   [CS.label_line(0, 'VD', phase=0), CS.label_line(100, 'VD', phase=1), CS.label_line(255, 'SD', 60)],
                   model='color', steps=100, write=50, capillary_model='csf', wall_gradient=wall_gradient)


def film_cases(nx=20, ny=8, nz=5, force=1e-5, transparent=False):
    """Half-permeable films (HF) in the colour model, in a channel between
    60-degree walls (y = 0, ny - 1), periodic along x:

    - film-block: phase-0 film planes at x = 2 and x = 12 enclose a phase-1
      slab (3 <= x <= 11) pushed along +x by a body force; a phase-1 film
      plane at x = 7 sits inside the slab; phase 0 elsewhere.
    - film-zh: a phase-0 velocity inlet (VB, +x) at x = 0 feeds through a
      phase-0 film at x = 1 into a channel held by phase 1 that leaves
      through a phase-1 film at x = nx - 2 and a phase-1 pressure outlet
      (PB, -x) at x = nx - 1.

    With transparent=True the film-block geometry holds phase 0 only (the
    phase-1 labels become phase 0), for the check that a film of the only
    phase present changes nothing."""
    out = []
    ph1 = 0 if transparent else 1
    for dims in (3, 2):
        z, y, x = grid_indices(nz if dims == 3 else 1, ny, nx)
        wall = (y == 0) | (y == ny - 1)
        block = np.where(wall, 255, np.where((x == 2) | (x == 12), 30, np.where(
            x == 7, 31, np.where((x >= 3) & (x <= 11), 101, 0))))
        zh = np.where(wall, 255, np.where(x == 0, 50, np.where(x == 1, 30, np.where(
            x == nx - 2, 31, np.where(x == nx - 1, 200, np.where(x < 5, 0, 100))))))
        for eq in ('SRT', 'MRT'):
            common = dict(model='color', dimensions=dims, eq=eq, tau=(1.0, 0.8), steps=300, write=100)
            out.append(CS.make(f'film-block-{dims}d-{eq}' + ('-transparent' if transparent else ''), block,
                               [CS.label_line(0, 'VD', phase=0), CS.label_line(30, 'HF', phase=0),
                                CS.label_line(31, 'HF', phase=ph1), CS.label_line(101, 'GV', force, 1, ph1),
                                CS.label_line(255, 'SD', 60)], **common))
            if not transparent:
                out.append(CS.make(f'film-zh-{dims}d-{eq}', zh,
                                   [CS.label_line(50, 'VB', 0.001, 1, 0), CS.label_line(30, 'HF', phase=0),
                                    CS.label_line(31, 'HF', phase=1),
                                    CS.label_line(200, 'PB', 1, 3 if dims == 2 else 2, 1),
                                    CS.label_line(0, 'VD', phase=0), CS.label_line(100, 'VD', phase=1),
                                    CS.label_line(255, 'SD', 60)], **common))
    return out


def sphere_case(n=(12, 14, 16), model='scsp', eq='MRT', tau=(0.8,), porosity=0.35, seed=3,
                angle=60, two_phase=False, dims=3):
    """Small periodic sphere pack with curved walls (3D) or discs (2D)."""
    from lbflow_amd import geometry as G
    shape = tuple(n) if dims == 3 else (1,) + tuple(n[-2:])
    labels = G.sphere_pack(shape, porosity, 2.0, 4.0, seed)
    if two_phase:
        labels = G.random_blocks(labels, 3, 0.5, seed + 100)
    indices = [CS.label_line(75, 'GV', 1e-5, 1, 0), CS.label_line(255, 'SD', angle)]
    if two_phase:
        indices.append(CS.label_line(76, 'GV', 1e-5, 1, 1))
    return CS.make(f'pack-{model}-{eq}-{dims}d', labels, indices, model=model, eq=eq, tau=tau,
                   dimensions=dims, sigma=0.01)

