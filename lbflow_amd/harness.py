"""Case construction, input files and VTI reading for LBFlow-compatible runs.

Mirrors the reference harness (scripts/benchmarks.py:26-149 ``spec``/``Study.run``,
scripts/vti.py:5-38 ``read_vti``) so that the same study cases can be written
as ``input.prm`` + ``geometry.raw`` for the drop-in ``LBFlow`` executable, or
handed to the C ABI as a plain lattice description.
"""
import math
import struct
import xml.etree.ElementTree as ET
from pathlib import Path

import numpy as np

LABEL_TYPES = ('VD', 'GV', 'PB', 'VB', 'AB', 'BB', 'HF', 'SD')


def spec(name, nx=8, ny=8, nz=4, dimensions=3, model='color', eq='MRT',
         tau=(1., 1.), steps=1000, write=250, label=None, indices=None, **kw):
    """A study case, same fields and defaults as scripts/benchmarks.py:26-34."""
    result = dict(name=name, size=(nx, ny, nz), dimensions=dimensions,
                  model=model, eq=eq, tau=tau, steps=steps, write=write,
                  restart=steps, density=(1., 1.) if model == 'color' else (1.,),
                  alpha=4 / 9 if dimensions == 2 else 1 / 3, sigma=.005, beta=.7,
                  label=label or (lambda x, y, z: 0),
                  indices=indices or ['0 "VD" "N/A" "N/A" "0"'])
    result.update(kw)
    return result


def force_cases():
    """Uniform periodic body force (scripts/benchmarks.py:176-188)."""
    cases = []
    for model in ('scsp', 'color'):
        for ndim in (2, 3):
            for eq in ('SRT', 'MRT'):
                for tau in (.6, 1.4):
                    name = f'force-{model}-d{ndim}-{eq}-tau{tau}'
                    cases.append(spec(name, ny=6, nz=1 if ndim == 2 else 4,
                                      dimensions=ndim, model=model, eq=eq,
                                      tau=(.8, tau) if model == 'color' else (tau,), steps=100, write=50,
                                      label=lambda x, y, z: 75,
                                      indices=[f'75 "GV" "0.000001" "1" "{1 if model == "color" else 0}"']))
    return cases


def channel_cases():
    """Poiseuille channels (scripts/benchmarks.py:191-208)."""
    cases = []
    for model, ndim in (('scsp', 3), ('color', 2)):
        for eq in ('SRT', 'MRT'):
            for tau in (.6, 1.4):
                for h in (8, 16, 32):
                    nu = (tau - .5) / 3
                    n = math.ceil(18 * h * h / (nu * math.pi ** 2))
                    steps = ((n + 499) // 500) * 500
                    force = 1e-6 * (8 / h) ** 3
                    cases.append(spec(
                        f'channel-{model}-d{ndim}-{eq}-tau{tau}-H{h}',
                        nx=4, ny=h + 2, nz=1 if ndim == 2 else 4, dimensions=ndim,
                        model=model, eq=eq, tau=(.8, tau) if model == 'color' else (tau,),
                        steps=steps, write=steps // 2, force=force, height=h,
                        label=lambda x, y, z, h=h: 255 if y in (0, h + 1) else 75,
                        indices=[f'75 "GV" "{force:.17g}" "1" "{1 if model == "color" else 0}"',
                                 '255 "SD" "90" "N/A" "N/A"']))
    return cases


def layered_cases(interpolation='arithmetic'):
    """Two-phase layered channels (scripts/layered_flow.py:10-23)."""
    runs = []
    for tau0, tau1 in ((.6, 1.4), (1.4, .6)):
        for h in (16, 32, 64):
            force = 1e-6 * (16 / h) ** 3
            steps = math.ceil(20 * h * h / ((min(tau0, tau1) - .5) / 3 * math.pi ** 2))
            steps = ((steps + 999) // 1000) * 1000
            runs.append(spec(
                f'layered-tau{tau0}-{tau1}-H{h}' + ('' if interpolation == 'arithmetic' else '-harmonic'),
                nx=4, ny=h + 2, nz=1, dimensions=2, tau=(tau0, tau1), steps=steps, write=steps // 2,
                height=h, force=force, viscosity_interpolation=interpolation,
                label=lambda x, y, z, h=h: 255 if y in (0, h + 1) else 75 if y <= h // 2 else 76,
                indices=[f'75 "GV" "{force:.17g}" "1" "0"', f'76 "GV" "{force:.17g}" "1" "1"',
                         '255 "SD" "90" "N/A" "N/A"']))
    return runs


def pressure_case(h=16, nx=65, nz=4, dp=1e-5):
    """Pressure-driven slit between Zou-He pressure boundaries
    (scripts/porous_media.py:257-264): inlet x = 0 at rho 1 + 3 dp along +x,
    outlet x = nx - 1 at rho 1 along -x, walls y = 0 and y = h + 1."""
    return spec('pressure-channel-report', nx=nx, ny=h + 2, nz=nz,
                model='scsp', tau=(1.,), steps=6000, write=2000, report_dir=1, stop_error=1e-12,
                label=lambda x, y, z: 255 if y in (0, h + 1) else 50 if x == 0 else 200 if x == nx - 1 else 75,
                overrides={'latticeLen': 1e-6}, height=h, pressure_drop=dp,
                indices=[f'50 "PB" "{1 + 3 * dp:.17g}" "1" "0"', '200 "PB" "1" "2" "0"',
                         '75 "VD" "N/A" "N/A" "0"', '255 "SD" "90" "N/A" "N/A"'])


def zouhe_cases(nx=20, ny=10, nz=6, u0=1e-3, rho_in=1.0003):
    """Small Zou-He boundary cases, every axial direction and both signs:
    'zh-x': velocity inlet (VB, +x) at x = 0, pressure outlet (PB, -x) at
    x = nx - 1, walls at y = 0, ny - 1 and a 2-voxel obstacle block;
    'zh-t': pressure inlet along +z (2D: +y) at the first plane, velocity
    boundary along -z (2D: -y) at the last plane, walls at x = 0, nx - 1."""
    cases = []
    for dims in (3, 2):
        zdim = nz if dims == 3 else 1
        for eq in ('SRT', 'MRT'):
            def lab_x(x, y, z):
                if y in (0, ny - 1) or (nx // 2 <= x < nx // 2 + 2 and ny // 2 - 1 <= y < ny // 2 + 1):
                    return 255
                return 50 if x == 0 else 200 if x == nx - 1 else 75

            cases.append(spec(f'zh-x-{dims}d-{eq}', nx=nx, ny=ny, nz=zdim, dimensions=dims, model='scsp',
                              eq=eq, tau=(0.8,), steps=300, write=100, label=lab_x,
                              indices=[f'50 "VB" "{u0:.17g}" "1" "0"', f'200 "PB" "1" "{3 if dims == 2 else 2}" "0"',
                                       '75 "VD" "N/A" "N/A" "0"', '255 "SD" "90" "N/A" "N/A"']))
            last = (nz if dims == 3 else ny) - 1
            din, dout = (5, 6) if dims == 3 else (2, 4)

            def lab_t(x, y, z, dims=dims, last=last):
                c = z if dims == 3 else y
                return 255 if x in (0, nx - 1) else 50 if c == 0 else 200 if c == last else 75

            cases.append(spec(f'zh-t-{dims}d-{eq}', nx=nx, ny=ny, nz=zdim, dimensions=dims, model='scsp',
                              eq=eq, tau=(0.8,), steps=300, write=100, label=lab_t,
                              indices=[f'50 "PB" "{rho_in:.17g}" "{din}" "0"',
                                       f'200 "VB" "{u0:.17g}" "{dout}" "0"',
                                       '75 "VD" "N/A" "N/A" "0"', '255 "BB" "N/A" "N/A" "N/A"']))
    return cases


def zouhe_color_cases(nx=20, ny=10, nz=6, u0=1e-3, rho_in=1.0003):
    """Colour-model Zou-He cases: phase 0 enters through a velocity inlet (VB,
    +x) at x = 0 into a channel held by phase 1 with a pressure outlet (PB,
    -x, phase 1) at x = nx - 1; and a pressure inlet (+z / 2D +y, phase 0)
    against a velocity boundary (-z / 2D -y, phase 1).  60-degree walls."""
    cases = []
    for dims in (3, 2):
        zdim = nz if dims == 3 else 1
        for eq in ('SRT', 'MRT'):
            def lab_x(x, y, z):
                if y in (0, ny - 1):
                    return 255
                return 50 if x == 0 else 200 if x == nx - 1 else 0 if x < 5 else 100

            cases.append(spec(f'zhc-x-{dims}d-{eq}', nx=nx, ny=ny, nz=zdim, dimensions=dims, model='color',
                              eq=eq, tau=(1.0, 0.8), steps=300, write=100, sigma=0.005, label=lab_x,
                              indices=[f'50 "VB" "{u0:.17g}" "1" "0"', f'200 "PB" "1" "{3 if dims == 2 else 2}" "1"',
                                       '0 "VD" "N/A" "N/A" "0"', '100 "VD" "N/A" "N/A" "1"',
                                       '255 "SD" "60" "N/A" "N/A"']))
            last = (nz if dims == 3 else ny) - 1
            din, dout = (5, 6) if dims == 3 else (2, 4)

            def lab_t(x, y, z, dims=dims, last=last):
                c = z if dims == 3 else y
                return (255 if x in (0, nx - 1) else 50 if c == 0 else 200 if c == last else
                        0 if c < 3 else 100)

            cases.append(spec(f'zhc-t-{dims}d-{eq}', nx=nx, ny=ny, nz=zdim, dimensions=dims, model='color',
                              eq=eq, tau=(1.0, 1.0), steps=300, write=100, sigma=0.005, label=lab_t,
                              indices=[f'50 "PB" "{rho_in:.17g}" "{din}" "0"',
                                       f'200 "VB" "{u0:.17g}" "{dout}" "1"',
                                       '0 "VD" "N/A" "N/A" "0"', '100 "VD" "N/A" "N/A" "1"',
                                       '255 "SD" "60" "N/A" "N/A"']))
    return cases


def csf_cases():
    """Experimental CSF capillary force (scripts/capillary_checks.py:23-45): a
    bulk planar interface (phase 0 at x < 16) and a uniform body force, 32 x 4
    (x 4) sites, D2Q9/D3Q19, SRT/MRT, 600 steps."""
    cases = []
    for dims in (2, 3):
        for eq in ('SRT', 'MRT'):
            common = dict(nx=32, ny=4, nz=1 if dims == 2 else 4, dimensions=dims, eq=eq,
                          capillary_model='csf', steps=600, write=300)
            cases.append(spec(f'csf-plane-d{dims}-{eq}', **common,
                              label=lambda x, y, z: 0 if x < 16 else 100,
                              indices=['0 "VD" "N/A" "N/A" "0"', '100 "VD" "N/A" "N/A" "1"']))
            cases.append(spec(f'csf-force-d{dims}-{eq}', **common, label=lambda x, y, z: 75,
                              indices=['75 "GV" "0.000001" "1" "0"']))
    return cases


def csf_drop(wall_gradient='tanh', curvature='normal'):
    """12^3 drop of radius 3 with CSF (scripts/capillary_checks.py:46-50)."""
    return spec('csf-drop-continuous', nx=12, ny=12, nz=12, steps=40, write=20, restart=20,
                capillary_model='csf', wall_gradient=wall_gradient, curvature_stencil=curvature,
                label=lambda x, y, z: 100 if (x - 5.5) ** 2 + (y - 5.5) ** 2 + (z - 5.5) ** 2 <= 9 else 0,
                indices=['0 "VD" "N/A" "N/A" "0"', '100 "VD" "N/A" "N/A" "1"'])


def csf_wall(wall_gradient='tanh'):
    """2D contact line between 60-degree walls with CSF (scripts/capillary_checks.py:57-62)."""
    return spec('csf-wall-continuous', nx=16, ny=10, nz=1, dimensions=2, steps=100, write=50,
                capillary_model='csf', wall_gradient=wall_gradient,
                label=lambda x, y, z: 255 if y in (0, 9) else (0 if 4 <= x < 12 else 100),
                indices=['0 "VD" "N/A" "N/A" "0"', '100 "VD" "N/A" "N/A" "1"', '255 "SD" "60" "N/A" "N/A"'])


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
    cases = []
    for dims in (3, 2):
        zdim = nz if dims == 3 else 1
        for eq in ('SRT', 'MRT'):
            def lab_block(x, y, z):
                if y in (0, ny - 1):
                    return 255
                if x in (2, 12):
                    return 30
                if x == 7:
                    return 31
                return 101 if 3 <= x <= 11 else 0

            ph1 = 0 if transparent else 1
            cases.append(spec(f'film-block-{dims}d-{eq}' + ('-transparent' if transparent else ''),
                              nx=nx, ny=ny, nz=zdim, dimensions=dims, model='color', eq=eq,
                              tau=(1.0, 0.8), steps=300, write=100, sigma=0.005, label=lab_block,
                              indices=['0 "VD" "N/A" "N/A" "0"', '30 "HF" "N/A" "N/A" "0"',
                                       f'31 "HF" "N/A" "N/A" "{ph1}"',
                                       f'101 "GV" "{force:.17g}" "1" "{ph1}"',
                                       '255 "SD" "60" "N/A" "N/A"']))
            if transparent:
                continue

            def lab_zh(x, y, z):
                if y in (0, ny - 1):
                    return 255
                return {0: 50, 1: 30, nx - 2: 31, nx - 1: 200}.get(x, 0 if x < 5 else 100)

            cases.append(spec(f'film-zh-{dims}d-{eq}', nx=nx, ny=ny, nz=zdim, dimensions=dims, model='color',
                              eq=eq, tau=(1.0, 0.8), steps=300, write=100, sigma=0.005, label=lab_zh,
                              indices=['50 "VB" "0.001" "1" "0"', '30 "HF" "N/A" "N/A" "0"',
                                       '31 "HF" "N/A" "N/A" "1"', f'200 "PB" "1" "{3 if dims == 2 else 2}" "1"',
                                       '0 "VD" "N/A" "N/A" "0"', '100 "VD" "N/A" "N/A" "1"',
                                       '255 "SD" "60" "N/A" "N/A"']))
    return cases


def entry_cases(scales=(1,), dimensions=(2,), walls=('legacy',)):
    """Capillary entry between Zou-He pressure boundaries in the colour model
    (scripts/porous_media.py:153-182): phase 0 enters a 60-degree slit (2D) or
    square tube (3D) held by phase 1, at 0.5 and 1.5 times the sharp entry
    pressure, with the given wall gradients (legacy, renormalized)."""
    cases = []
    for ndim in dimensions:
        for scale in scales:
            h, nx = 16 * scale, 64 * scale + 1
            sigma = .005 / scale
            pc = (2 if ndim == 2 else 4) * sigma * math.cos(math.radians(60)) / h
            for wall, ratio in [(w, r) for w in walls for r in (.5, 1.5)]:
                dp = ratio * pc

                def label(x, y, z, h=h, nx=nx, ndim=ndim):
                    return (255 if y in (0, h + 1) or (ndim == 3 and z in (0, h + 1)) else 50 if x == 0 else
                            200 if x == nx - 1 else 0 if x < h else 100)

                cases.append(spec(
                    'entry-' + ('d3-' if ndim == 3 else '') + f's{scale}-{wall}-ratio{ratio}',
                    nx=nx, ny=h + 2, nz=1 if ndim == 2 else h + 2, dimensions=ndim,
                    steps=8000 * scale ** 2, write=2000 * scale ** 2, tau=(1., 1.), sigma=sigma,
                    height=h, pressure_drop=dp, label=label, wall_gradient=wall,
                    indices=[f'50 "PB" "{1 + 3 * dp:.17g}" "1" "0"',
                             f'200 "PB" "1" "{3 if ndim == 2 else 2}" "1"', '0 "VD" "N/A" "N/A" "0"',
                             '100 "VD" "N/A" "N/A" "1"', '255 "SD" "60" "N/A" "N/A"']))
    return cases


def bundle_cases(scales=(1, 2, 3), dimensions=(2, 3)):
    """Two isolated slit pores (scripts/porous_media.py:67-100)."""
    cases = []
    for ndim in dimensions:
        for scale in scales:
            h, ny = 8 * scale, 36 * scale
            widths = (h, 2 * h)
            force = 1e-6 / scale ** 3
            for taus in ((1.,), (.6, 1.4), (1.4, .6)):
                model = 'scsp' if len(taus) == 1 else 'color'
                viscosities = [(tau - .5) / 3 for tau in taus]
                decay = max(widths[i] ** 2 / (viscosities[i % len(taus)] * math.pi ** 2) for i in (0, 1))
                steps = math.ceil(18 * decay / 1000) * 1000

                def label(x, y, z, h=h, scale=scale, model=model):
                    if 4 * scale <= y < 4 * scale + h:
                        return 75
                    if 8 * scale + h <= y < 8 * scale + 3 * h:
                        return 75 if model == 'scsp' else 76
                    return 255

                cases.append(spec(
                    f'bundle-d{ndim}-s{scale}-{model}-tau' + '-'.join(map(str, taus)),
                    nx=4 * scale, ny=ny, nz=1 if ndim == 2 else 4 * scale,
                    dimensions=ndim, model=model, tau=taus, steps=steps, write=steps // 2,
                    force=force, scale=scale, widths=widths, label=label,
                    indices=[f'75 "GV" "{force:.17g}" "1" "0"'] +
                            ([f'76 "GV" "{force:.17g}" "1" "1"'] if model == 'color' else []) +
                            ['255 "SD" "90" "N/A" "N/A"']))
    return cases


def geometry(s):
    """uint8 labels, x fastest then y then z (scripts/benchmarks.py:70-71)."""
    if s.get('geometry') is not None:
        return np.ascontiguousarray(s['geometry'], dtype=np.uint8).ravel()
    nx, ny, nz = s['size']
    return np.frombuffer(bytes(s['label'](x, y, z) for z in range(nz)
                               for y in range(ny) for x in range(nx)), dtype=np.uint8)


def _quoted(line):
    out, cur, inside = [], '', False
    for ch in line:
        if ch == '"':
            if inside:
                out.append(cur)
                cur = ''
            inside = not inside
        elif inside:
            cur += ch
    return out


def parse_index(line):
    """One label line ``<u8> "TYPE" "value" "dir" "phase"`` (fileprm.cpp:563-575, :633-670)."""
    key = int(line.split()[0])
    fields = _quoted(line)
    typ = fields[0]

    def num(text, conv):
        try:
            return conv(text)
        except ValueError:
            return conv(0)

    value = num(fields[1], float) if typ in ('GV', 'PB', 'VB', 'SD') else 0.0
    direction = num(fields[2], int) if typ in ('GV', 'PB', 'VB') else 0
    phase = num(fields[3], int) if typ in ('VD', 'GV', 'PB', 'VB', 'HF') else 0
    return key, (typ, value, direction, phase)


def lattice_case(s):
    """Plain lattice description of a study case (sizes in reference order)."""
    nx, ny, nz = s['size']
    ndim = s['dimensions']
    return dict(D=ndim, Q=9 if ndim == 2 else 19,
                size=(ny, nx) if ndim == 2 else (nz, ny, nx),
                model=s['model'], eq=s['eq'], tau=tuple(s['tau']), density=tuple(s['density']),
                alpha=s['alpha'], sigma=s['sigma'], beta=s['beta'],
                harmonic=s.get('viscosity_interpolation', 'arithmetic') == 'harmonic',
                wall_gradient=s.get('wall_gradient', 'legacy'),
                capillary=s.get('capillary_model', 'perturbation'),
                curvature=s.get('curvature_stencil', 'normal'),
                labels=dict(parse_index(line) for line in s['indices']))


def prm_text(s, folder, start=0):
    """input.prm text (scripts/benchmarks.py:72-104)."""
    folder = Path(folder)
    nx, ny, nz = s['size']
    dims = (ny, nx) if s['dimensions'] == 2 else (nz, ny, nx)
    params = dict(gridFile=f'"{folder / "geometry.raw"}"', dimension=' '.join(map(str, dims)),
                  directions=9 if s['dimensions'] == 2 else 19, fileType='"uint8"',
                  latticeLen=1, startTime=start, endTime=s['steps'],
                  writeInterval=s['write'], reportInterval=s.get('report_interval', s['write']),
                  restartInterval=s['restart'], stopError=s.get('stop_error', 1e-8),
                  outPath=f'"{folder / "field"}"', logPath=f'"{folder}"',
                  reportDir=s.get('report_dir', 0), reportSlice=f'1 {nx - 2}',
                  fluidModel=f'"{s["model"]}"', equilModel=f'"{s["eq"]}"',
                  rexlationTime=' '.join(format(t, '.17g') for t in s['tau']),
                  latticeSpeed=1, alpha0=format(s['alpha'], '.17g'),
                  density=' '.join(map(str, s['density'])), contactAngle=90,
                  surfaceTension=s['sigma'], interfaceThickness=s['beta'],
                  wallGradient=f'"{s.get("wall_gradient", "legacy")}"')
    params['viscosityInterpolation'] = f'"{s.get("viscosity_interpolation", "arithmetic")}"'
    if 'capillary_model' in s:
        params['capillaryModel'] = f'"{s["capillary_model"]}"'
    if 'curvature_stencil' in s:
        params['curvatureStencil'] = f'"{s["curvature_stencil"]}"'
    params.update(s.get('overrides', {}))
    text = '\n'.join(f'{key} {val};' for key, val in params.items())
    return text + '\n' + '\n'.join(s['indices']) + '\n'


def write_case(s, folder, start=0):
    folder = Path(folder)
    folder.mkdir(parents=True, exist_ok=True)
    (folder / 'geometry.raw').write_bytes(geometry(s).tobytes())
    (folder / 'input.prm').write_text(prm_text(s, folder, start))
    return folder / 'input.prm'


def read_vti(path, pieces=None):
    """Parse an appended-raw VTI with one or more pieces (scripts/vti.py:5-38):
    every DataArray offset must point at a block of exactly its piece's
    length and the pieces must cover the grid exactly once.  Returns
    ((nx, ny, nz), {name: array [nz, ny, nx, n]}); the piece extents
    (x0 x1 y0 y1 z0 z1 each) are appended to the list `pieces` if given."""
    data = Path(path).read_bytes()
    pos = data.index(b'<AppendedData')
    doc = ET.fromstring(data[:pos] + b'</VTKFile>')
    begin = data.index(b'>_', pos) + 2
    im = doc.find('ImageData')
    extent = list(map(int, im.attrib['WholeExtent'].split()))
    nx, ny, nz = [extent[i + 1] - extent[i] + 1 for i in (0, 2, 4)]
    fields = {}
    coverage = np.zeros((nz, ny, nx), np.int64)
    for piece in im.findall('Piece'):
        e = list(map(int, piece.attrib['Extent'].split()))
        if pieces is not None:
            pieces.append(e)
        box = (slice(e[4], e[5] + 1), slice(e[2], e[3] + 1), slice(e[0], e[1] + 1))
        shape = (e[5] - e[4] + 1, e[3] - e[2] + 1, e[1] - e[0] + 1)
        coverage[box] += 1
        for a in piece.find('PointData').findall('DataArray'):
            name = a.attrib['Name']
            n = int(a.attrib.get('NumberOfComponents', '1'))
            off = begin + int(a.attrib['offset'])
            length = struct.unpack_from('<Q', data, off)[0]
            expected = shape[0] * shape[1] * shape[2] * n * 8
            if length != expected:
                raise ValueError(f'{name}: block {length} != {expected}')
            values = np.frombuffer(data, dtype='<f8', count=length // 8, offset=off + 8)
            field = fields.setdefault(name, np.full((nz, ny, nx, n), np.nan))
            field[box] = values.reshape(shape + (n,))
    if (coverage != 1).any():
        raise ValueError('Pieces do not cover whole grid exactly once')
    if any(np.isnan(f).any() for f in fields.values()):
        raise ValueError('Missing field values')
    return (nx, ny, nz), fields
