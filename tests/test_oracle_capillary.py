"""Pin the oracle's experimental CSF capillary force and tanh wall gradient
(CPU only).

The reference records no CSF or tanh-wall output.  What pins them:

  - the reference's own CSF checks (scripts/capillary_checks.py:23-45): a
    bulk planar interface stays at rest (|u| <= 2e-14 at t = 0 and after 600
    steps) and the capillary force leaves a uniform body force's integration
    exact (|u_x - 600 F| <= 1e-12, transverse <= 1e-12), D2Q9/D3Q19, SRT/MRT,
    with both curvature stencils;
  - Laplace's law, which fixes the force's sign and size: a 2D drop of
    radius R relaxes to an inside-outside pressure jump of sigma / R (the
    planar checks cannot see a sign error, this one can);
  - the tanh wall gradient (colorfluid.cpp:494-518) reconstructs an
    equilibrium tanh profile exactly: at a flat 90-degree wall, for a colour
    field chi = tanh(c + k x) along the wall, the wall site's gradient is the
    full-stencil bulk gradient 3 sum_q w_q e_q tanh(c + e_q . k) of the
    profile continued into the wall, which the clipped legacy stencil misses.
"""
import numpy as np
import pytest

from lbflow_amd import cases as CS
from tests import cases as TC
from tests.helpers import oracle_run, velocity_grid
from tests.test_oracle_film import lattice_vectors

CASES = TC.csf_cases()


def speeds(o, case):
    _, _, u = o.fields()
    return np.linalg.norm(velocity_grid(u, case), axis=-1)


@pytest.mark.parametrize('curvature', ['normal', 'gradient'])
@pytest.mark.parametrize('spec', CASES, ids=lambda s: s['name'])
def test_csf_reference_checks(spec, curvature):
    spec = dict(spec, curvature_stencil=curvature)
    o, case = oracle_run(spec, 0)
    initial = speeds(o, case).max()
    o.step(spec['steps'])
    if spec['name'].startswith('csf-plane'):
        assert initial <= 2e-14 and speeds(o, case).max() <= 2e-14
    else:
        v = velocity_grid(o.fields()[2], case)
        assert np.max(np.abs(v[..., 0] - 6e-4)) <= 1e-12
        assert np.max(np.abs(v[..., 1:])) <= 1e-12


def test_csf_laplace_law():
    """Pressure jump of a 2D drop (radius 8, sigma 0.01) within 10 % of
    sigma / R, inside higher than outside."""
    n, R, sigma = 40, 8.0, 0.01
    _, y, x = np.indices((1, n, n))
    spec = CS.make('csf-laplace', np.where((x - 19.5) ** 2 + (y - 19.5) ** 2 <= R * R, 100, 0), [CS.label_line(0, 'VD', phase=0), CS.label_line(100, 'VD', phase=1)],
                   model='color', capillary_model='csf', sigma=sigma, steps=4000, write=4000)
    o, case = oracle_run(spec, 4000)
    r0, r1, u = o.fields()
    rho = (r0 + r1).reshape(n, n)
    y, x = np.mgrid[0:n, 0:n]
    r = np.hypot(x - 19.5, y - 19.5)
    dp = (rho[r < R - 3].mean() - rho[r > R + 4].mean()) / 3.0
    assert abs(dp - sigma / R) < 0.1 * sigma / R, dp
    assert np.abs(u).max() < 1e-4


def test_tanh_wall_reconstructs_a_tanh_profile():
    from oracle.pyoracle import Oracle
    nx, ny, nz = 10, 6, 5
    _, y, _ = np.indices((nz, ny, nx))
    spec = CS.make('tanh-flat-wall', np.where((y == 0) | (y == ny - 1), 255, 0),
                   [CS.label_line(0, 'VD', phase=0), CS.label_line(255, 'SD', 90)], model='color', steps=1, write=1)
    e = lattice_vectors(3)
    w = np.array([1 / 3] + [1 / 18 if (v ** 2).sum() == 1 else 1 / 36 for v in e[1:]])
    c0, k = 0.2, np.array([0.0, 0.0, 0.35])          # chi = tanh(c0 + k . (z, y, x - 4))
    grids = {}
    for wall in ('tanh', 'legacy'):
        case = CS.lattice(dict(spec, wall_gradient=wall))
        o = Oracle(case, CS.geometry(spec)).init()
        f = np.zeros((2, 19, nz, ny, nx))
        chi = np.tanh(c0 + k[2] * (np.arange(nx) - 4.0))
        f[0, 0] = (1 + chi) / 2
        f[1, 0] = (1 - chi) / 2
        o.set_populations(f.ravel(), 0)
        o.macro_only()
        grids[wall] = o
    for x in range(2, nx - 3):
        site = (2 * ny + 1) * nx + x                    # z = 2, y = 1 (next to the wall)
        rho0, rho1 = (1 + np.tanh(c0 + k[2] * (x - 4.0))) / 2, (1 - np.tanh(c0 + k[2] * (x - 4.0))) / 2
        c = np.arctanh((rho0 - rho1) / (rho0 + rho1))
        want = 3 * sum(w[q] * e[q] * np.tanh(c + e[q] @ k) for q in range(1, 19))
        got, norm = grids['tanh'].color_gradient(site)
        assert np.allclose(got, want, rtol=1e-10, atol=1e-13), (x, got, want)
        assert abs(norm - np.linalg.norm(want)) < 1e-12
        legacy, _ = grids['legacy'].color_gradient(site)
        assert abs(legacy[2] - want[2]) > 1e-3 * abs(want[2])
