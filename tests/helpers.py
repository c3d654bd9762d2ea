"""Shared helpers of the test suite: oracle runs and VTI-like views of fields."""
import numpy as np

from lbflow_amd import cases as CS
from tests import cases as TC

C1_SHA256 = '03d845deefdaf322bbffd2db52e60aa5aba43938e20aa83d3461237df07ed664'


def c1_spec():
    """The 8x10x4 SRT channel, 20 000 steps."""
    return TC.c1()


def c1_profile_error(got, tau=1.0, force=1e-6):
    """Largest deviation of the 8x10x4 channel's u_x(y), y = 1..8, from the
    closed form of body-force Poiseuille flow between halfway bounce-back
    walls at y = 1/2 and 17/2 with the BGK slip, relative to the peak:
    u(y) = F (y - 1/2)(17/2 - y) / (2 nu) + F (16 L - 3) / (24 nu), L = (tau - 1/2)^2."""
    nu, lam = (tau - 0.5) / 3, (tau - 0.5) ** 2
    y = np.arange(1, 9, dtype=float)
    want = force * (y - 0.5) * (8.5 - y) / (2 * nu) + force * (16 * lam - 3) / (24 * nu)
    return float(np.max(np.abs(np.asarray(got) - want)) / np.max(want))


def velocity_grid(u, case):
    """u[d][site] (reference component order) -> array (nz, ny, nx, D) with components x, y(, z),
    the VTI 'velocity'/'v' layout (filevti.cpp:153-167)."""
    D = case['D']
    size = case['size']
    shape = (1,) + tuple(size) if D == 2 else tuple(size)
    return np.stack([u[D - 1 - k].reshape(shape) for k in range(D)], axis=-1)


def profile(u, case):
    """Mean of u_x over x and z, solid rows removed (scripts/benchmarks.py:230)."""
    return velocity_grid(u, case)[..., 0].mean(axis=(0, 2))[1:-1]


def darcy(u, case):
    return velocity_grid(u, case).mean(axis=(0, 1, 2)).tolist()


def oracle_run(spec, steps, geometry=None):
    from oracle.pyoracle import Oracle
    case = CS.lattice(spec)
    o = Oracle(case, CS.geometry(spec) if geometry is None else geometry).init()
    if steps:
        o.step(steps)
    return o, case


sphere_case = TC.sphere_case


def lattice_shifts(D):
    """The neighbour vectors of D3Q19 (D2Q9) in (z, y, x) order: components in
    {-1, 0, 1}, one or two of them non-zero."""
    import itertools
    if D == 2:
        return [(0,) + v for v in itertools.product((-1, 0, 1), repeat=2) if any(v)]
    return [v for v in itertools.product((-1, 0, 1), repeat=3) if 0 < sum(map(abs, v)) <= 2]


def solid_shift_map(before, after, geometry, case, solid=255):
    """Streaming identity of solid sites (simulator.cpp:73-101: every voxel is
    streamed, bounce-back only writes fluid sites): after one step the rest
    population of a solid site is unchanged and its population q is the
    previous population q of its upstream site.  Returns, for q = 1..Q-1, the
    neighbour vectors under which plane q of `after` equals plane q of
    `before` shifted, on all solid sites."""
    Q, D = case['Q'], case['D']
    shape = (1,) + tuple(case['size']) if D == 2 else tuple(case['size'])
    mask = (np.asarray(geometry) == solid).reshape(shape)
    a = np.asarray(before).reshape((Q,) + shape)
    b = np.asarray(after).reshape((Q,) + shape)
    assert np.array_equal(a[0][mask].view(np.uint64), b[0][mask].view(np.uint64))
    out = []
    for q in range(1, Q):
        out.append([v for v in lattice_shifts(D)
                    if np.array_equal(np.roll(a[q], v, axis=(0, 1, 2))[mask].view(np.uint64),
                                      b[q][mask].view(np.uint64))])
    return out
