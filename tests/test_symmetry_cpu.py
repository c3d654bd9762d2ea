"""The lattice symmetries used to augment the warm-start training data (CPU).

`lbflow_amd.surrogate.symmetry` maps a fluid mask and its normalised steady
field.  For flips of z and y and the y <-> z swap the D3Q19 MRT lattice with
halfway bounce-back is exactly invariant, so running the oracle on the
mapped geometry must give the mapped field (to summation-order rounding);
the x flip with the transverse velocity and the density deviation negated
is a symmetry of the linear (Stokes) limit only, so it holds to the size of
the quadratic equilibrium terms (1e-4 relative measured at F = 1e-5).
"""
import numpy as np
import pytest

from lbflow_amd import geometry as G
from lbflow_amd import cases as CS
from lbflow_amd.surrogate import symmetry
from oracle.pyoracle import Oracle

N, STEPS = 12, 120


def run(labels):
    s = CS.make('sym', labels, [CS.label_line(75, 'GV', 1e-5, 1, 0), CS.label_line(255, 'SD', 90)], tau=(0.8,))
    o = Oracle(CS.lattice(s), CS.geometry(s)).init()
    o.step(STEPS)
    r, _, u = o.fields()
    # (u_z, u_y, u_x, rho - 1) per site, unnormalised (rms 1 per channel),
    # 0 on solid sites like the training targets
    fluid = (labels.ravel() == 75)[:, None]
    return np.stack([u[0], u[1], u[2], r - 1.0], axis=-1) * fluid


@pytest.fixture(scope='module')
def base():
    labels = G.sphere_pack((N, N, N), 0.45, 2.0, 3.5, seed=11)
    return labels, run(labels)


@pytest.mark.parametrize('k,tol', [(1, 1e-10), (2, 1e-10), (4, 1e-10), (7, 1e-10), (8, 2e-3), (15, 2e-3)])
def test_symmetry_maps_solutions_to_solutions(base, k, tol):
    labels, field = base
    # masks as the training pipeline holds them: [n, S, S, S]
    mask = (labels == 75).reshape(1, -1).astype(np.uint8)
    tgt = field.reshape(1, -1, 4)
    m2, t2 = symmetry(mask, tgt, k, np.ones(4))
    lab2 = np.where(m2.reshape(N, N, N) == 1, 75, 255).astype(np.uint8)
    want = run(lab2)
    got = t2.reshape(-1, 4)
    err = np.abs(got - want).max(axis=0) / np.abs(want).max(axis=0)
    assert np.all(err < tol), (k, err)


def test_density_parity_is_pinned(base):
    """The x flip must negate the density deviation: keeping its sign is caught."""
    labels, field = base
    mask = (labels == 75).reshape(1, -1).astype(np.uint8)
    m2, t2 = symmetry(mask, field.reshape(1, -1, 4), 8, np.ones(4))
    lab2 = np.where(m2.reshape(N, N, N) == 1, 75, 255).astype(np.uint8)
    want = run(lab2)
    wrong = t2.reshape(-1, 4)[:, 3] * -1.0
    assert np.abs(wrong - want[:, 3]).max() / np.abs(want[:, 3]).max() > 1.0
