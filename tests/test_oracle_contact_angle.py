"""The oracle's geometric contact-angle rotation against what geometry fixes (CPU).

At a wall site the reference replaces the colour gradient g by the vector of
the same length that makes the contact angle theta with the wall normal n,
in the plane of n and g, on the tangential side g already points to
(colorfluid.cpp:550-567 with the frames of calcTibNorm, colorfluid.cpp:
1111-1129, and RotateT2D/Tx3D/Tz3D, fn.cpp:118-164).  That is the closed form

    g' = |g| (cos(theta) n^ + sin(theta) t^),  t^ = tangential part of g, normalised,

so |g'| = |g|, n^ . g' = |g| cos(theta) and g' lies in the plane of n and g.
The oracle's rotation is checked against it for every integer wall normal
that D3Q19 (and D2Q9) can produce, n = -sum of e_q over the solid links, with
gradients in general position, along the normal, against it and tangential,
and angles over the whole range.  Each planted mistake (tangential sign,
frame not inverted, sin/cos swapped, gradient reversed) must fail the check.
"""
import itertools

import numpy as np
import pytest

from oracle.pyoracle import contact_angle

ANGLES = (0.0, 30.0, 45.0, 60.0, 90.0, 120.0, 135.0, 150.0, 180.0, 17.3, 101.7)


def lattice(D):
    if D == 2:
        return [v for v in itertools.product((-1, 0, 1), repeat=2) if any(v)]
    return [v for v in itertools.product((-1, 0, 1), repeat=3) if 0 < sum(map(abs, v)) <= 2]


def wall_normals(D):
    """Every distinct n = -sum of e_q over a non-empty set of solid links."""
    e = np.array(lattice(D))
    subsets = np.array(list(itertools.product((0, 1), repeat=len(e))), dtype=np.int64)
    normals = np.unique(-(subsets @ e), axis=0)
    return [n for n in normals if np.any(n)]


def gradients(n, rng):
    nh = n / np.linalg.norm(n)
    t = rng.normal(size=len(n))
    t -= (t @ nh) * nh
    out = [rng.normal(size=len(n)) * 10.0 ** rng.uniform(-6, 0) for _ in range(3)]
    out += [0.37 * nh, -0.8 * nh, t / np.linalg.norm(t) * 0.05, 0.2 * nh + 1e-3 * t]
    return out


def closed_form_error(n, g, angle, gp):
    """Largest deviation of g' from |g| (cos n^ + sin t^), relative to |g|
    and divided by the conditioning of t^, max(1, |g| / |t|) (for g along
    the normal only the length and the angle are fixed)."""
    nh = n / np.linalg.norm(n)
    norm = np.linalg.norm(g)
    a = np.radians(angle)
    t = g - (g @ nh) * nh
    if np.linalg.norm(t) <= 1e-9 * norm:
        return max(abs(np.linalg.norm(gp) - norm), abs(gp @ nh - norm * np.cos(a))) / norm
    want = norm * (np.cos(a) * nh + np.sin(a) * t / np.linalg.norm(t))
    return float(np.max(np.abs(gp - want))) / norm / max(1.0, norm / np.linalg.norm(t))


@pytest.mark.parametrize('D', (3, 2))
def test_rotation_closed_form_every_wall_normal(D):
    rng = np.random.default_rng(7 + D)
    normals = wall_normals(D)
    assert len(normals) == (870 if D == 3 else 36)
    worst = 0.0
    for n in normals:
        n = n.astype(np.float64)
        for g in gradients(n, rng):
            for angle in ANGLES:
                gp = contact_angle(n, g, angle)
                worst = max(worst, closed_form_error(n, g, angle, gp))
    assert worst < 1e-12, worst


@pytest.mark.parametrize('plant', (1, 2, 3, 4), ids=('sin-sign', 'frame-not-inverted', 'sin-cos-swapped',
                                                      'gradient-reversed'))
@pytest.mark.parametrize('D', (3, 2))
def test_planted_rotation_mistake_is_caught(D, plant):
    rng = np.random.default_rng(11)
    caught = 0
    for n in wall_normals(D)[::7]:
        n = n.astype(np.float64)
        for g in gradients(n, rng)[:3]:
            for angle in (60.0, 120.0):
                caught += closed_form_error(n, g, angle, contact_angle(n, g, angle, plant)) > 1e-6
    assert caught > 0


def test_tangential_gradient_at_ninety_degrees_is_unchanged():
    """theta = 90 degrees: a gradient already tangential to the wall stays as it is."""
    for n in wall_normals(3)[::11]:
        n = n.astype(np.float64)
        nh = n / np.linalg.norm(n)
        g = np.cross(nh, [0.3, -0.2, 0.9])
        if np.linalg.norm(g) < 1e-3:
            continue
        assert np.allclose(contact_angle(n, g, 90.0), g, rtol=0, atol=1e-15)
