"""Pin the oracle's half-permeable film (HF) in the colour model (CPU only).

The reference has no recorded output for HF labels; its own check only
expects HF to be rejected together with the experimental CSF force
(scripts/capillary_checks.py:78-82).  The film's rule (ColorFluid::Run,
colorfluid.cpp:156-159; buildIb, simulator.cpp:418-421; clearHalfFilm,
colorfluid.cpp:752-764) fixes, independently of how the oracle states it:

  - one streaming half-step equals a push of every population to x + e_q,
    with each population whose destination is a solid site or a film site of
    the other phase reflected back into its own site with the opposite
    direction, and the other phase emptied at film sites (checked bit for
    bit against a numpy restatement in that destination form, with lattice
    vectors measured from the oracle's own streaming);
  - a film of the only phase present changes nothing: the run equals the run
    with the film labelled VD bit for bit;
  - a film plane that spans the channel is a perfect barrier for the other
    phase: its density is exactly 0 beyond the film at every step, and each
    phase's total mass is conserved to rounding.
"""
import numpy as np
import pytest

from lbflow_amd import cases as CS
from tests import cases as TC
from tests.helpers import oracle_run

CASES = TC.film_cases()
SOLID = ('SD', 'BB')
FLUID = ('VD', 'GV', 'PB', 'VB', 'HF')


def lattice_vectors(D):
    """e_q in reference component order, measured by streaming a single unit
    population through an all-fluid periodic lattice with the oracle."""
    from oracle.pyoracle import Oracle
    n = 5
    spec = CS.make('probe', np.zeros((n if D == 3 else 1, n, n)), [CS.label_line(0, 'VD', phase=0)], model='color', dimensions=D,
                   steps=1, write=1)
    case = CS.lattice(spec)
    Q = case['Q']
    size = case['size']
    centre = tuple(k // 2 for k in size)
    e = np.zeros((Q, D), int)
    for q in range(1, Q):
        o = Oracle(case, CS.geometry(spec)).init()
        f = np.zeros((2, Q) + tuple(size))
        f[(0, q) + centre] = 1.0
        o.set_populations(f.ravel(), 0)
        g = o.stream_only().populations().reshape((2, Q) + tuple(size))
        hit = np.argwhere(g[0, q] == 1.0)
        assert hit.shape == (1, D)
        e[q] = hit[0] - np.array(centre)
    return e


def stream_half_numpy(F, g, case, e):
    """Destination-form restatement of streaming + bounce-back + film."""
    labels = case['labels']
    Q, D = case['Q'], case['D']
    axes = tuple(range(-D, 0))
    kind = {lab: v[0] for lab, v in labels.items()}
    phase = {lab: v[3] for lab, v in labels.items()}
    solid = np.isin(g, [k for k, t in kind.items() if t in SOLID])
    fluid = np.isin(g, [k for k, t in kind.items() if t in FLUID])
    film_phase = np.full(g.shape, -1)
    for k, t in kind.items():
        if t == 'HF':
            film_phase[g == k] = phase[k]
    op = [int(np.flatnonzero((e == -e[q]).all(axis=1))[0]) for q in range(Q)]
    out = np.empty_like(F)
    for q in range(Q):
        out[:, q] = np.roll(F[:, q], tuple(e[q]), axis=axes)
    for q in range(1, Q):
        shift = tuple(-e[q])                      # value at x + e_q lands on x
        dest_solid = np.roll(solid, shift, axis=axes)
        dest_film = np.roll(film_phase, shift, axis=axes)
        for ph in (0, 1):
            m = fluid & (dest_solid | (dest_film == 1 - ph))
            out[ph, op[q]][m] = F[ph, q][m]
    for k in (0, 1):
        m = film_phase == k
        out[1 - k][:, m] = 0.0
    return out


@pytest.mark.parametrize('spec', [s for s in CASES if s['name'].startswith('film-block')], ids=lambda s: s['name'])
def test_film_streaming_matches_destination_form(spec):
    o, case = oracle_run(spec, 0)
    Q, size = case['Q'], tuple(case['size'])
    g = CS.geometry(spec).reshape(size)
    e = lattice_vectors(case['D'])
    rng = np.random.default_rng(7)
    F = rng.uniform(0.01, 0.1, (2, Q) + size)
    o.set_populations(F.ravel(), 5)
    got = o.stream_only().populations().reshape(F.shape)
    want = stream_half_numpy(F, g, case, e)
    fluid = np.isin(g, [k for k, v in case['labels'].items() if v[0] in FLUID])
    assert np.array_equal(got[:, :, fluid].view(np.uint64), want[:, :, fluid].view(np.uint64))
    # films are exercised in both directions and on both phases
    assert (want[1][:, g == 30] == 0).all() and (want[0][:, g == 31] == 0).all()


@pytest.mark.parametrize('spec', TC.film_cases(transparent=True), ids=lambda s: s['name'])
def test_film_of_the_only_phase_is_transparent(spec):
    plain = dict(spec, indices=[i.replace('"HF"', '"VD"') for i in spec['indices']])
    a, _ = oracle_run(spec, 150)
    b, _ = oracle_run(plain, 150)
    pa, pb = a.populations(), b.populations()
    assert np.array_equal(pa.view(np.uint64), pb.view(np.uint64))


def barrier_masks(spec, case):
    """Sites where each phase must stay absent: (no phase 1, no phase 0)."""
    nz, ny, nx = (1,) + tuple(case['size']) if case['D'] == 2 else tuple(case['size'])
    x = np.broadcast_to(np.arange(nx), (nz, ny, nx)).reshape(case['size'])
    if spec['name'].startswith('film-block'):
        return (x < 3) | (x > 11), np.zeros_like(x, bool)
    return x <= 1, x >= nx - 2


@pytest.mark.parametrize('spec', CASES, ids=lambda s: s['name'])
def test_film_is_a_barrier_and_conserves_each_phase(spec):
    o, case = oracle_run(spec, 0)
    g = CS.geometry(spec).reshape(case['size'])
    fluid = np.isin(g, [k for k, v in case['labels'].items() if v[0] in FLUID])
    no1, no0 = barrier_masks(spec, case)
    block = spec['name'].startswith('film-block')
    r0, r1, _ = o.fields()
    m0, m1 = r0[fluid.ravel()].sum(), r1[fluid.ravel()].sum()
    for _ in range(6):
        o.step(50)
        r0, r1, u = o.fields()
        r0, r1 = r0.reshape(g.shape), r1.reshape(g.shape)
        assert (r1[no1 & fluid] == 0).all()
        assert (r0[no0 & fluid] == 0).all()
        if block:
            # phase 0 cannot cross the phase-1 film at x = 7 either: its mass
            # reaches the slab only through the phase-0 films
            assert abs(r0[fluid].sum() - m0) <= 1e-11 * m0
            assert abs(r1[fluid].sum() - m1) <= 1e-11 * m1
            assert (r1[g == 101] > 0).all()
    if block:
        assert np.abs(u).max() > 1e-6          # the slab is driven against the films
    else:
        assert r0[(~no1) & (~no0) & fluid].max() > 0.1   # phase 0 has entered the channel
