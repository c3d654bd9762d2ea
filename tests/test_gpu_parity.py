"""Parity of the HIP engine (through the C ABI) with the CPU oracle.

FP64 mode keeps the reference's operation order, so populations, macro
fields and link tables must be bit-identical to the oracle's.  FP32 mode is
checked against FP64 with the tolerances written in each test.
"""
import hashlib
import sys

import numpy as np
import pytest

from lbflow_amd import cases as CS
from tests import cases as TC
from lbflow_amd.engine import FP32, FP64, STEP_ADJACENT, STEP_FIELDS, Engine
from tests.helpers import C1_SHA256, c1_spec, darcy, oracle_run, profile, solid_shift_map, sphere_case

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def engine_run(spec, steps, precision=FP64, stream_solids=True, geometry=None, dense=False):
    case = CS.lattice(spec)
    e = Engine(case, CS.geometry(spec) if geometry is None else geometry, precision, stream_solids,
               dense).init()
    for k in range(steps):
        e.step(STEP_FIELDS if k == steps - 1 else 0)
    e.synchronize()
    return e, case


def assert_same(eng, orc):
    pe, po = eng.populations(), orc.populations()
    assert pe.shape == po.shape
    diff = np.flatnonzero(pe.view(np.uint64) != po.view(np.uint64))
    assert diff.size == 0, f'{diff.size} populations differ, first {diff[:5]}: ' \
                           f'{pe[diff[:3]]} vs {po[diff[:3]]}'
    fe, fo = eng.fields(), orc.fields()
    for a, b in zip(fe, fo):
        if a is not None:
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize('spec', TC.reference('force'), ids=lambda s: s['name'])
def test_force_cases_bitwise(spec):
    say('case', spec['name'])
    e, case = engine_run(spec, spec['steps'])
    o, _ = oracle_run(spec, spec['steps'])
    assert_same(e, o)
    e.check_status()


@pytest.mark.parametrize('spec', TC.reference('channel', height=8), ids=lambda s: s['name'])
def test_channel_bitwise(spec, records):
    say('case', spec['name'])
    e, case = engine_run(spec, spec['steps'])
    got = profile(e.fields()[2], case)
    want = np.array(records['channel'][spec['name'] + '-1rank']['measured_profile'])
    assert np.array_equal(got, want)
    o, _ = oracle_run(spec, spec['steps'])
    assert_same(e, o)


@pytest.mark.parametrize('spec', TC.reference('layered', viscosity_interpolation='arithmetic', height=16), ids=lambda s: s['name'])
def test_layered_bitwise(spec, records):
    say('case', spec['name'])
    e, case = engine_run(spec, spec['steps'])
    got = profile(e.fields()[2], case)
    want = np.array(records['layered'][spec['name'] + '-1rank']['profile'])
    assert np.array_equal(got, want)


@pytest.mark.parametrize('spec', TC.reference('bundle', scale=1), ids=lambda s: s['name'])
def test_bundle_bitwise(spec, records):
    say('case', spec['name'])
    rec = records['bundle'][spec['name'] + '-1rank']['history'][1]
    e, case = engine_run(spec, rec['step'])
    assert darcy(e.fields()[2], case) == rec['darcy_velocity']


PACKS = [
    dict(model='scsp', eq='SRT', tau=(0.8,)),
    dict(model='scsp', eq='MRT', tau=(0.8,)),
    dict(model='scsp', eq='MRT', tau=(1.3,), dims=2),
    dict(model='color', eq='MRT', tau=(1.0, 1.0), two_phase=True),
    dict(model='color', eq='MRT', tau=(0.7, 1.2), two_phase=True),
    dict(model='color', eq='SRT', tau=(1.0, 0.8), two_phase=True, angle=120),
    dict(model='color', eq='MRT', tau=(1.0, 1.0), two_phase=True, dims=2, angle=45),
]


@pytest.mark.parametrize('kw', PACKS, ids=lambda k: '-'.join(str(v) for v in k.values()))
def test_sphere_pack_bitwise(kw):
    """Curved walls: 3D contact-angle rotation, link masks, solids streamed like the reference."""
    spec = sphere_case(**kw)
    say('case', spec['name'], kw)
    steps = 150
    e, case = engine_run(spec, steps)
    o, _ = oracle_run(spec, steps)
    m_e, w_e, n_e, s_e, c_e = e.links()
    m_o, w_o, n_o, s_o, c_o = o.links()
    assert np.array_equal(m_e, m_o) and np.array_equal(w_e, w_o)
    assert np.array_equal(n_e, n_o) and np.array_equal(s_e, s_o) and np.array_equal(c_e, c_o)
    assert_same(e, o)
    e.check_status()


@pytest.mark.parametrize('precision', [FP64, FP32], ids=['fp64', 'fp32'])
@pytest.mark.parametrize('eq', ['SRT', 'MRT'])
@pytest.mark.parametrize('dims', [3, 2])
def test_fluid_only_layouts_match(precision, eq, dims):
    """Fluid-only layouts (dense with solids skipped, sparse fluid list) evolve the fluid
    sites exactly like the reference-exact dense layout; solid entries read 0."""
    spec = sphere_case(model='scsp', eq=eq, tau=(0.8,), dims=dims)
    a, case = engine_run(spec, 60, precision, stream_solids=True)
    b, _ = engine_run(spec, 60, precision, stream_solids=False, dense=True)
    c, _ = engine_run(spec, 60, precision, stream_solids=False)
    assert 'sparse' in c.step_kernel and 'dense-fluid' in b.step_kernel
    g = CS.geometry(spec)
    fluid = np.tile(g == 75, case['Q'])
    pa, pb, pc = a.populations(), b.populations(), c.populations()
    assert np.array_equal(pa[fluid], pb[fluid]) and np.array_equal(pa[fluid], pc[fluid])
    assert np.all(pb[~fluid] == 0.0) and np.all(pc[~fluid] == 0.0)
    for x, y in zip(a.fields(), c.fields()):
        if x is not None:
            assert np.array_equal(x, y)


def test_restart_equivalence():
    """Continuous run == checkpoint + resume (ci_smoke.py:42-56)."""
    spec = sphere_case(model='color', eq='MRT', tau=(0.7, 1.2), two_phase=True)
    cont, case = engine_run(spec, 80)
    half, _ = engine_run(spec, 40)
    pops = half.populations()
    res = Engine(case, CS.geometry(spec), FP64, True)
    res.set_populations(pops, 40)
    for k in range(40):
        res.step(STEP_FIELDS if k == 39 else 0)
    assert res.t == 80
    assert np.array_equal(res.populations(), cont.populations())


@pytest.mark.parametrize('precision,stream_solids', [(FP64, True), (FP32, False)], ids=['fp64-dense', 'fp32-sparse'])
def test_set_populations_waits_for_queued_steps(precision, stream_solids):
    """lbf_set_populations on an engine with steps still queued on its
    stream (lbf_advance returns without waiting): the upload waits for them,
    so restoring the step-10 state right after queueing 7 more steps goes on
    exactly like the same restore with nothing queued."""
    spec = sphere_case(model='scsp', eq='MRT', tau=(0.8,))
    case, geo = CS.lattice(spec), CS.geometry(spec)
    a = Engine(case, geo, precision, stream_solids).init()
    a.advance(10)
    a.set_populations(a.populations(), 10)
    a.advance(10)
    want = a.populations()
    b = Engine(case, geo, precision, stream_solids).init()
    b.advance(10)
    p10 = b.populations()
    b.advance(7)                 # queued, not waited for
    b.set_populations(p10, 10)
    b.advance(10)
    assert b.t == 20
    assert np.array_equal(b.populations().view(np.uint64), want.view(np.uint64))


def test_solid_entries_are_upstream_copies_gpu():
    """The engine streams solid sites like the reference: every solid entry
    of plane q after a step is the previous entry of the upstream site, one
    distinct lattice vector per plane."""
    spec = sphere_case(n=(12, 14, 16), model='scsp', eq='MRT', tau=(0.8,))
    e, case = engine_run(spec, 50)
    before = e.populations()
    e.step(0)
    shifts = solid_shift_map(before, e.populations(), CS.geometry(spec), case)
    assert all(len(m) == 1 for m in shifts), shifts
    assert len({m[0] for m in shifts}) == case['Q'] - 1


def test_c1_checkpoint_digest_gpu():
    """The 8x10x4 SRT channel after 20000 GPU steps has the reference's checkpoint bytes."""
    spec = c1_spec()
    e, case = engine_run(spec, 20000)
    digest = hashlib.sha256(e.populations().astype('<f8').tobytes()).hexdigest()
    assert digest == C1_SHA256


def test_report_sums_match_oracle():
    spec = sphere_case(model='color', eq='MRT', tau=(1.0, 1.0), two_phase=True)
    e, case = engine_run(spec, 30)
    o, _ = oracle_run(spec, 30)
    axis = 2
    r = e.report(axis, 1, case['size'][axis] - 2, convergence=True, adjacent=False)
    assert r.has_previous == 0
    ref = o.report_sums(axis, 1, case['size'][axis] - 2)
    got = [r.vel_range, r.rho_range[0], r.rho_range[1], r.rho_slice[0][0], r.rho_slice[0][1],
           r.rho_slice[1][0], r.rho_slice[1][1], r.count_slice[0], r.count_slice[1]]
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-18)
    # second report: convergence sums against the stored previous velocity
    prev = o.fields()[2].ravel().copy()
    e.step(STEP_ADJACENT)
    o.step(1)
    adj = o.fields()[2].ravel().copy()
    e.step(STEP_FIELDS)
    o.step(1)
    r = e.report(axis, 1, case['size'][axis] - 2, convergence=True, adjacent=True)
    assert r.has_previous == 1
    ref = o.convergence_sums(prev, adj)
    np.testing.assert_allclose(list(r.conv), ref, rtol=1e-12, atol=1e-30)


@pytest.mark.parametrize('eq', ['SRT', 'MRT'])
def test_fp32_tracks_fp64(eq):
    """FP32 shifted populations: velocity field within 1e-5 relative (L2) of FP64 after 300 steps."""
    spec = sphere_case(n=(16, 16, 16), model='scsp', eq=eq, tau=(0.8,), porosity=0.4)
    a, case = engine_run(spec, 300, FP64)
    b, _ = engine_run(spec, 300, FP32, stream_solids=False)
    ua, ub = a.fields()[2], b.fields()[2]
    rel = np.linalg.norm(ua - ub) / np.linalg.norm(ua)
    say('fp32 rel', rel)
    assert rel < 1e-5
    ra, rb = a.fields()[0], b.fields()[0]
    assert np.max(np.abs(ra - rb)) < 1e-6
