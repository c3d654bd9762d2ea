"""Zou-He pressure / velocity boundaries (PB/VB) on the GPU.

FP64 runs through the C ABI must equal the CPU oracle bit for bit (every
axial direction, both signs, SRT and MRT, D2Q9 and D3Q19); the fluid-only
layouts must evolve the fluid sites like the reference-exact dense layout;
the pressure-driven slit reproduces the reference's recorded report line
(dP and K to 7 significant digits) and the FP32 engine stays within the
tolerance written below of it.
"""
import sys

import numpy as np
import pytest

from lbflow_amd import cases as CS
from tests import cases as TC
from lbflow_amd.engine import FP32, FP64, STEP_ADJACENT, STEP_FIELDS, Engine
from tests.helpers import oracle_run

pytestmark = pytest.mark.gpu

CASES = TC.zouhe_cases()


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def engine_run(spec, steps, precision=FP64, stream_solids=True, dense=False):
    case = CS.lattice(spec)
    e = Engine(case, CS.geometry(spec), precision, stream_solids, dense).init()
    for k in range(steps):
        e.step(STEP_FIELDS if k == steps - 1 else 0)
    e.synchronize()
    e.check_status()
    return e, case


@pytest.mark.parametrize('spec', CASES, ids=lambda s: s['name'])
def test_zouhe_bitwise(spec):
    say('case', spec['name'])
    e, _ = engine_run(spec, 200)
    o, _ = oracle_run(spec, 200)
    pe, po = e.populations(), o.populations()
    diff = np.flatnonzero(pe.view(np.uint64) != po.view(np.uint64))
    assert diff.size == 0, f'{diff.size} populations differ, first {diff[:5]}'
    for a, b in zip(e.fields(), o.fields()):
        if a is not None:
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize('precision', [FP64, FP32], ids=['fp64', 'fp32'])
@pytest.mark.parametrize('spec', [s for s in CASES if s['eq'] == 'MRT'], ids=lambda s: s['name'])
def test_zouhe_layouts_match(spec, precision):
    say('case', spec['name'], precision)
    a, case = engine_run(spec, 120, precision, stream_solids=True)
    b, _ = engine_run(spec, 120, precision, stream_solids=False, dense=True)
    c, _ = engine_run(spec, 120, precision, stream_solids=False)
    assert 'sparse' in c.step_kernel and 'dense-fluid' in b.step_kernel
    g = CS.geometry(spec)
    solid = np.isin(g, [k for k, v in case['labels'].items() if v[0] in ('SD', 'BB')])
    fluid = np.tile(~solid, case['Q'])
    pa, pb, pc = a.populations(), b.populations(), c.populations()
    assert np.array_equal(pa[fluid], pb[fluid]) and np.array_equal(pa[fluid], pc[fluid])
    for x, y in zip(a.fields(), c.fields()):
        if x is not None:
            assert np.array_equal(x, y)


def slit_report(e, spec, case, first, last, tau=1.0, lattice_len=1e-6):
    """dP and K of SCSP::report (scsp.cpp:325-352) from the engine's report sums."""
    r = e.report(2, first, last, convergence=False, adjacent=False)
    dP = (r.rho_slice[0][0] / r.count_slice[0] - r.rho_slice[1][0] / r.count_slice[1]) / 3.0
    Q = r.vel_range / float(last - first + 1)
    nz, ny, nx = case['size']
    kpar = (last - first) * ((tau - 0.5) / 3) / (ny * nz) * lattice_len ** 2 * 1.0e12
    return dP, kpar * Q / dP


def test_pressure_slit_matches_reference_report(records):
    spec = TC.reference('pressure')[0]
    nx = spec['size'][0]
    e, case = engine_run(spec, spec['steps'])
    o, _ = oracle_run(spec, spec['steps'])
    assert np.array_equal(e.populations().view(np.uint64), o.populations().view(np.uint64))
    dP, K = slit_report(e, spec, case, 1, nx - 2)
    rec = records['pressure_report']
    say('dP', dP, 'K', K)
    assert f'{dP:.6e}' == f"{rec['report_slice_pressure_drop']:.6e}"
    assert f'{K:.6e}' == f"{rec['reported_K_um2']:.6e}"


def test_pressure_slit_fp32():
    """FP32 shifted populations on the slit: K within 2e-5 and dP within 2e-5 (relative)
    of the FP64 engine (dP is 1e-5 of the density, so FP32 rounding of rho is ~1e-3
    of dP per site; the slab averages bring it down)."""
    spec = TC.reference('pressure')[0]
    nx = spec['size'][0]
    a, case = engine_run(spec, spec['steps'])
    b, _ = engine_run(spec, spec['steps'], FP32, stream_solids=False)
    dPa, Ka = slit_report(a, spec, case, 1, nx - 2)
    dPb, Kb = slit_report(b, spec, case, 1, nx - 2)
    say('fp64', dPa, Ka, 'fp32', dPb, Kb, 'rel', abs(dPb / dPa - 1), abs(Kb / Ka - 1))
    assert abs(dPb / dPa - 1) < 2e-5
    assert abs(Kb / Ka - 1) < 2e-5


def test_zouhe_adjacent_report_matches_oracle():
    """The adjacent velocity of a PB/VB site is its value after this step's boundary
    (the reference copies it after zouHeBoundary, scsp.cpp:79-83): its term is 0."""
    spec = [s for s in CASES if s['name'] == 'zh-x-3d-MRT'][0]
    e, case = engine_run(spec, 30)
    o, _ = oracle_run(spec, 30)
    axis = 2
    e.report(axis, 1, case['size'][axis] - 2, convergence=True, adjacent=False)
    prev = o.fields()[2].ravel().copy()
    e.step(STEP_ADJACENT)
    o.step(1)
    adj = o.fields()[2].ravel().copy()
    e.step(STEP_FIELDS)
    o.step(1)
    g = np.tile(CS.geometry(spec), case['D'])
    zh = np.isin(g, [50, 200])
    adj[zh] = o.fields()[2].ravel()[zh]
    r = e.report(axis, 1, case['size'][axis] - 2, convergence=True, adjacent=True)
    ref = o.convergence_sums(prev, adj)
    np.testing.assert_allclose(list(r.conv), ref, rtol=1e-12, atol=1e-30)


# ---------------------------------------------------------------- colour model
COLOR_CASES = TC.zouhe_color_cases()


@pytest.mark.parametrize('spec', COLOR_CASES, ids=lambda s: s['name'])
def test_zouhe_color_bitwise(spec):
    say('case', spec['name'])
    e, _ = engine_run(spec, 200)
    o, _ = oracle_run(spec, 200)
    pe, po = e.populations(), o.populations()
    diff = np.flatnonzero(pe.view(np.uint64) != po.view(np.uint64))
    assert diff.size == 0, f'{diff.size} populations differ, first {diff[:5]}'
    for a, b in zip(e.fields(), o.fields()):
        if a is not None:
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize('spec', [s for s in COLOR_CASES if s['eq'] == 'MRT'], ids=lambda s: s['name'])
def test_zouhe_color_fp32_layouts_and_tracking(spec):
    """FP32 colour: sparse == dense fluid-only bit for bit; within the FP32 colour
    tolerances of the FP64 engine (rho0 5e-4 absolute, u 1e-2 relative L2)."""
    say('case', spec['name'])
    a, case = engine_run(spec, 200)
    b, _ = engine_run(spec, 200, FP32, stream_solids=False, dense=True)
    c, _ = engine_run(spec, 200, FP32, stream_solids=False)
    assert 'sparse' in c.step_kernel and 'dense-fluid' in b.step_kernel
    assert np.array_equal(b.populations(), c.populations())
    ra0, _, ua = a.fields()
    rc0, _, uc = c.fields()
    du = np.linalg.norm(ua - uc) / np.linalg.norm(ua)
    say('max |rho0 diff|', np.max(np.abs(ra0 - rc0)), 'rel L2 u', du)
    assert np.max(np.abs(ra0 - rc0)) < 5e-4
    assert du < 1e-2


def entry_measure(r0, r1, u, case, height):
    D = case['D']
    shape = (1,) + tuple(case['size']) if D == 2 else tuple(case['size'])
    r0, r1 = r0.reshape(shape), r1.reshape(shape)
    vx = u[D - 1].reshape(shape)
    rho = r0 + r1
    phi = np.divide(r0, rho, out=np.zeros_like(rho), where=rho > 0)
    interior = (0 if D == 2 else slice(1, -1), slice(1, -1), slice(1, -1))
    return (float(phi[interior].sum() / height ** (D - 1)), float(vx[interior].mean()), float(rho.sum()),
            [float(r0.sum()), float(r1.sum())])


@pytest.mark.parametrize('spec', TC.reference('entry', wall_gradient='legacy'), ids=lambda s: s['name'])
def test_entry_history_gpu(spec, records):
    """Capillary entry (porous_media.py:153-253) on the GPU in FP64: the reference's
    recorded history at every output step (3D has no oracle run here; the recorded
    values are the check).  2D compares exactly; in 3D the interior is a
    non-contiguous view whose numpy summation order differs between numpy versions,
    so the sums are compared to 4 ulp."""
    say('case', spec['name'])
    case = CS.lattice(spec)
    e = Engine(case, CS.geometry(spec), FP64, True).init()
    rec = records['entry'][spec['name'] + '-1rank']
    for h in rec['history'][1:]:
        n = h['step'] - e.t
        for k in range(n):
            e.step(STEP_FIELDS if k == n - 1 else 0)
        e.synchronize()
        e.check_status()
        got = entry_measure(*e.fields(), case, spec['height'])
        want = (h['phase0_equivalent_length'], h['pore_mean_ux'], h['total_mass'], h['phase_masses'])
        say(h['step'], got, want)
        if case['D'] == 2:
            assert got == want
        else:
            np.testing.assert_allclose(np.hstack(got), np.hstack(want), rtol=4 * np.finfo(float).eps, atol=0)


@pytest.mark.parametrize('spec', TC.reference('entry', dimensions=2, wall_gradient='legacy'), ids=lambda s: s['name'])
def test_entry_fp32(spec, records):
    """FP32 colour on the entry cases: the phase-0 length after 8000 steps within 5e-3
    (absolute, lattice units; 1.9e-3 measured) and the pore velocity within 1e-3
    relative (1.8e-4 measured) of the record."""
    case = CS.lattice(spec)
    e, _ = engine_run(spec, spec['steps'], FP32, stream_solids=False)
    rec = records['entry'][spec['name'] + '-1rank']['history'][-1]
    L, ux, _, _ = entry_measure(*e.fields(), case, spec['height'])
    say('fp32', L, rec['phase0_equivalent_length'], ux, rec['pore_mean_ux'])
    assert abs(L - rec['phase0_equivalent_length']) < 5e-3
    assert abs(ux / rec['pore_mean_ux'] - 1) < 1e-3


@pytest.mark.parametrize('precision', [FP64, FP32], ids=['fp64', 'fp32'])
@pytest.mark.parametrize('name', ['zhc-x-3d-MRT', 'zhc-t-3d-MRT'])
def test_zouhe_color_adjacent_report(precision, name):
    """Adjacent history at PB/VB sites of the colour model: the Zou-He velocity and
    (rho0 - rho1 of the previous step) / (Zou-He density) (colorfluid.cpp:166-171).
    On the -z velocity face of zhc-t the Zou-He velocity differs from the
    macroscopic one, so the history must come from the boundary, not the fields."""
    spec = [s for s in COLOR_CASES if s['name'] == name][0]
    e, case = engine_run(spec, 30, precision, stream_solids=precision == FP64)
    o, _ = oracle_run(spec, 30)
    axis = 2
    e.report(axis, 1, case['size'][axis] - 2, convergence=True, adjacent=False)
    prev = o.fields()[2].ravel().copy()
    r0, r1, _ = o.fields()
    rt = r0 + r1
    chi_prev = np.divide(r0 - r1, rt, out=np.zeros_like(rt), where=rt > 0)
    e.step(STEP_ADJACENT)
    o.step(1)
    e.step(STEP_FIELDS)
    o.step(1)
    adj_u, adj_chi = o.adjacent()
    r = e.report(axis, 1, case['size'][axis] - 2, convergence=True, adjacent=True)
    ref = o.convergence_sums(prev, adj_u)
    r0, r1, _ = o.fields()
    rt = r0 + r1
    chi = np.divide(r0 - r1, rt, out=np.zeros_like(rt), where=rt > 0)
    if precision == FP64:
        np.testing.assert_allclose(list(r.conv), ref, rtol=1e-12, atol=1e-30)
        assert r.phase_change[0] == np.max(np.abs(chi - chi_prev))
        assert r.phase_change[1] == np.max(np.abs(chi - adj_chi))
    else:
        np.testing.assert_allclose(list(r.conv)[1:3], ref[1:3], rtol=1e-4)
        assert abs(r.phase_change[1] - np.max(np.abs(chi - adj_chi))) < 1e-4


# ------------------------------------------------- renormalized wall gradient
@pytest.mark.parametrize('kw', [dict(model='color', eq='MRT', tau=(1.0, 1.0), two_phase=True, porosity=0.8),
                                dict(model='color', eq='SRT', tau=(1.0, 0.8), two_phase=True, angle=120,
                                     porosity=0.8, seed=4),
                                dict(model='color', eq='MRT', tau=(1.0, 1.0), two_phase=True, dims=2, angle=45,
                                     porosity=0.6, seed=4)],
                         ids=['mrt-3d', 'srt-3d-120', 'mrt-2d-45'])
def test_renormalized_wall_bitwise(kw):
    """wallGradient renormalized (colorfluid.cpp:487-492, 917-957) on curved walls:
    FP64 populations equal to the oracle's bit for bit; FP32 sparse == dense.  (Denser
    3D packs have wall sites whose fluid links span fewer than 3 directions; the
    reference rejects those, see the next test.)"""
    from tests.helpers import sphere_case
    spec = dict(sphere_case(**kw), wall_gradient='renormalized')
    say('case', spec['name'], kw)
    e, _ = engine_run(spec, 150)
    o, _ = oracle_run(spec, 150)
    assert np.array_equal(e.populations().view(np.uint64), o.populations().view(np.uint64))
    b, _ = engine_run(spec, 100, FP32, stream_solids=False, dense=True)
    c, _ = engine_run(spec, 100, FP32, stream_solids=False)
    assert np.array_equal(b.populations(), c.populations())


@pytest.mark.parametrize('spec', TC.reference('entry', wall_gradient='renormalized'), ids=lambda s: s['name'])
def test_entry_history_renormalized_gpu(spec, records):
    test_entry_history_gpu(spec, records)


def test_renormalized_wall_rejects_rank_deficient_stencils():
    """A wall site whose fluid links do not span all axes: the reference throws
    (colorfluid.cpp:940-941); the engine refuses the geometry with the same text."""
    from tests.helpers import sphere_case
    from lbflow_amd.engine import EngineError
    from oracle.pyoracle import Oracle
    spec = dict(sphere_case(model='color', eq='MRT', tau=(1.0, 1.0), two_phase=True), wall_gradient='renormalized')
    msg = 'Rank-deficient fluid wall stencil; reconstructed derivatives are unsupported here.'
    with pytest.raises(Exception, match=msg):
        Oracle(CS.lattice(spec), CS.geometry(spec))
    with pytest.raises(EngineError, match=msg):
        Engine(CS.lattice(spec), CS.geometry(spec), FP64, True)
