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

from lbflow_amd import harness as H
from lbflow_amd.engine import FP32, FP64, STEP_ADJACENT, STEP_FIELDS, Engine
from tests.helpers import oracle_run

pytestmark = pytest.mark.gpu

CASES = H.zouhe_cases()


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def engine_run(spec, steps, precision=FP64, stream_solids=True, dense=False):
    case = H.lattice_case(spec)
    e = Engine(case, H.geometry(spec), precision, stream_solids, dense).init()
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
    g = H.geometry(spec)
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
    spec = H.pressure_case()
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
    spec = H.pressure_case()
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
    g = np.tile(H.geometry(spec), case['D'])
    zh = np.isin(g, [50, 200])
    adj[zh] = o.fields()[2].ravel()[zh]
    r = e.report(axis, 1, case['size'][axis] - 2, convergence=True, adjacent=True)
    ref = o.convergence_sums(prev, adj)
    np.testing.assert_allclose(list(r.conv), ref, rtol=1e-12, atol=1e-30)
