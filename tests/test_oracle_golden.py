"""Pin the CPU oracle to the reference's own recorded results (CPU only).

Every expected value below was produced by the reference solver and is
stored in the reference tree (research/drlab-publication/paper/data/*.json,
extracted into tests/golden/reference_records.json by make_golden.py), except
the 8x10x4 checkpoint digest: the SHA-256 of the checkpoint that the
reference, built from its own sources, wrote after 20 000 steps.  Comparisons are exact (==): the oracle keeps
the reference's floating-point operation order.
"""
import hashlib

import numpy as np
import pytest

from lbflow_amd import cases as CS
from tests import cases as TC
from tests.helpers import (C1_SHA256, c1_profile_error, c1_spec, darcy, oracle_run, profile, solid_shift_map,
                           sphere_case, velocity_grid)


@pytest.mark.parametrize('spec', TC.reference('force'), ids=lambda s: s['name'])
def test_uniform_force(spec, records):
    """u(t) = t F after 100 steps; recorded max errors reproduced exactly (benchmarks.py:176-188)."""
    o, case = oracle_run(spec, spec['steps'])
    v = velocity_grid(o.fields()[2], case)
    err = float(np.max(np.abs(v[..., 0] - 1e-4)))
    transverse = float(np.max(np.abs(v[..., 1:])))
    rec = records['force'][spec['name'] + '-1rank']
    assert err == rec['absolute_error']
    assert transverse == rec['transverse_max']
    assert err <= 1e-12


CHANNELS = TC.reference('channel', height=(8, 16))


@pytest.mark.parametrize('spec', CHANNELS, ids=lambda s: s['name'])
def test_channel_profile(spec, records):
    """Poiseuille channels: recorded measured_profile reproduced exactly (benchmarks.py:191-242)."""
    o, case = oracle_run(spec, spec['steps'])
    got = profile(o.fields()[2], case)
    want = np.array(records['channel'][spec['name'] + '-1rank']['measured_profile'])
    assert np.array_equal(got, want)


def _layered(interp):
    return TC.reference('layered', viscosity_interpolation=interp, height=16)


@pytest.mark.parametrize('spec', _layered('arithmetic') + _layered('harmonic'), ids=lambda s: s['name'])
def test_layered_two_phase(spec, records):
    """Two-phase D2Q9 channel with a flat interface and 90-degree walls (layered_flow.py:10-23):
    perturbation, recolouring, wall rotation and both viscosity interpolations."""
    o, case = oracle_run(spec, spec['steps'])
    got = profile(o.fields()[2], case)
    key = spec['name'].replace('-harmonic', '') + ('-harmonic' if case['harmonic'] else '') + '-1rank'
    want = np.array(records['layered'][key]['profile'])
    assert np.array_equal(got, want)


@pytest.mark.parametrize('spec', TC.reference('bundle', scale=1), ids=lambda s: s['name'])
def test_bundle_history(spec, records):
    """Slit bundles (porous_media.py:67-151): Darcy velocity and total mass at every output step."""
    rec = records['bundle'][spec['name'] + '-1rank']['history']
    o, case = oracle_run(spec, 0)
    t = 0
    for h in rec[1:]:
        o.step(h['step'] - t)
        t = h['step']
        r0, r1, u = o.fields()
        assert darcy(u, case) == h['darcy_velocity']
        rho = r0 if r1 is None else r0 + r1
        assert float(rho.sum()) == h['total_mass']


def pressure_report(sums, case, tau, first, last, lattice_len=1e-6):
    """dP and K of SCSP::report (scsp.cpp:325-352, simulator.cpp:151-160) from the report sums."""
    vel, _, _, s0, _, s1, _, c0, c1 = sums
    dP = (s0 / c0 - s1 / c1) / 3.0
    Q = vel / float(last - first + 1)
    nz, ny, nx = case['size']
    kpar = (last - first) * ((tau - 0.5) * 1 / 3) / (ny * nz) * lattice_len ** 2 * 1.0e12
    return dP, kpar * Q / dP


def test_pressure_boundaries_report(records):
    """Zou-He pressure inlet/outlet slit (porous_media.py:257-279): the report line the
    reference recorded (dP and K printed with 7 significant digits) is reproduced, and
    both lie within 2e-4 of the analytic slit."""
    spec = TC.reference('pressure')[0]
    o, case = oracle_run(spec, spec['steps'])
    nx = spec['size'][0]
    dP, K = pressure_report(o.report_sums(2, 1, nx - 2), case, 1.0, 1, nx - 2)
    rec = records['pressure_report']
    assert f'{dP:.6e}' == f"{rec['report_slice_pressure_drop']:.6e}"
    assert f'{K:.6e}' == f"{rec['reported_K_um2']:.6e}"
    h, dp = spec['height'], spec['pressure_drop']
    assert abs(dP / (dp * (nx - 3) / (nx - 1)) - 1) < 2e-4
    assert abs(K / ((h ** 3 / 12 + h / 24) / (h + 2)) - 1) < 2e-4


def entry_measure(r0, r1, u, case, height):
    """phase0_equivalent_length, pore_mean_ux and total_mass as scripts/porous_media.py:
    187-197 and :20-31 compute them from the VTI fields (x-fastest grids, interior
    without the boundary planes)."""
    D = case['D']
    shape = (1,) + tuple(case['size']) if D == 2 else tuple(case['size'])
    r0, r1 = r0.reshape(shape), r1.reshape(shape)
    vx = u[D - 1].reshape(shape)
    rho = r0 + r1
    phi = np.divide(r0, rho, out=np.zeros_like(rho), where=rho > 0)
    interior = (0 if D == 2 else slice(1, -1), slice(1, -1), slice(1, -1))
    return (float(phi[interior].sum() / height ** (D - 1)), float(vx[interior].mean()), float(rho.sum()),
            [float(r0.sum()), float(r1.sum())])


@pytest.mark.parametrize('spec', TC.reference('entry', dimensions=2),
                         ids=lambda s: s['name'])
def test_entry_history(spec, records):
    """Colour-model capillary entry between Zou-He pressure boundaries
    (porous_media.py:153-253), legacy and renormalized wall gradients: the recorded
    displacement history is reproduced exactly."""
    from oracle.pyoracle import Oracle
    case = CS.lattice(spec)
    o = Oracle(case, CS.geometry(spec)).init()
    rec = records['entry'][spec['name'] + '-1rank']
    for h in rec['history'][1:]:
        o.step(h['step'] - o.t)
        r0, r1, u = o.fields()
        L, ux, mass, phases = entry_measure(r0, r1, u, case, spec['height'])
        assert L == h['phase0_equivalent_length']
        assert ux == h['pore_mean_ux']
        assert mass == h['total_mass']
        assert phases == h['phase_masses']


def test_c1_profile_closed_form():
    """8x10x4 SRT channel after 20000 steps: u_x(y) is the halfway bounce-back
    Poiseuille profile with the BGK slip F (16 L - 3) / (24 nu) to roundoff;
    without the slip term it would be off by 5e-3."""
    o, case = oracle_run(c1_spec(), 20000)
    assert c1_profile_error(profile(o.fields()[2], case)) < 1e-10


def test_solid_entries_are_upstream_copies():
    """Solid-site checkpoint bytes follow from streaming alone: on a sphere
    pack every solid entry of plane q after a step is the previous entry of
    the upstream site along one lattice vector, a different vector per plane;
    with those vectors the same holds for the 8x10x4 channel's solid rows at
    step 20000, whose checkpoint digest the reference recorded."""
    pack = sphere_case(n=(12, 14, 16), model='scsp', eq='MRT', tau=(0.8,))
    o, case = oracle_run(pack, 50)
    before = o.populations()
    o.step(1)
    shifts = solid_shift_map(before, o.populations(), CS.geometry(pack), case)
    assert all(len(m) == 1 for m in shifts), shifts
    vectors = [m[0] for m in shifts]
    assert len(set(vectors)) == case['Q'] - 1
    c1, ccase = oracle_run(c1_spec(), 19999)
    before = c1.populations()
    c1.step(1)
    on_c1 = solid_shift_map(before, c1.populations(), CS.geometry(c1_spec()), ccase)
    assert all(v in m for v, m in zip(vectors, on_c1))


def test_c1_checkpoint_digest():
    """8x10x4 SRT channel, 20000 steps: checkpoint bytes equal the reference's (solid entries included)."""
    s = c1_spec()
    o, case = oracle_run(s, 20000)
    digest = hashlib.sha256(o.populations().astype('<f8').tobytes()).hexdigest()
    assert digest == C1_SHA256


@pytest.mark.slow
def test_bentheimer48_history(records):
    """Bentheimer48 crop, D3Q19 MRT, tau 1, F 1e-6 (rock_workflow.py:19-39): step-1000 record."""
    s = TC.reference_case('rock', 'rock-scsp-F1e-06')
    rec = records['rock_scsp']['rock-scsp-F1e-06-2rank']
    assert hashlib.sha256(CS.geometry(s).tobytes()).hexdigest() == rec['geometry_sha256']
    o, case = oracle_run(s, 1000)
    r0, _, u = o.fields()
    h = rec['history'][1]
    assert h['step'] == 1000
    assert darcy(u, case) == h['darcy_velocity']
    assert float(r0.sum()) == h['total_mass']
    pore = r0 > 0
    assert [float(r0[pore].min()), float(r0[pore].max())] == h['density_range']


def test_mrt_operator_identities():
    """M MInv = I, conserved rates 0, B = MInv (I - S/2) M (moment_audit.py:23-35)."""
    from oracle.pyoracle import mrt_matrices
    for Q in (9, 19):
        for tau in (0.6, 1.0, 1.4):
            A, B, M, MInv = mrt_matrices(Q, tau)
            assert np.allclose(M @ MInv, np.eye(Q), atol=3e-14)
            S = np.diag(M @ A @ MInv)
            conserved = [0, 3, 5] if Q == 9 else [0, 3, 5, 7]
            assert np.allclose(S[conserved], 0.0, atol=3e-14)
            assert np.allclose(M @ B @ MInv, np.eye(Q) - 0.5 * np.diag(S), atol=3e-14)
            su = 1 / tau
            assert np.isclose(S[1], su)


def test_planted_error_is_caught():
    """A plausible mistake (wrong relaxation time) in the oracle input changes the recorded profile."""
    s = TC.reference_case('channel', 'channel-scsp-d3-MRT-tau0.6-H8')
    bad = dict(s, tau=(0.6000001,))
    o, case = oracle_run(bad, s['steps'])
    got = profile(o.fields()[2], case)
    assert not np.array_equal(got, np.array([0]))
    from tests.conftest import GOLDEN
    import json
    want = np.array(json.loads((GOLDEN / 'reference_records.json').read_text())
                    ['channel'][s['name'] + '-1rank']['measured_profile'])
    assert not np.array_equal(got, want)
