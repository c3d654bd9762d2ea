"""The post-collision population check on the GPU (Simulator::validatePopulations,
solver/src/lbm/simulator.cpp:327-333, called after the collision at
solver/src/lbm/scsp.cpp:185 and colorfluid.cpp:431-432; message
"Non-finite post-collision population; no restart will be written.").

The condition is forced from the host, with every kernel ending normally: in
a slit one voxel wide along y (solid planes at y = 0 and y = 2, periodic in x
and z) both y links of a fluid site bounce back, so a pair of huge opposite
populations on them stays at that site.  The pair cancels in the density (the
density check passes), while the velocity it carries, about 2 x big / rho,
overflows in the collision's u^2 terms: 1e200 in FP64, 3e37 in FP32 (finite
in storage, their squares are not).  One step later exactly the population
bit is raised, and the engine reports the reference's message, the one the
oracle raises on the same input (FP64 pair).
"""
import numpy as np
import pytest

from lbflow_amd import cases as CS
from lbflow_amd.engine import FP32, FP64, Engine, EngineError
from oracle.pyoracle import Oracle, OracleError

pytestmark = pytest.mark.gpu

NZ, NY, NX = 6, 3, 8
PLUS_Y, MINUS_Y = 3, 4          # D3Q19 directions (0, 1, 0) and (0, -1, 0), e in (z, y, x) order
POPULATION_BIT, DENSITY_BIT = 0x10, 0x1


def slit_labels(two_phase):
    g = np.full((NZ, NY, NX), 75, np.uint8)
    if two_phase:
        g[:, :, NX // 2:] = 76
    g[:, 0, :] = 255
    g[:, 2, :] = 255
    return g


def plant(obj, phases, big):
    """The opposite pair on the y links of site (z, y, x) = (3, 1, 2)."""
    pops = obj.populations().reshape(phases * 19, NZ * NY * NX)
    i = (3 * NY + 1) * NX + 2              # phase 0 in the two-phase slit
    pops[PLUS_Y, i] = big                  # phase 0's planes come first
    pops[MINUS_Y, i] = -big
    obj.set_populations(pops.ravel(), obj.t)


def oracle_message(spec, phases):
    o = Oracle(CS.lattice(spec), CS.geometry(spec)).init().step(3)
    plant(o, phases, 1e200)
    with pytest.raises(OracleError) as err:
        o.step(1)
    return str(err.value)


def run_forced(spec, precision, stream_solids, phases):
    e = Engine(CS.lattice(spec), CS.geometry(spec), precision, stream_solids).init()
    try:
        for _ in range(3):
            e.step(0)
        e.synchronize()
        assert e.status()[0] == 0
        plant(e, phases, 1e200 if precision == FP64 else 3e37)
        e.step(0)
        e.synchronize()
        bits, _ = e.status()
        msg = None
        try:
            e.check_status()
        except EngineError as err:
            msg = str(err)
        return bits, msg
    finally:
        e.close()


@pytest.mark.parametrize('precision,stream_solids,eq', [
    (FP64, True, 'MRT'), (FP64, False, 'SRT'), (FP32, False, 'MRT'), (FP32, True, 'SRT')],
    ids=['fp64-dense-MRT', 'fp64-sparse-SRT', 'fp32-sparse-MRT', 'fp32-dense-SRT'])
def test_single_phase_population_check(precision, stream_solids, eq):
    spec = CS.make('slit-y1', slit_labels(False), [CS.label_line(75, 'GV', 1e-6, 1, 0),
                                                   CS.label_line(255, 'SD', 90)],
                   model='scsp', eq=eq, tau=(0.8,))
    bits, msg = run_forced(spec, precision, stream_solids, 1)
    assert bits == POPULATION_BIT, hex(bits)
    assert msg == oracle_message(spec, 1) == 'Non-finite post-collision population; no restart will be written.'


@pytest.mark.parametrize('precision', [FP64, FP32], ids=['fp64', 'fp32'])
def test_colour_population_check(precision):
    spec = CS.make('slit-y1-colour', slit_labels(True),
                   [CS.label_line(75, 'GV', 1e-6, 1, 0), CS.label_line(76, 'GV', 1e-6, 1, 1),
                    CS.label_line(255, 'SD', 60)],
                   model='color', eq='MRT', tau=(1.0, 1.0), sigma=1e-3, beta=0.7, alpha=1 / 3)
    bits, msg = run_forced(spec, precision, precision == FP64, 2)
    assert bits & POPULATION_BIT, hex(bits)
    assert not bits & DENSITY_BIT, hex(bits)
    want = oracle_message(spec, 2)
    assert want == 'Non-finite post-collision population; no restart will be written.'
    assert want in msg
