"""Edge cases of the lattice on the GPU: a lattice with no fluid site and a
lattice with a single fluid site, in both models, both precisions and every
layout.  The reference runs them (every label defined, links from buildIb);
the engine must too: no failure bits, fields 0 on solid sites, and the
single site (all 18 links bounce back) keeps its rest state: density 1 and
velocity 0, bit for bit with the oracle in FP64.
"""
import sys

import numpy as np
import pytest

from lbflow_amd import cases as CS
from lbflow_amd.engine import FP32, FP64, STEP_FIELDS, Engine
from tests.helpers import oracle_run

pytestmark = pytest.mark.gpu

LAYOUTS = [(FP64, True, False), (FP64, False, False), (FP32, False, False), (FP32, False, True)]


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def lattice(model, fluid_site):
    indices = [CS.label_line(255, 'SD', 90), CS.label_line(0, 'VD', phase=0)]
    labels = np.full((3, 3, 3), 255, np.uint8)
    if fluid_site:
        labels[1, 1, 1] = 0
    return CS.make(f'edge-{model}-{"one" if fluid_site else "none"}', labels, indices,
                   model=model, eq='MRT', tau=(0.8, 0.8) if model == 'color' else (0.8,))


@pytest.mark.parametrize('precision,stream_solids,dense', LAYOUTS, ids=['fp64-dense', 'fp64-fluid', 'fp32-sparse',
                                                                         'fp32-dense-fluid'])
@pytest.mark.parametrize('model', ['scsp', 'color'])
@pytest.mark.parametrize('fluid_site', [False, True], ids=['no-fluid', 'one-fluid-site'])
def test_degenerate_lattices(model, fluid_site, precision, stream_solids, dense):
    spec = lattice(model, fluid_site)
    case = CS.lattice(spec)
    say('case', spec['name'], precision, stream_solids, dense)
    e = Engine(case, CS.geometry(spec), precision, stream_solids, dense).init()
    for k in range(20):
        e.step(STEP_FIELDS if k == 19 else 0)
    e.synchronize()
    e.check_status()
    r0, r1, u = e.fields()
    solid = CS.geometry(spec) == 255
    assert (r0[solid] == 0).all() and (u[:, solid] == 0).all()
    if fluid_site:
        site = 13
        rho = r0[site] + (r1[site] if r1 is not None else 0.0)
        assert abs(rho - 1.0) < (1e-14 if precision == FP64 else 1e-6)
        assert np.abs(u[:, site]).max() < (1e-15 if precision == FP64 else 1e-7)
        if precision == FP64:
            o, _ = oracle_run(spec, 20)
            if stream_solids:
                assert np.array_equal(e.populations().view(np.uint64), o.populations().view(np.uint64))
            for a, b in zip(e.fields(), o.fields()):
                if a is not None:
                    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
