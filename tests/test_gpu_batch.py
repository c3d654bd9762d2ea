"""Batched crops on the GPU: every crop equals its own single-domain run.

The batch has no reference counterpart; it replaces one LBFlow process per
crop (scripts/rock_workflow.py:63-69).  It is pinned by the invariant that a
crop's populations after n steps are bitwise those of the single-domain FP32
fluid-only engine on that crop alone (the same operator and FP32 arithmetic,
SCSP::equilSRT/equilMRT, scsp.cpp:374-464), and that each crop stops at the
step the reference's rule (Simulator::updateConvergence, simulator.cpp:
248-325) gives for that crop alone.
"""
import sys

import numpy as np
import pytest

from lbflow_amd import geometry as G
from lbflow_amd.engine import FP32, Batch, Engine
from lbflow_amd.permeability import converge

pytestmark = pytest.mark.gpu

LABELS = {75: ('GV', 1e-5, 1, 0), 255: ('SD', 90.0, 0, 0)}


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def crops_3d():
    shapes = [(12, 10, 8), (9, 14, 11), (16, 16, 16), (7, 5, 9), (20, 6, 13), (10, 10, 10)]
    out = []
    for i, s in enumerate(shapes):
        g = G.sphere_pack(s, 0.35 + 0.05 * (i % 3), 1.5, 3.5, 40 + i)
        out.append(g)
    out[-1] = np.full((10, 10, 10), 255, np.uint8)    # a crop without fluid
    return out


def crops_2d():
    shapes = [(10, 12), (17, 9), (8, 8), (24, 30)]
    return [G.sphere_pack((1,) + s, 0.4, 1.5, 3.0, 60 + i).reshape(s) for i, s in enumerate(shapes)]


def single_case(D, Q, eq, shape, tau=0.8):
    return dict(D=D, Q=Q, size=tuple(shape), model='scsp', eq=eq, tau=(tau,), density=(1.0,),
                labels=LABELS)


@pytest.mark.parametrize('D,Q,eq', [(3, 19, 'SRT'), (3, 19, 'MRT'), (2, 9, 'MRT')],
                         ids=['d3-SRT', 'd3-MRT', 'd2-MRT'])
def test_batch_populations_equal_single_domain(D, Q, eq):
    crops = crops_3d() if D == 3 else crops_2d()
    say('batch', D, Q, eq, [c.shape for c in crops])
    b = Batch(D, Q, eq, 0.8, 1.0, LABELS, crops)
    b.init().advance(37)
    assert b.t == 37
    for c, g in enumerate(crops):
        e = Engine(single_case(D, Q, eq, g.shape), g, FP32, stream_solids=False).init()
        e.advance(37)
        want = e.populations()
        got = b.populations(c)
        assert got.shape == want.shape
        assert np.array_equal(got, want), (c, np.max(np.abs(got - want)))
        e.close()
    b.close()


def test_batch_tune_then_init_restarts():
    crops = crops_3d()[:4]
    b = Batch(3, 19, 'MRT', 0.8, 1.0, LABELS, crops)
    per, ms = b.tune(steps=4)
    say('tuned', per, ms)
    assert per in (1, 2, 4, 8, 16, 32, 64, 128) and ms > 0
    b.init().advance(11)
    e = Engine(single_case(3, 19, 'MRT', crops[2].shape), crops[2], FP32, stream_solids=False).init()
    e.advance(11)
    assert np.array_equal(b.populations(2), e.populations())


def test_batch_convergence_matches_single_domain():
    crops = [G.sphere_pack(s, 0.45, 1.5, 3.0, 70 + i) for i, s in enumerate([(10, 12, 14), (8, 8, 8), (14, 9, 12)])]
    b = Batch(3, 19, 'MRT', 0.8, 1.0, LABELS, crops).init()
    res, sec = b.run(axis=2, tol=1e-5, interval=40, max_steps=20000)
    say('batch results', res, sec)
    for c, g in enumerate(crops):
        e = Engine(single_case(3, 19, 'MRT', g.shape), g, FP32, stream_solids=False).init()
        one = converge(e, axis=2, first=0, last=g.shape[2] - 1, tol=1e-5, interval=40, max_steps=20000,
                       force=1e-5, tau=0.8)
        say('single', c, one['steps'], one['K'])
        assert res[c]['converged'] and one['converged']
        assert res[c]['steps'] == one['steps']
        assert abs(res[c]['K'] - one['K']) <= 1e-12 * abs(one['K'])
        # the crop stopped stepping at its own convergence step
        e2 = Engine(single_case(3, 19, 'MRT', g.shape), g, FP32, stream_solids=False).init()
        e2.advance(one['steps'])
        assert np.array_equal(b.populations(c), e2.populations())
