"""lbf_report_static, the report entry point that allocates nothing, against lbf_report.

Two engines on the same lattice take the same steps; one reports through
lbf_report (the context's reduction buffers), the other through
lbf_report_static (the library's static device array).  Same kernels, same
block partition, so every sum must be bitwise equal.  The static array keeps
the previous call's contents, so the second lattice is smaller than the
first: a stale partial would show up in its sums.
"""
import numpy as np
import pytest

from lbflow_amd import cases as CS
from lbflow_amd.engine import STEP_ADJACENT, STEP_FIELDS, FP32, FP64, Engine
from tests.helpers import sphere_case

pytestmark = pytest.mark.gpu


def sums(r):
    return np.array([r.vel_range, *r.rho_range, r.rho_slice[0][0], r.rho_slice[0][1], r.rho_slice[1][0],
                     r.rho_slice[1][1], *r.count_slice, *r.conv, *r.phase_change, r.nonfinite, r.has_previous])


def pair(spec, precision):
    case = CS.lattice(spec)
    return [Engine(case, CS.geometry(spec), precision, stream_solids=precision == FP64).init() for _ in range(2)]


@pytest.mark.parametrize('precision', [FP64, FP32], ids=['fp64', 'fp32'])
def test_report_static_equals_report(precision):
    big = sphere_case(n=(40, 44, 48), model='color', eq='MRT', tau=(1.0, 1.0), two_phase=True)
    small = sphere_case(n=(12, 14, 16), model='scsp', eq='MRT', tau=(0.8,))
    for spec in (big, small):          # the second call's lattice is the smaller one
        a, b = pair(spec, precision)
        axis = 2
        last = CS.lattice(spec)['size'][axis] - 2
        for k in range(3):             # reports at three steps: convergence history in use
            for e in (a, b):
                e.advance(7)
                e.step(STEP_ADJACENT)
                e.step(STEP_FIELDS)
            ra = a.report(axis, 1, last, convergence=True, adjacent=k > 0)
            rb = b.report(axis, 1, last, convergence=True, adjacent=k > 0, static=True)
            sa, sb = sums(ra), sums(rb)
            assert np.array_equal(sa.view(np.uint64), sb.view(np.uint64)), (spec['name'], k, sa, sb)
        a.close()
        b.close()
