"""Time-to-permeability on the GPU: FP64 vs FP32, and the tube-bundle closed form.

The slit bundle's Darcy permeability has the midpoint closed form
K = sum_i (h_i^3/12 + h_i/24) / n_y (scripts/porous_media.py:111-126); the
reference itself meets it to 2e-6.
"""
import sys

import numpy as np
import pytest

from lbflow_amd import cases as CS
from tests import cases as TC
from lbflow_amd.engine import FP32, FP64, Engine
from lbflow_amd.permeability import converge
from tests.helpers import sphere_case

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.mark.parametrize('precision', [FP64, FP32], ids=['fp64', 'fp32'])
def test_bundle_permeability_closed_form(precision):
    spec = TC.reference('bundle', scale=2, dimensions=3, model='scsp')[0]
    case = CS.lattice(spec)
    eng = Engine(case, CS.geometry(spec), precision, stream_solids=False).init()
    nx = spec['size'][0]
    # FP32 residuals level off near 1e-7, so its stop tolerance is the FP32 one
    tol = 1e-9 if precision == FP64 else 1e-6
    res = converge(eng, axis=2, first=1, last=nx - 2, tol=tol, interval=500, max_steps=40000,
                   force=spec['force'], tau=spec['tau'][0])
    widths, ny = spec['widths'], spec['size'][1]
    expected = sum(h ** 3 / 12 + h / 24 for h in widths) / ny
    say(precision, res['steps'], res['K'], expected)
    assert res['converged']
    assert abs(res['K'] / expected - 1) < 2e-6


def test_fp32_permeability_within_1e5_of_fp64():
    spec = sphere_case(n=(32, 32, 32), model='scsp', eq='MRT', tau=(0.8,), porosity=0.3, seed=5)
    spec = dict(spec, indices=['75 "GV" "1e-06" "1" "0"', '255 "SD" "90" "N/A" "N/A"'])
    case = CS.lattice(spec)
    out = {}
    for prec in (FP64, FP32):
        eng = Engine(case, CS.geometry(spec), prec, stream_solids=False).init()
        out[prec] = converge(eng, axis=2, first=1, last=30, tol=1e-6, interval=200, max_steps=60000,
                             force=1e-6, tau=0.8)
        say(prec, out[prec]['steps'], out[prec]['K'])
        assert out[prec]['converged']
    rel = abs(out[FP32]['K'] - out[FP64]['K']) / out[FP64]['K']
    say('relative K difference', rel)
    assert rel < 1e-5
