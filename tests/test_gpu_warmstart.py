"""What the trained warm-start network gains over the cold start.

Uses the network trained on the GPU by `python -m lbflow_amd.surrogate`
(16384 solver-generated 64^3 sphere-pack crops, seeds 10000+i), stored in
lbflow_amd/weights/warm64.npz.  On unseen crops (seeds 1000000+i, a range
no training set reaches) the same
batched solver runs to the same stop rule (Simulator::updateConvergence,
simulator.cpp:248-325, tolerance 1e-3) from rest and from the network's
prediction; the warm runs must need fewer iterations and less wall clock
(inference included), by the floors below, and their
permeability must be as close to the tightly converged one (tolerance 1e-5)
as the cold runs' is: a stop at residual 1e-3 leaves K uncertain by about
1e-3 either way, so the warm K is checked against the tight K with the cold
K's own error as the yardstick.
"""
import sys
from pathlib import Path

import numpy as np
import pytest

from lbflow_amd import surrogate as S

pytestmark = pytest.mark.gpu

WEIGHTS = Path(S.__file__).resolve().parent / 'weights' / 'warm64.npz'


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.mark.skipif(not WEIGHTS.exists(), reason='no trained weights in lbflow_amd/weights')
def test_warm_start_beats_cold_start():
    # 32 crops: with 8 the network's inference (about 12 ms) is a large part
    # of a 30 ms batch, and the wall-clock ratio swings around its floor
    net, scaling = S.load(WEIGHTS, capacity=8)
    crops = S.make_crops(32, S.EVAL_SEED0, 64, workers=1)
    tau, force, tol, interval = 0.8, 1e-6, 1e-3, 50
    cold, t_cold, _ = S.run_crops(crops, tau, force, tol, interval, 60000)
    warm, t_warm, t_inf = S.run_crops(crops, tau, force, tol, interval, 60000, net, scaling, relax=5)
    tight, _, _ = S.run_crops(crops, tau, force, 1e-5, interval, 60000)
    sc = np.array([r['steps'] for r in cold])
    sw = np.array([r['steps'] for r in warm])
    kc = np.array([r['K'] for r in cold])
    kw = np.array([r['K'] for r in warm])
    kt = np.array([r['K'] for r in tight])
    ec, ew = np.abs(kc - kt) / kt, np.abs(kw - kt) / kt
    say('cold steps', sc, 'warm steps', sw)
    say('K error vs tight: cold', ec, 'warm', ew)
    say('cold s', t_cold, 'warm s', t_warm, 'inference', t_inf)
    assert all(r['converged'] for r in cold + warm + tight)
    # measured on these 32 crops: 1.96x fewer iterations, 1.25x less wall
    # clock with the inference and the relaxation sweeps counted
    assert sc.sum() / sw.sum() >= 1.6
    assert t_cold / t_warm >= 1.1
    assert ew.mean() <= 1.5 * ec.mean() + 1e-4
    assert ew.max() < 2e-2
