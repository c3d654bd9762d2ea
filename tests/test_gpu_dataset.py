"""The warm-start training set on the device (lbf_dataset_*, lbf_net_train_samples).

- Target scaling: the fitted RMS equals a float64 host sum over pore voxels
  (to summation order), and the normalised targets equal the same FP32
  operations done on the host, bit for bit.
- Every one of the 16 lattice symmetries gathers exactly what the host
  statement of the mapping (surrogate.symmetry) gives, masks and targets,
  bit for bit (the CPU test pins that statement against the oracle).
- A training step on dataset samples equals a training step on the same
  samples mapped on the host and uploaded: loss and parameters bit for bit.
- Steady fields taken over from a solved batch equal lbf_batch_crop_fields.
- Bad indices, symmetries and sides are refused.
"""
import sys

import numpy as np
import pytest

from lbflow_amd import geometry as G
from lbflow_amd import surrogate as S
from lbflow_amd.engine import Batch, Buffer, EngineError

pytestmark = pytest.mark.gpu

SIDE, N = 16, 5
TAU, FORCE = 0.8, 1e-5


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def raw_data(seed=0):
    rng = np.random.default_rng(seed)
    masks = np.stack([(G.sphere_pack((SIDE,) * 3, 0.5, 1.5, 3.0, 400 + i) == 75) for i in range(N)])
    masks = masks.astype(np.uint8).reshape(N, -1)
    fields = np.empty((N, SIDE ** 3, 4), np.float32)
    fields[..., :3] = rng.standard_normal((N, SIDE ** 3, 3)) * 2e-3
    fields[..., 3] = 1.0 + rng.standard_normal((N, SIDE ** 3)) * 3e-5
    fields *= masks[..., None]
    fields[..., 3] += (1 - masks)   # rho 1 on solid voxels, as the batch writes 0 for u only
    return masks, fields


def host_normalise(masks, fields, rms):
    nu = (TAU - 0.5) / 3.0
    f = fields.astype(np.float32)
    out = np.empty_like(f)
    a = np.float32(nu / FORCE)
    for c in range(3):
        out[..., c] = (f[..., c] * a) / np.float32(rms[c]) * masks.astype(np.float32)
    out[..., 3] = ((f[..., 3] - np.float32(1.0)) / np.float32(FORCE)) / np.float32(rms[3]) * masks.astype(np.float32)
    return out


@pytest.fixture(scope='module')
def data():
    masks, fields = raw_data()
    d = S.Dataset(SIDE, N)
    d.set_masks(0, masks)
    d.set_targets(0, fields)
    rms = d.normalise(TAU, FORCE)
    return d, masks, fields, rms


def test_scaling_fit_and_normalise(data):
    d, masks, fields, rms = data
    nu = (TAU - 0.5) / 3.0
    pore = masks.reshape(-1).astype(bool)
    raw = fields.reshape(-1, 4)[pore].astype(np.float64)
    raw[:, :3] *= nu / FORCE
    raw[:, 3] = (raw[:, 3] - 1.0) / FORCE
    want = np.sqrt((raw ** 2).mean(axis=0))
    say('rms', rms, want)
    assert np.allclose(rms, want, rtol=1e-12, atol=0)
    targets = host_normalise(masks, fields, rms)
    for i in range(N):
        m, t = d.sample(i)
        assert np.array_equal(m, masks[i])
        assert np.array_equal(t.view(np.uint32), targets[i].view(np.uint32))


@pytest.mark.parametrize('sym', range(16))
def test_symmetry_gather_equals_host_mapping(data, sym):
    d, masks, fields, rms = data
    targets = host_normalise(masks, fields, rms)
    for i in (0, 3):
        m, t = d.sample(i, sym)
        wm, wt = S.symmetry(masks[i:i + 1], targets[i:i + 1], sym, rms)
        assert np.array_equal(m, wm[0])
        assert np.array_equal(t.view(np.uint32), wt[0].view(np.uint32))


def test_train_samples_equals_train_step(data):
    d, masks, fields, rms = data
    targets = host_normalise(masks, fields, rms)
    a = S.Net(side=SIDE, capacity=3, seed=9)
    b = S.Net(side=SIDE, capacity=3, seed=9)
    opt = S.adam(2e-3)
    for idx, sym in (([0, 2, 4], 5), ([1, 3, 0], 12), ([4, 4, 1], 0)):
        la = a.train_samples(d, idx, sym, opt)
        wm, wt = S.symmetry(masks[idx], targets[idx], sym, rms)
        lb = b.train_step(wm, wt, opt)
        say('loss', sym, la, lb)
        assert la == lb
    assert np.array_equal(a.params().view(np.uint32), b.params().view(np.uint32))


def test_take_fields_equals_crop_fields():
    crops = [G.sphere_pack((SIDE,) * 3, 0.45, 1.5, 3.0, 500 + i) for i in range(3)]
    d, res = S.solve(crops, tau=TAU, force=FORCE, tol=1e-4, interval=40, max_steps=20000)
    assert all(r['converged'] for r in res)
    b = Batch(3, 19, 'MRT', TAU, 1.0, S.labels(FORCE), crops).init()
    b.run(axis=2, tol=1e-4, interval=40, max_steps=20000)
    L = S.lib()
    for c in range(len(crops)):
        buf = Buffer((SIDE ** 3, 4), np.float32)
        assert L.lbf_batch_crop_fields(b.h, c, buf.ptr) == 0
        m, t = d.sample(c)
        assert np.array_equal(m, (crops[c] == 75).reshape(-1).astype(np.uint8))
        assert np.array_equal(t.view(np.uint32), buf.numpy().view(np.uint32))
    b.close()


def test_refusals(data):
    d, _, _, _ = data
    with pytest.raises(EngineError, match='index outside'):
        d.sample(N, 0)
    with pytest.raises(EngineError, match='symmetry outside'):
        d.sample(0, 16)
    with pytest.raises(EngineError, match='normalised already'):
        d.normalise(TAU, FORCE)
    net = S.Net(side=2 * SIDE, capacity=2, seed=1)
    with pytest.raises(EngineError, match="network's side"):
        net.train_samples(d, [0, 1], 0, S.adam(1e-3))


def test_pipeline_end_to_end():
    """Generate, solve, scale, train (with symmetry augmentation) and evaluate
    at a small size, all on the engine's runtime."""
    rep = S.pipeline(n_train=8, n_eval=8, side=16, epochs=4, capacity=4, lr=2e-3, tau=0.8, force=1e-5,
                     train_force=1e-5, train_tol=1e-4, eval_tol=1e-3, interval=20, max_steps=20000, log=say)
    # at 16^3 the benchmark's sphere radii (4-10) close some crops: no flowing
    # cluster, no steady state to converge to, in the cold and the warm runs alike
    assert rep['training_data']['converged'] >= 4
    loss = rep['training']['loss']
    assert np.all(np.isfinite(loss)) and loss[-1] < loss[0]
    ev = rep['evaluation']
    assert ev['warm_relax5']['converged'] == ev['cold']['converged'] >= 4
