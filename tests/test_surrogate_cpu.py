"""CPU checks of the warm-start network's reference (no GPU needed).

The reference network in oracle/surrogate_ref.py is the checker of the
device network; these tests pin it: the PyTorch circular convolution with
the packed weight layout against a brute-force loop convolution, and its
autograd gradients against central finite differences of its own forward.
"""
import numpy as np
import pytest
import torch

from oracle import surrogate_ref as R


@pytest.mark.parametrize('cin,cout,taps', [(8, 32, 27), (16, 8, 27), (32, 4, 1)])
def test_packed_conv_matches_bruteforce(cin, cout, taps):
    rng = np.random.default_rng(cin * 100 + cout)
    s, n = 4, 2
    x = rng.standard_normal((n, s, s, s, cin))
    w = rng.standard_normal((cout, taps, cin))
    b = rng.standard_normal(cout)
    want = R.conv_bruteforce(x, w, b)
    k = 3 if taps == 27 else 1
    wt = torch.from_numpy(np.ascontiguousarray(w.reshape(cout, k, k, k, cin).transpose(0, 4, 1, 2, 3)))
    got = R.conv(torch.from_numpy(x.transpose(0, 4, 1, 2, 3).copy()), wt, torch.from_numpy(b), False, False)
    got = got.numpy().transpose(0, 2, 3, 4, 1)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-11)


def test_tap_order_is_dz_dy_dx():
    """A single tap picks the neighbour at (dz, dy, dx) = (tap/9-1, tap/3%3-1, tap%3-1)."""
    s = 4
    x = np.zeros((1, s, s, s, 1))
    x[0, 1, 2, 3, 0] = 1.0
    for tap in (0, 5, 13, 21, 26):
        w = np.zeros((1, 27, 1))
        w[0, tap, 0] = 1.0
        out = R.conv_bruteforce(x, w, [0.0])
        dz, dy, dx = tap // 9 - 1, (tap // 3) % 3 - 1, tap % 3 - 1
        # out[v] = x[v + d] is 1 where v + d = (1, 2, 3)
        z, y, xx = (1 - dz) % s, (2 - dy) % s, (3 - dx) % s
        assert out[0, z, y, xx, 0] == 1.0 and out.sum() == 1.0


def test_parameter_count():
    assert R.param_count() == 5842180


def test_reference_gradients_match_finite_differences():
    rng = np.random.default_rng(7)
    s, n = 8, 1
    mask = (rng.uniform(size=(n, s, s, s)) < 0.6).astype(np.float64)
    target = rng.standard_normal((n, s ** 3, 4)) * 0.1
    flat = np.concatenate([rng.standard_normal(co * (t * ci + 1)) * np.sqrt(2.0 / (t * ci))
                           for ci, co, t in R.LAYERS])
    w4 = np.array([1.0, 1.0, 2.0, 0.5])
    L0, g = R.gradients(mask, target, flat, w4)
    # central differences on parameters of the first, a middle and the head layer
    offsets = np.cumsum([0] + [co * (t * ci + 1) for ci, co, t in R.LAYERS])
    picks = [offsets[0] + 5, offsets[7] + 1234, offsets[12] + 77, offsets[14] + 3, offsets[15] - 1]
    for i in picks:
        h = 1e-6
        fp, fm = flat.copy(), flat.copy()
        fp[i] += h
        fm[i] -= h
        lp = float(R.loss(R.forward(mask, R.unpack(fp)), target, w4))
        lm = float(R.loss(R.forward(mask, R.unpack(fm)), target, w4))
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-9, (i, fd, g[i])
