"""CPU reference of the warm-start network (test infrastructure only).

The learned warm-start has no counterpart in the reference; this module is
the checker of the build's own BF16-MFMA network: the same periodic 3D U-Net
written with PyTorch CPU operators in float64 (autograd supplies the
gradients), plus a brute-force loop convolution that pins the tap order and
the periodic wrap on tiny shapes.  Only tests/ import it.

Conventions shared with the device network (lbflow_amd/csrc/surrogate.cpp):
* volumes are [sample][z][y][x] with a power-of-two side, periodic;
* a 3x3x3 weight is packed W[cout][tap][cin], tap = (dz+1)*9 + (dy+1)*3 + (dx+1),
  and out[v] = sum_tap sum_ci W[co][tap][ci] * in[v + (dz,dy,dx)][ci] + b[co];
* layers: 2 convolutions per level (ReLU), 2x average pooling, nearest 2x
  upsampling, skip connections concatenated as [upsampled, skip], and a
  linear 1x1x1 head with 4 outputs; input channel 0 = fluid indicator,
  channels 1..7 = 0.
"""
import numpy as np
import torch
import torch.nn.functional as F

# cin, cout, taps per layer, in order (the head has 4 real outputs)
LAYERS = [(8, 32, 27), (32, 32, 27), (32, 64, 27), (64, 64, 27), (64, 128, 27), (128, 128, 27),
          (128, 256, 27), (256, 256, 27), (384, 128, 27), (128, 128, 27), (192, 64, 27), (64, 64, 27),
          (96, 32, 27), (32, 32, 27), (32, 4, 1)]


def param_count():
    return sum(co * (t * ci + 1) for ci, co, t in LAYERS)


def unpack(flat):
    """Packed parameter vector -> list of (torch weight [cout][cin][k][k][k], bias [cout]) in float64."""
    flat = np.asarray(flat, dtype=np.float64)
    out, o = [], 0
    for ci, co, t in LAYERS:
        w = flat[o:o + co * t * ci].reshape(co, t, ci)
        o += co * t * ci
        b = flat[o:o + co]
        o += co
        k = 3 if t == 27 else 1
        wt = torch.from_numpy(np.ascontiguousarray(w.reshape(co, k, k, k, ci).transpose(0, 4, 1, 2, 3)))
        out.append((wt, torch.from_numpy(b.copy())))
    assert o == flat.size
    return out


def pack_grads(params):
    """Gradients of unpack()'s tensors back into the packed layout."""
    parts = []
    for (w, b), (ci, co, t) in zip(params, LAYERS):
        g = w.grad.detach().numpy().transpose(0, 2, 3, 4, 1).reshape(co, t * ci)
        parts.append(g.ravel())
        parts.append(b.grad.detach().numpy().ravel())
    return np.concatenate(parts)


def bf16(x):
    """Round to bfloat16 (nearest even) and back, keeping the dtype."""
    return x.to(torch.bfloat16).to(x.dtype)   # autograd rounds the gradient through it as well


def conv(x, w, b, relu, round_bf16):
    """Periodic convolution of x [n][c][z][y][x]."""
    if w.shape[-1] == 3:
        x = F.pad(x, (1, 1, 1, 1, 1, 1), mode='circular')
    y = F.conv3d(x, w, b)
    if relu:
        y = torch.relu(y)
    return bf16(y) if round_bf16 else y


def forward(mask, params, round_bf16=False):
    """mask [n][S][S][S] (1 = fluid) -> output [n][S^3][4] (float64 tensor)."""
    m = torch.as_tensor(np.asarray(mask, dtype=np.float64))
    n, s = m.shape[0], m.shape[1]
    x0 = torch.zeros((n, 8, s, s, s), dtype=torch.float64)
    x0[:, 0] = m
    P = params
    if round_bf16:
        P = [(bf16(w), b) for w, b in params]
    r = round_bf16

    def pool(x):
        y = F.avg_pool3d(x, 2)
        return bf16(y) if r else y

    def up(x):
        return x.repeat_interleave(2, 2).repeat_interleave(2, 3).repeat_interleave(2, 4)

    def cat(a, b):
        # the device stores the gradient of a concatenation (a convolution's input
        # gradient) in BF16; bf16() above rounds gradients only where it is applied
        c = torch.cat([a, b], 1)
        if r and c.requires_grad:
            c.register_hook(bf16)
        return c

    e1 = conv(conv(x0, *P[0], True, r), *P[1], True, r)
    e2 = conv(conv(pool(e1), *P[2], True, r), *P[3], True, r)
    e3 = conv(conv(pool(e2), *P[4], True, r), *P[5], True, r)
    e4 = conv(conv(pool(e3), *P[6], True, r), *P[7], True, r)
    d3 = conv(conv(cat(up(e4), e3), *P[8], True, r), *P[9], True, r)
    d2 = conv(conv(cat(up(d3), e2), *P[10], True, r), *P[11], True, r)
    d1 = conv(conv(cat(up(d2), e1), *P[12], True, r), *P[13], True, r)
    out = conv(d1, *P[14], False, False)
    if r and out.requires_grad:
        out.register_hook(bf16)   # the loss gradient is stored in BF16 too
    return out.permute(0, 2, 3, 4, 1).reshape(n, s * s * s, 4)


def loss(out, target, weights):
    """mean over all sites of sum_c w_c (out_c - target_c)^2."""
    t = torch.as_tensor(np.asarray(target, dtype=np.float64)).reshape(out.shape)
    w = torch.as_tensor(np.asarray(weights, dtype=np.float64))
    n_sites = out.shape[0] * out.shape[1]
    return (((out - t) ** 2) * w).sum() / n_sites


def gradients(mask, target, flat, weights, round_bf16=False):
    """(loss, packed gradient) of the float64 network at the packed parameters."""
    params = unpack(flat)
    for w, b in params:
        w.requires_grad_(True)
        b.requires_grad_(True)
    out = forward(mask, params, round_bf16)
    L = loss(out, target, weights)
    L.backward()
    return float(L.detach()), pack_grads(params)


def conv_bruteforce(x, w, b):
    """Loop form of one periodic convolution, x [n][S][S][S][cin] (channels last),
    w packed [cout][taps][cin]; returns [n][S][S][S][cout] (float64).  Small shapes only."""
    x = np.asarray(x, dtype=np.float64)
    n, s = x.shape[0], x.shape[1]
    cout, taps, cin = w.shape
    out = np.zeros((n, s, s, s, cout))
    for t in range(taps):
        if taps == 27:
            dz, dy, dx = t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1
        else:
            dz = dy = dx = 0
        shifted = np.roll(x, shift=(-dz, -dy, -dx), axis=(1, 2, 3))   # shifted[z] = x[z + dz]
        out += np.einsum('nzyxc,oc->nzyxo', shifted, w[:, t, :])
    return out + np.asarray(b, dtype=np.float64)
