"""Deterministic synthetic porous media for the LBFlow benchmark cases.

Periodic overlapping-sphere packs: sphere centres uniform in the box, radii
uniform in [rmin, rmax], spheres added one at a time until the pore fraction
drops to the target porosity; distances use the periodic minimum image and
voxel centres at integer + 0.5.  Output is a uint8 label volume in the
reference's raw layout (x fastest, then y, then z; solver/src/file/fileraw.h),
pore = 75 and solid = 255 by default as in the reference's examples.
"""
import hashlib

import numpy as np


def sphere_pack(shape_zyx, porosity, rmin, rmax, seed, pore=75, solid=255):
    nz, ny, nx = (int(n) for n in shape_zyx)
    rng = np.random.default_rng(seed)
    mask = np.zeros((nz, ny, nx), dtype=bool)
    total = nz * ny * nx
    nsolid = 0
    target = (1.0 - porosity) * total
    while nsolid < target:
        c = rng.uniform(0.0, 1.0, 3) * (nz, ny, nx)
        r = rng.uniform(rmin, rmax)
        lo = np.floor(c - r - 0.5).astype(int)
        hi = np.ceil(c + r - 0.5).astype(int)
        iz = np.arange(lo[0], hi[0] + 1)
        iy = np.arange(lo[1], hi[1] + 1)
        ix = np.arange(lo[2], hi[2] + 1)
        dz = (iz + 0.5 - c[0])[:, None, None]
        dy = (iy + 0.5 - c[1])[None, :, None]
        dx = (ix + 0.5 - c[2])[None, None, :]
        inside = dz * dz + dy * dy + dx * dx <= r * r
        # periodic images: where the sphere is wider than an axis (a single
        # z-plane in 2D) several of its indices wrap onto one site, so the
        # covered sites are collected as a set before marking them
        flat = ((iz % nz)[:, None, None] * ny + (iy % ny)[None, :, None]) * nx + (ix % nx)[None, None, :]
        covered = np.unique(flat[inside])
        flat_mask = mask.reshape(-1)
        nsolid += int(np.count_nonzero(~flat_mask[covered]))
        flat_mask[covered] = True
    labels = np.full((nz, ny, nx), pore, dtype=np.uint8)
    labels[mask] = solid
    return labels


def porosity(labels, pore_values=(75,)):
    return float(np.isin(labels, pore_values).mean())


def sha256(labels):
    return hashlib.sha256(np.ascontiguousarray(labels, dtype=np.uint8).tobytes()).hexdigest()


def random_blocks(labels, block, fraction, seed, phase0=75, phase1=76, pore=75):
    """Two-phase initial state: pore voxels get phase0/phase1 in random cubic blocks."""
    nz, ny, nx = labels.shape
    rng = np.random.default_rng(seed)
    bz, by, bx = -(-nz // block), -(-ny // block), -(-nx // block)
    choice = rng.uniform(size=(bz, by, bx)) < fraction
    grid = np.repeat(np.repeat(np.repeat(choice, block, 0), block, 1), block, 2)[:nz, :ny, :nx]
    out = labels.copy()
    poremask = labels == pore
    out[poremask & grid] = phase0
    out[poremask & ~grid] = phase1
    return out


def _batch_crop(args):
    i, seed0, side, phi, rmin, rmax = args
    rng = np.random.default_rng(seed0 + i)
    s = int(rng.integers(side[0], side[1] + 1))
    p = float(rng.uniform(phi[0], phi[1]))
    return sphere_pack((s, s, s), p, rmin, rmax, seed0 + i), p


def batch_crops(n=512, seed0=1000, side=(32, 96), phi=(0.15, 0.35), rmin=4.0, rmax=10.0, workers=1):
    """Independent cubic sphere-pack crops: crop i draws its side uniformly from
    the integers side[0]..side[1] and its porosity uniformly from phi with
    seed seed0 + i, then packs spheres (radius uniform in [rmin, rmax]) with
    that seed.  Returns (list of uint8 volumes, list of target porosities).
    workers > 1 uses a forked process pool (call before the GPU is used)."""
    args = [(i, seed0, side, phi, rmin, rmax) for i in range(n)]
    if workers > 1 and n > 8:
        from multiprocessing import get_context
        with get_context('fork').Pool(workers) as pool:
            out = pool.map(_batch_crop, args, chunksize=max(1, n // (workers * 4)))
    else:
        out = [_batch_crop(a) for a in args]
    return [c for c, _ in out], [p for _, p in out]
