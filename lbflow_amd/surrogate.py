"""Learned warm-start for single-phase permeability runs.

A periodic 3D U-Net (lbf_net_* in include/lbflow.h; BF16 MFMA kernels with
FP32 accumulation) maps a crop's fluid mask to its steady velocity and
density.  Its prediction initialises the populations at equilibrium
(lbf_batch_set_fields), and the plain lattice Boltzmann iteration then runs
to the unchanged stop rule (Simulator::updateConvergence,
solver/src/lbm/simulator.cpp:248-325), which alone decides convergence: a
poor prediction costs iterations, never correctness.

Pipeline: generate sphere-pack crops -> solve them to steady state with the
batched engine -> train the network on (mask, fields) -> evaluate cold vs
warm starts on unseen crops.  A warm start (run_crops) runs entirely in
the engine's library: masks go in from host memory and the prediction stays
in the network's device buffers (lbf_batch_warm_start).  Training runs there
too: the solved crops' steady fields move from the batch into a device-side
training set (lbf_dataset_*), which scales them into targets and gathers
each minibatch, mapped by a random lattice symmetry, into the buffers the
network trains on.  No other GPU runtime is involved.
"""
import ctypes as C
import json
import time
from multiprocessing import get_context

import numpy as np

from . import geometry as G
from .engine import Batch, Buffer, EngineError, load_library

PORE, SOLID = 75, 255


class Adam(C.Structure):
    _fields_ = [('lr', C.c_float), ('beta1', C.c_float), ('beta2', C.c_float), ('eps', C.c_float),
                ('weights', C.c_float * 4)]


def _bind(L):
    if getattr(L, '_net_bound', False):
        return L
    vp, u32, u64, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
    L.lbf_net_create.argtypes = [u32, u32, u64, C.POINTER(vp)]
    L.lbf_net_create.restype = i32
    L.lbf_net_destroy.argtypes = [vp]
    L.lbf_net_error.argtypes = [vp]
    L.lbf_net_error.restype = C.c_char_p
    L.lbf_net_param_count.argtypes = [vp]
    L.lbf_net_param_count.restype = u64
    L.lbf_net_get_params.argtypes = [vp, vp]
    L.lbf_net_get_params.restype = i32
    L.lbf_net_set_params.argtypes = [vp, vp]
    L.lbf_net_set_params.restype = i32
    L.lbf_net_forward.argtypes = [vp, vp, u32, vp]
    L.lbf_net_forward.restype = i32
    L.lbf_net_gradients.argtypes = [vp, vp, vp, u32, vp, C.POINTER(C.c_double), vp]
    L.lbf_net_gradients.restype = i32
    L.lbf_net_train_step.argtypes = [vp, vp, vp, u32, C.POINTER(Adam), C.POINTER(C.c_double)]
    L.lbf_net_train_step.restype = i32
    L.lbf_net_timed_train.argtypes = [vp, vp, vp, u32, C.POINTER(Adam), u32, C.POINTER(C.c_double),
                                      C.POINTER(C.c_double)]
    L.lbf_net_timed_train.restype = i32
    L.lbf_conv3d_forward.argtypes = [i32, i32, i32, u32, u32, vp, u32, vp, vp, vp, u32, i32, i32]
    L.lbf_conv3d_forward.restype = i32
    L.lbf_conv3d_wgrad.argtypes = [i32, i32, i32, u32, u32, vp, u32, vp, u32, u32, vp]
    L.lbf_conv3d_wgrad.restype = i32
    L.lbf_batch_crop_fields.argtypes = [vp, u32, vp]
    L.lbf_batch_crop_fields.restype = i32
    L.lbf_batch_set_fields.argtypes = [vp, u32, vp]
    L.lbf_batch_set_fields.restype = i32
    L.lbf_batch_relax.argtypes = [vp, u32]
    L.lbf_batch_relax.restype = i32
    L.lbf_batch_warm_start.argtypes = [vp, vp, vp, u32, vp, vp, u32]
    L.lbf_batch_warm_start.restype = i32
    L.lbf_net_side.argtypes = [vp]
    L.lbf_net_side.restype = u32
    L.lbf_dataset_create.argtypes = [u32, u32, C.POINTER(vp)]
    L.lbf_dataset_create.restype = i32
    L.lbf_dataset_destroy.argtypes = [vp]
    L.lbf_dataset_error.argtypes = [vp]
    L.lbf_dataset_error.restype = C.c_char_p
    for name in ('lbf_dataset_set_masks', 'lbf_dataset_set_targets'):
        getattr(L, name).argtypes = [vp, u32, u32, vp]
        getattr(L, name).restype = i32
    L.lbf_dataset_take_fields.argtypes = [vp, u32, vp, u32, u32]
    L.lbf_dataset_take_fields.restype = i32
    L.lbf_dataset_normalise.argtypes = [vp, C.c_double, C.c_double, i32, vp]
    L.lbf_dataset_normalise.restype = i32
    L.lbf_dataset_sample.argtypes = [vp, u32, u32, vp, vp]
    L.lbf_dataset_sample.restype = i32
    L.lbf_dataset_device.argtypes = [vp, C.POINTER(vp), C.POINTER(vp)]
    L.lbf_dataset_device.restype = i32
    L.lbf_net_train_samples.argtypes = [vp, vp, vp, u32, u32, C.POINTER(Adam), C.POINTER(C.c_double)]
    L.lbf_net_train_samples.restype = i32
    L._net_bound = True
    return L


def lib():
    return _bind(load_library())


class _View:
    """Device memory owned elsewhere (a Dataset's samples): address and shape."""

    def __init__(self, ptr, shape):
        self.ptr, self.shape = ptr, tuple(shape)


def _dev(a, dtype, shape=None):
    """Device memory (a Buffer or a _View) as is, or host data uploaded into a
    new Buffer."""
    if isinstance(a, (Buffer, _View)):
        return a
    a = np.ascontiguousarray(a, dtype=dtype)
    return Buffer(a.shape if shape is None else shape, dtype, a)


class Net:
    """The warm-start network on the GPU.  forward, gradients and train_step
    take host arrays (uploaded per call) or device Buffers (masks uint8
    [n, S^3], targets float32 [n, S^3, 4]); training from a Dataset
    (train_samples) and warm starts (run_crops) keep everything on the device."""

    def __init__(self, side=64, capacity=4, seed=1):
        self.L = lib()
        h = C.c_void_p()
        if self.L.lbf_net_create(side, capacity, seed, C.byref(h)):
            raise EngineError(self.L.lbf_net_error(None).decode())
        self.h = h
        self.side, self.capacity = side, capacity
        self.n_params = int(self.L.lbf_net_param_count(h))

    def close(self):
        if getattr(self, 'h', None):
            self.L.lbf_net_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc:
            raise EngineError(self.L.lbf_net_error(self.h).decode())

    def params(self):
        out = np.zeros(self.n_params, np.float32)
        self._check(self.L.lbf_net_get_params(self.h, out.ctypes.data_as(C.c_void_p)))
        return out

    def set_params(self, flat):
        flat = np.ascontiguousarray(flat, dtype=np.float32)
        assert flat.size == self.n_params
        self._check(self.L.lbf_net_set_params(self.h, flat.ctypes.data_as(C.c_void_p)))

    def forward(self, mask, out=None):
        """Network output (u_z, u_y, u_x, rho channels, normalised) of masks
        [n, S^3]; into the device Buffer `out` [n, S^3, 4] when given, else
        returned as a host array."""
        m = _dev(mask, np.uint8)
        n = m.shape[0]
        o = out if out is not None else Buffer((n, self.side ** 3, 4), np.float32)
        self._check(self.L.lbf_net_forward(self.h, m.ptr, n, o.ptr))
        return None if out is not None else o.numpy()

    def gradients(self, mask, target, weights=(1, 1, 1, 1)):
        m, t = _dev(mask, np.uint8), _dev(target, np.float32)
        w = np.ascontiguousarray(weights, dtype=np.float32)
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_double()
        self._check(self.L.lbf_net_gradients(self.h, m.ptr, t.ptr, m.shape[0], w.ctypes.data_as(C.c_void_p),
                                             C.byref(loss), g.ctypes.data_as(C.c_void_p)))
        return loss.value, g

    def train_step(self, mask, target, opt):
        m, t = _dev(mask, np.uint8), _dev(target, np.float32)
        loss = C.c_double()
        self._check(self.L.lbf_net_train_step(self.h, m.ptr, t.ptr, m.shape[0], C.byref(opt), C.byref(loss)))
        return loss.value

    def train_samples(self, data, idx, sym, opt):
        """One Adam step on samples idx of a Dataset mapped by symmetry sym."""
        i = np.ascontiguousarray(idx, dtype=np.uint32)
        loss = C.c_double()
        if self.L.lbf_net_train_samples(self.h, data.h, i.ctypes.data_as(C.c_void_p), i.size, int(sym),
                                        C.byref(opt), C.byref(loss)):
            raise EngineError(data.error())
        return loss.value

    def timed_train(self, mask, target, opt, steps):
        m, t = _dev(mask, np.uint8), _dev(target, np.float32)
        ms, fl = C.c_double(), C.c_double()
        self._check(self.L.lbf_net_timed_train(self.h, m.ptr, t.ptr, m.shape[0], C.byref(opt), steps,
                                               C.byref(ms), C.byref(fl)))
        return ms.value, fl.value


def adam(lr=1e-3, weights=(1.0, 1.0, 1.0, 1.0), beta1=0.9, beta2=0.999, eps=1e-8):
    o = Adam(lr=lr, beta1=beta1, beta2=beta2, eps=eps)
    for i, w in enumerate(weights):
        o.weights[i] = w
    return o


# ---------------------------------------------------------------------------
# data
# ---------------------------------------------------------------------------
def _crop(args):
    seed, side, phi, rmin, rmax = args
    rng = np.random.default_rng(seed)
    p = float(rng.uniform(*phi))
    return G.sphere_pack((side, side, side), p, rmin, rmax, seed)


# Crop seeds: training crops are TRAIN_SEED0 + i, evaluation crops
# EVAL_SEED0 + i; pipeline() refuses a training set that would reach the
# evaluation seeds (crop i uses seed seed0 + i, so ranges must not overlap).
TRAIN_SEED0 = 10000
EVAL_SEED0 = 1000000


def make_crops(n, seed0, side=64, phi=(0.15, 0.35), rmin=4.0, rmax=10.0, workers=16):
    """n cubic sphere-pack crops of the given side; crop i uses seed seed0 + i for
    its porosity (uniform in phi) and its spheres.  Call before the GPU is used
    (the pool forks)."""
    args = [(seed0 + i, side, phi, rmin, rmax) for i in range(n)]
    if workers > 1 and n > 8:
        with get_context('fork').Pool(workers) as pool:
            return pool.map(_crop, args, chunksize=max(1, n // (workers * 4)))
    return [_crop(a) for a in args]


def labels(force):
    return {PORE: ('GV', force, 1, 0), SOLID: ('SD', 90.0, 0, 0)}


class Dataset:
    """Training set on the device (lbf_dataset_*): per sample a pore mask and
    (u_z, u_y, u_x, rho) per voxel, raw steady fields until normalise() turns
    them into targets."""

    def __init__(self, side, count):
        self.L = lib()
        h = C.c_void_p()
        if self.L.lbf_dataset_create(side, count, C.byref(h)):
            raise EngineError(self.L.lbf_dataset_error(None).decode())
        self.h = h
        self.side, self.count = side, count
        self.rms = None

    def error(self):
        return self.L.lbf_dataset_error(self.h).decode()

    def _check(self, rc):
        if rc:
            raise EngineError(self.error())

    def close(self):
        if getattr(self, 'h', None):
            self.L.lbf_dataset_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_masks(self, first, masks):
        m = np.ascontiguousarray(masks, dtype=np.uint8)
        self._check(self.L.lbf_dataset_set_masks(self.h, first, m.shape[0], m.ctypes.data_as(C.c_void_p)))

    def set_targets(self, first, targets):
        t = np.ascontiguousarray(targets, dtype=np.float32)
        self._check(self.L.lbf_dataset_set_targets(self.h, first, t.shape[0], t.ctypes.data_as(C.c_void_p)))

    def take_fields(self, first, batch, crop0, n):
        self._check(self.L.lbf_dataset_take_fields(self.h, first, batch.h, crop0, n))

    def normalise(self, tau, force, rms=None):
        """Raw fields -> targets in place; fits the per-channel RMS unless given.
        Returns the RMS."""
        r = np.zeros(4) if rms is None else np.ascontiguousarray(rms, dtype=np.float64)
        self._check(self.L.lbf_dataset_normalise(self.h, tau, force, int(rms is None), r.ctypes.data_as(C.c_void_p)))
        self.rms = r
        return r

    def sample(self, i, sym=0):
        """(mask [S^3] uint8, fields or targets [S^3, 4] float32) of sample i
        mapped by symmetry sym."""
        s3 = self.side ** 3
        m = np.empty(s3, np.uint8)
        t = np.empty((s3, 4), np.float32)
        self._check(self.L.lbf_dataset_sample(self.h, i, sym, m.ctypes.data_as(C.c_void_p),
                                              t.ctypes.data_as(C.c_void_p)))
        return m, t

    def device(self):
        """Device addresses of the masks and targets of sample 0."""
        m, t = C.c_void_p(), C.c_void_p()
        self._check(self.L.lbf_dataset_device(self.h, C.byref(m), C.byref(t)))
        return m, t


def solve(crops, tau=0.8, force=1e-5, tol=1e-5, interval=100, max_steps=40000, log=None, chunk=4096):
    """Steady fields of every crop with the batched engine, `chunk` crops per
    batch.  Returns (a Dataset holding the crops' pore masks and raw fields
    (u_z, u_y, u_x, rho), per-crop results)."""
    side = crops[0].shape[0]
    data = Dataset(side, len(crops))
    res = []
    for c0 in range(0, len(crops), chunk):
        part = crops[c0:c0 + chunk]
        data.set_masks(c0, np.stack([(c == PORE) for c in part]).astype(np.uint8).reshape(len(part), -1))
        b = Batch(3, 19, 'MRT', tau, 1.0, labels(force), part)
        b.tune(steps=4)
        b.init()
        r, sec = b.run(axis=2, tol=tol, interval=interval, max_steps=max_steps)
        res += r
        if log:
            log(f'solved crops {c0}-{c0 + len(part) - 1} in {sec:.1f} s, steps {min(x["steps"] for x in r)}-'
                f'{max(x["steps"] for x in r)}, converged {sum(x["converged"] for x in r)}')
        data.take_fields(c0, b, 0, len(part))
        b.close()
    return data, res


class Scaling:
    """Targets are (u * nu / F, (rho - rho0) / F) divided by their RMS over fluid sites."""

    def __init__(self, rms, tau, force):
        self.rms = np.asarray(rms, dtype=np.float64)
        self.nu = (tau - 0.5) / 3.0
        self.force = force

    @classmethod
    def fit(cls, data, tau, force):
        """Fits the RMS on a Dataset of raw fields and normalises it in place."""
        return cls(data.normalise(tau, force), tau, force)

    def fields(self, pred, tau, force, mask):
        """Network output [n, S^3, 4] -> physical (u, rho) for a run with
        (tau, force); velocities 0 on solid voxels (mask [n, S^3])."""
        nu = (tau - 0.5) / 3.0
        out = np.array(pred, dtype=np.float32)
        for c in range(4):
            out[..., c] *= np.float32(self.rms[c])
        out[..., :3] *= np.float32(force / nu)
        out[..., 3] = np.float32(1.0) + out[..., 3] * np.float32(force)
        out[..., :3] *= np.asarray(mask, dtype=np.float32).reshape(out.shape[0], -1, 1)
        return out

    def affine(self, tau, force):
        """(scale, offset) of Scaling.fields as per-channel factors, for
        lbf_batch_warm_start: f_c = p_c scale_c + offset_c."""
        nu = (tau - 0.5) / 3.0
        scale = [self.rms[c] * force / nu for c in range(3)] + [self.rms[3] * force]
        return np.asarray(scale, np.float32), np.asarray([0.0, 0.0, 0.0, 1.0], np.float32)

    def to_json(self):
        return dict(rms=self.rms.tolist(), tau=(self.nu * 3.0 + 0.5), force=self.force)

    @classmethod
    def from_json(cls, d):
        return cls(d['rms'], d['tau'], d['force'])


def symmetry(masks, targets, k, rms):
    """Host statement of the lattice symmetry the training set applies on the
    device (lbf_dataset_sample, lbf_net_train_samples), for masks [n, S^3]
    and normalised targets [n, S^3, 4] = (u_z, u_y, u_x, rho): bit 0 flips z
    (u_z -> -u_z), bit 1 flips y (u_y -> -u_y), bit 3 flips x (the mirrored
    field of the reversed force, negated: u_y, u_z and the density deviation
    change sign), then bit 2 swaps y and z (u_y and u_z swapped and carried
    across their RMS).  Targets are 0 on solid sites.  The periodic D3Q19
    lattice with halfway bounce-back is invariant under bits 0-2, so the field
    of the mapped mask is the mapped field exactly; bit 3 holds in the Stokes
    (linear) limit, to relative size ~1e-4 at the training force."""
    n = masks.shape[0]
    s = round(targets.shape[1] ** (1.0 / 3.0))
    m = np.asarray(masks).reshape(n, s, s, s)
    t = np.asarray(targets).reshape(n, s, s, s, 4)
    sign = np.ones(4, t.dtype)
    axes = []
    if k & 1:
        axes.append(1)
        sign[0] = -sign[0]
    if k & 2:
        axes.append(2)
        sign[1] = -sign[1]
    if k & 8:
        axes.append(3)
        sign[0], sign[1], sign[3] = -sign[0], -sign[1], -sign[3]
    if axes:
        m = np.flip(m, axes)
        t = np.flip(t, axes)
    t = t * sign
    if k & 4:
        m = m.transpose(0, 2, 1, 3)
        t = t.transpose(0, 2, 1, 3, 4)[..., [1, 0, 2, 3]]
        ratio = np.array([rms[1] / rms[0], rms[0] / rms[1], 1.0, 1.0], dtype=t.dtype)
        t = t * ratio
    return np.ascontiguousarray(m).reshape(n, -1), np.ascontiguousarray(t).reshape(n, -1, 4)


def train(net, data, epochs, lr=1e-3, batch=None, log=None, seed=0, lr_final=None, augment=False):
    """Adam over shuffled minibatches of a normalised Dataset; cosine
    learning-rate decay to lr_final.  With `augment` every minibatch is mapped
    by a random one of the 16 lattice symmetries."""
    batch = batch or net.capacity
    n = data.count
    rng = np.random.default_rng(seed)
    steps = epochs * (n // batch)
    lr_final = lr * 0.05 if lr_final is None else lr_final
    opt = adam(lr)
    k = 0
    history = []
    t0 = time.perf_counter()
    for ep in range(epochs):
        order = rng.permutation(n).astype(np.uint32)
        tot = 0.0
        for i in range(n // batch):
            sym = int(rng.integers(16)) if augment else 0
            opt.lr = lr_final + 0.5 * (lr - lr_final) * (1 + np.cos(np.pi * k / max(1, steps)))
            tot += net.train_samples(data, order[i * batch:(i + 1) * batch], sym, opt)
            k += 1
        history.append(tot / max(1, n // batch))
        if log:
            log(f'epoch {ep + 1}/{epochs} loss {history[-1]:.5f} ({time.perf_counter() - t0:.1f} s)')
    return history


def save(path, net, scaling, half=True):
    """Parameters (float16 when half, else float32), the target scaling and the side."""
    p = net.params()
    np.savez_compressed(path, params=p.astype(np.float16) if half else p, scaling=json.dumps(scaling.to_json()),
                        side=net.side)


def load(path, capacity=4, side=None):
    """Weights and target scaling; the network is fully convolutional and
    periodic, so it can be built for another cubic side than it was trained on."""
    d = np.load(path, allow_pickle=False)
    net = Net(side=int(d['side']) if side is None else side, capacity=capacity)
    net.set_params(d['params'].astype(np.float32))
    return net, Scaling.from_json(json.loads(str(d['scaling'])))


# ---------------------------------------------------------------------------
# warm vs cold
# ---------------------------------------------------------------------------
def run_crops(crops, tau, force, tol, interval, max_steps, net=None, scaling=None, relax=0):
    """Runs all crops to the stop rule in one batch, cold (rest state) or warm
    (network prediction, lbf_batch_warm_start).  Returns (results, seconds
    from init to the stop, inference seconds)."""
    b = Batch(3, 19, 'MRT', tau, 1.0, labels(force), crops)
    b.tune(steps=4)
    L = lib()
    t0 = time.perf_counter()
    b.init()
    t_inf = 0.0
    if net is not None:
        t1 = time.perf_counter()
        masks = np.ascontiguousarray(np.stack([(c == PORE) for c in crops]).astype(np.uint8))
        scale, offset = scaling.affine(tau, force)
        if L.lbf_batch_warm_start(b.h, net.h, masks.ctypes.data_as(C.c_void_p), len(crops),
                                  scale.ctypes.data_as(C.c_void_p), offset.ctypes.data_as(C.c_void_p), relax):
            raise EngineError(L.lbf_net_error(net.h).decode())
        t_inf = time.perf_counter() - t1
    res, _ = b.run(axis=2, tol=tol, interval=interval, max_steps=max_steps)
    sec = time.perf_counter() - t0
    b.close()
    return res, sec, t_inf


def pipeline(n_train=4096, n_eval=256, side=64, epochs=8, capacity=8, lr=1e-3, tau=0.8, force=1e-6,
             train_force=1e-5, train_tol=1e-5, eval_tol=1e-3, interval=50, max_steps=60000, out=None, log=print,
             relax_sweeps=(5,), init=None, augment=True):
    """Generate, solve, train (from `init` weights when given), evaluate.
    Returns a report dict."""
    rep = dict(side=side, n_train=n_train, n_eval=n_eval, epochs=epochs, capacity=capacity, lr=lr, tau=tau,
               force=force, train_force=train_force, train_tol=train_tol, eval_tol=eval_tol,
               report_interval=interval)
    t0 = time.perf_counter()
    if TRAIN_SEED0 + n_train > EVAL_SEED0:
        raise ValueError(f'{n_train} training crops would reach the evaluation seeds ({EVAL_SEED0}+)')
    train_crops = make_crops(n_train, TRAIN_SEED0, side)
    eval_crops = make_crops(n_eval, EVAL_SEED0, side)
    rep['crop_generation_s'] = round(time.perf_counter() - t0, 2)
    log(f'crops generated in {rep["crop_generation_s"]} s')
    t0 = time.perf_counter()
    data, res = solve(train_crops, tau=tau, force=train_force, tol=train_tol, interval=100, max_steps=max_steps,
                      log=log)
    rep['training_data'] = dict(seconds=round(time.perf_counter() - t0, 2),
                                converged=int(sum(r['converged'] for r in res)),
                                steps_mean=float(np.mean([r['steps'] for r in res])),
                                fluid_sites=int(sum(int((c == PORE).sum()) for c in train_crops)))
    del train_crops
    scaling = Scaling.fit(data, tau, train_force)
    net = Net(side=side, capacity=capacity, seed=1)
    masks, targets = data.device()
    ms, flops = net.timed_train(_View(masks, (capacity, side ** 3)), _View(targets, (capacity, side ** 3, 4)),
                                adam(0.0), 3)
    rep['train_step'] = dict(ms=round(ms, 3), tflops=round(flops / ms / 1e9, 1), samples_per_step=capacity,
                             conv_flops_per_step=flops)
    log(f'train step {ms:.2f} ms, {flops / ms / 1e9:.1f} TFLOP/s (convolutions)')
    net.close()   # the timing steps ran with lr 0; train a fresh network
    net = Net(side=side, capacity=capacity, seed=1)
    if init:
        net.set_params(np.load(init, allow_pickle=False)['params'].astype(np.float32))
        rep['init'] = str(init)
    t0 = time.perf_counter()
    hist = train(net, data, epochs, lr=lr, batch=capacity, log=log, augment=augment)
    rep['augment'] = bool(augment)
    rep['training'] = dict(seconds=round(time.perf_counter() - t0, 2), loss=[round(h, 5) for h in hist])
    data.close()
    if out:
        save(out, net, scaling)
    rep['evaluation'] = evaluate(net, scaling, eval_crops, tau, force, eval_tol, interval, max_steps, relax_sweeps,
                                 log)
    return rep


def evaluate(net, scaling, eval_crops, tau=0.8, force=1e-6, eval_tol=1e-3, interval=50, max_steps=60000,
             relax_sweeps=(5,), log=print):
    """Cold vs warm starts of the same crops to the same stop rule, plus a tight
    (1e-5) cold run as the permeability reference."""
    cold, t_cold, _ = run_crops(eval_crops, tau, force, eval_tol, interval, max_steps)
    ref, _, _ = run_crops(eval_crops, tau, force, 1e-5, interval, max_steps)
    kc = np.array([r['K'] for r in cold])
    kr = np.array([r['K'] for r in ref])
    sc = np.array([r['steps'] for r in cold])
    ev = dict(cold=dict(seconds=round(t_cold, 3), steps_mean=float(sc.mean()), steps_total=int(sc.sum()),
                        converged=int(sum(r['converged'] for r in cold))),
              k_rel_cold_vs_tight_mean=float(np.mean(np.abs(kc - kr) / kr)))
    for relax in relax_sweeps:
        warm, t_warm, t_inf = run_crops(eval_crops, tau, force, eval_tol, interval, max_steps, net, scaling, relax)
        kw = np.array([r['K'] for r in warm])
        sw = np.array([r['steps'] for r in warm])
        ev[f'warm_relax{relax}'] = dict(
            seconds=round(t_warm, 3), inference_and_relax_seconds=round(t_inf, 3), steps_mean=float(sw.mean()),
            steps_total=int(sw.sum()), converged=int(sum(r['converged'] for r in warm)),
            speedup_wall=round(t_cold / t_warm, 3), speedup_iterations=round(float(sc.sum() / sw.sum()), 3),
            k_rel_vs_cold_max=float(np.max(np.abs(kw - kc) / kc)),
            k_rel_vs_cold_mean=float(np.mean(np.abs(kw - kc) / kc)),
            k_rel_vs_tight_mean=float(np.mean(np.abs(kw - kr) / kr)))
        log(f'relax {relax}: ' + json.dumps(ev[f'warm_relax{relax}']))
    log('cold: ' + json.dumps(ev['cold']))
    return ev


def diagnose(net, scaling, crops, tau=0.8, force=1e-6, log=print):
    """Prediction error on crops solved to 1e-5: relative L2 error of u over
    fluid sites, the flux ratio sum(u_x pred) / sum(u_x true), and the error
    left after the best single scale factor (how much of the error is the
    global magnitude, the slowest mode of the warm-started run)."""
    data, _ = solve(crops, tau=tau, force=force, tol=1e-5, interval=100)
    masks = np.stack([(c == PORE) for c in crops]).astype(np.uint8).reshape(len(crops), -1)
    rows = []
    for i in range(0, len(crops), net.capacity):
        m = masks[i:i + net.capacity]
        f = scaling.fields(net.forward(m), tau, force, m)
        for j in range(m.shape[0]):
            up = f[j, :, :3].astype(np.float64)
            ut = data.sample(i + j)[1][:, :3].astype(np.float64)
            err = float(np.linalg.norm(up - ut) / np.linalg.norm(ut))
            alpha = float((up * ut).sum() / (up * up).sum())
            err_s = float(np.linalg.norm(alpha * up - ut) / np.linalg.norm(ut))
            flux = float(up[:, 2].sum() / ut[:, 2].sum())
            rows.append((err, flux, alpha, err_s))
    data.close()
    r = np.array(rows)
    rep = dict(n=len(rows), rel_l2_mean=float(r[:, 0].mean()), flux_ratio_mean=float(r[:, 1].mean()),
               flux_ratio_std=float(r[:, 1].std()), best_scale_mean=float(r[:, 2].mean()),
               rel_l2_after_best_scale_mean=float(r[:, 3].mean()))
    log('diagnose: ' + json.dumps(rep))
    return rep


def train_bench(side, capacity, steps, log=print):
    """Training-step throughput on synthetic masks and targets (no solver):
    `steps` Adam steps of `capacity` crops, timed with HIP events."""
    crops = make_crops(capacity, 30000, side, workers=1)
    masks = Buffer((capacity, side ** 3), np.uint8,
                   np.stack([(c == PORE) for c in crops]).astype(np.uint8).reshape(capacity, -1))
    rng = np.random.default_rng(0)
    targets = Buffer((capacity, side ** 3, 4), np.float32,
                     rng.standard_normal((capacity, side ** 3, 4)).astype(np.float32))
    net = Net(side=side, capacity=capacity, seed=1)
    net.timed_train(masks, targets, adam(1e-4), 2)        # warm-up
    ms, flops = net.timed_train(masks, targets, adam(1e-4), steps)
    rep = dict(side=side, samples_per_step=capacity, steps=steps, ms_per_step=round(ms, 3),
               conv_flops_per_step=flops, tflops=round(flops / ms / 1e9, 1),
               mfma_frac=round(flops / ms / 1e9 / 2500.0, 4))
    log(json.dumps(rep))
    return rep


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description='train and evaluate the warm-start network')
    ap.add_argument('--train', type=int, default=4096)
    ap.add_argument('--eval', type=int, default=256)
    ap.add_argument('--side', type=int, default=64)
    ap.add_argument('--epochs', type=int, default=8)
    ap.add_argument('--capacity', type=int, default=8)
    ap.add_argument('--lr', type=float, default=1e-3)
    ap.add_argument('--out', default=None, help='weights file (.npz)')
    ap.add_argument('--init', default=None, help='start training from these weights (.npz)')
    ap.add_argument('--report', default=None, help='report file (.json)')
    ap.add_argument('--evaluate', default=None, help='only evaluate these weights (.npz) on --eval crops')
    ap.add_argument('--diagnose', default=None, help='only report the prediction error of these weights (.npz)')
    ap.add_argument('--train-bench', type=int, default=0,
                    help='only time this many training steps on --capacity synthetic crops')
    a = ap.parse_args(argv)

    def log(*x):
        print(*x, flush=True)

    if a.train_bench:
        rep = train_bench(a.side, a.capacity, a.train_bench, log)
    elif a.diagnose:
        crops = make_crops(a.eval, EVAL_SEED0, a.side)
        net, scaling = load(a.diagnose, capacity=a.capacity)
        rep = dict(weights=a.diagnose, diagnose=diagnose(net, scaling, crops, log=log))
    elif a.evaluate:
        crops = make_crops(a.eval, EVAL_SEED0, a.side)
        net, scaling = load(a.evaluate, capacity=a.capacity)
        rep = dict(weights=a.evaluate, n_eval=a.eval, evaluation=evaluate(net, scaling, crops, log=log))
    else:
        rep = pipeline(a.train, a.eval, a.side, a.epochs, a.capacity, a.lr, out=a.out, log=log, init=a.init)
    if a.report:
        with open(a.report, 'w') as f:
            json.dump(rep, f, indent=1)


if __name__ == '__main__':
    main()
