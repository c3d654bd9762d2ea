"""ctypes binding of the engine's C ABI (include/lbflow.h).

This is the Python face of the same entry points the drop-in ``LBFlow``
executable calls.  It loads ``lbflow_amd/lib/liblbflow.so`` (built in-tree by
``__graft_entry__.build()`` or ``make -C lbflow_amd/csrc``) and fails loudly if
the library is missing: there is no CPU fallback.
"""
import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / 'lib' / 'liblbflow.so'

TYPES = {'VD': 0x1, 'GV': 0x2, 'PB': 0x4, 'VB': 0x8, 'AB': 0x10, 'BB': 0x20, 'HF': 0x40, 'SD': 0x80}
FP64, FP32 = 0, 1
STREAM_SOLIDS = 0x1
LAYOUT_DENSE = 0x2
STEP_FIELDS, STEP_ADJACENT = 0x1, 0x2
STATUS_BITS = (0x1, 0x2, 0x4, 0x8, 0x10, 0x20)

SYMBOLS = ('lbf_create', 'lbf_destroy', 'lbf_last_error', 'lbf_init', 'lbf_step', 'lbf_advance',
           'lbf_time', 'lbf_synchronize', 'lbf_status', 'lbf_status_message', 'lbf_fields',
           'lbf_get_populations', 'lbf_set_populations', 'lbf_report', 'lbf_report_static', 'lbf_links', 'lbf_sites',
           'lbf_device_bytes', 'lbf_step_kernel', 'lbf_step_bytes', 'lbf_timed_advance',
           'lbf_copy_bandwidth', 'lbf_batch_create', 'lbf_batch_destroy', 'lbf_batch_error',
           'lbf_batch_sites', 'lbf_batch_tune', 'lbf_batch_init', 'lbf_batch_advance', 'lbf_batch_run',
           'lbf_batch_crop_populations', 'lbf_batch_time', 'lbf_batch_shape', 'lbf_buffer_alloc', 'lbf_buffer_free',
           'lbf_buffer_upload', 'lbf_buffer_download')


class Label(C.Structure):
    _fields_ = [('type', C.c_int32), ('dir', C.c_int32), ('phase', C.c_int32),
                ('reserved', C.c_int32), ('value', C.c_double)]


class Desc(C.Structure):
    _fields_ = [('D', C.c_int32), ('Q', C.c_int32), ('model', C.c_int32), ('equil', C.c_int32),
                ('precision', C.c_int32), ('flags', C.c_int32), ('harmonic', C.c_int32),
                ('wall_gradient', C.c_int32), ('sz', C.c_uint64 * 3), ('tau', C.c_double * 2),
                ('density', C.c_double * 2), ('alpha0', C.c_double), ('sigma', C.c_double),
                ('beta', C.c_double), ('capillary', C.c_int32), ('curvature', C.c_int32)]


class ReportReq(C.Structure):
    _fields_ = [('axis', C.c_int32), ('convergence', C.c_int32), ('adjacent', C.c_int32),
                ('reserved', C.c_int32), ('first', C.c_uint64), ('last', C.c_uint64)]


class ReportOut(C.Structure):
    _fields_ = [('vel_range', C.c_double), ('rho_range', C.c_double * 2),
                ('rho_slice', (C.c_double * 2) * 2), ('count_slice', C.c_double * 2),
                ('conv', C.c_double * 4), ('phase_change', C.c_double * 2),
                ('has_previous', C.c_int32), ('nonfinite', C.c_int32)]


class BatchDesc(C.Structure):
    _fields_ = [('D', C.c_int32), ('Q', C.c_int32), ('equil', C.c_int32), ('ncrops', C.c_uint32),
                ('tau', C.c_double), ('density', C.c_double)]


class Crop(C.Structure):
    _fields_ = [('sz', C.c_uint64 * 3), ('geometry', C.c_void_p)]


class CropResult(C.Structure):
    _fields_ = [('steps', C.c_uint64), ('converged', C.c_int32), ('reserved', C.c_int32),
                ('residual', C.c_double), ('adjacent_residual', C.c_double), ('k_darcy', C.c_double)]


_LIB = None


def load_library(path=None):
    """Load liblbflow.so; raises if it has not been built."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f'{p} is missing: build the HIP engine first '
                           '(python -c "import __graft_entry__ as g; g.build()")')
    L = C.CDLL(str(p))
    vp, u64, i32 = C.c_void_p, C.c_uint64, C.c_int
    L.lbf_create.argtypes = [C.POINTER(Desc), C.POINTER(Label), C.c_char_p, C.POINTER(vp)]
    L.lbf_create.restype = i32
    L.lbf_destroy.argtypes = [vp]
    L.lbf_last_error.argtypes = [vp]
    L.lbf_last_error.restype = C.c_char_p
    for name in ('lbf_init', 'lbf_synchronize'):
        getattr(L, name).argtypes = [vp]
        getattr(L, name).restype = i32
    L.lbf_step.argtypes = [vp, C.c_uint32]
    L.lbf_step.restype = i32
    L.lbf_advance.argtypes = [vp, u64]
    L.lbf_advance.restype = i32
    L.lbf_time.argtypes = [vp]
    L.lbf_time.restype = u64
    L.lbf_status.argtypes = [vp, C.POINTER(C.c_uint32), vp]
    L.lbf_status.restype = i32
    L.lbf_status_message.argtypes = [C.c_uint32]
    L.lbf_status_message.restype = C.c_char_p
    L.lbf_fields.argtypes = [vp, vp, vp, vp]
    L.lbf_fields.restype = i32
    L.lbf_get_populations.argtypes = [vp, vp]
    L.lbf_get_populations.restype = i32
    L.lbf_set_populations.argtypes = [vp, vp, u64]
    L.lbf_set_populations.restype = i32
    for name in ('lbf_report', 'lbf_report_static'):
        getattr(L, name).argtypes = [vp, C.POINTER(ReportReq), C.POINTER(ReportOut)]
        getattr(L, name).restype = i32
    L.lbf_links.argtypes = [vp, vp, vp, vp, vp, vp]
    L.lbf_links.restype = u64
    L.lbf_sites.argtypes = [vp]
    L.lbf_sites.restype = u64
    L.lbf_device_bytes.argtypes = [vp]
    L.lbf_device_bytes.restype = u64
    L.lbf_step_kernel.argtypes = [vp]
    L.lbf_step_kernel.restype = C.c_char_p
    L.lbf_step_bytes.argtypes = [vp]
    L.lbf_step_bytes.restype = C.c_double
    L.lbf_timed_advance.argtypes = [vp, u64, C.POINTER(C.c_double)]
    L.lbf_timed_advance.restype = i32
    L.lbf_copy_bandwidth.argtypes = [u64, C.c_int]
    L.lbf_copy_bandwidth.restype = C.c_double
    L.lbf_batch_create.argtypes = [C.POINTER(BatchDesc), C.POINTER(Label), C.POINTER(Crop), C.POINTER(vp)]
    L.lbf_batch_create.restype = i32
    L.lbf_batch_destroy.argtypes = [vp]
    L.lbf_batch_error.argtypes = [vp]
    L.lbf_batch_error.restype = C.c_char_p
    L.lbf_batch_sites.argtypes = [vp, C.POINTER(u64)]
    L.lbf_batch_sites.restype = u64
    L.lbf_batch_tune.argtypes = [vp, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_double)]
    L.lbf_batch_tune.restype = i32
    L.lbf_batch_shape.argtypes = [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    L.lbf_batch_shape.restype = i32
    L.lbf_batch_init.argtypes = [vp]
    L.lbf_batch_init.restype = i32
    L.lbf_batch_advance.argtypes = [vp, u64]
    L.lbf_batch_advance.restype = i32
    L.lbf_batch_run.argtypes = [vp, i32, C.c_double, C.c_uint32, u64, C.POINTER(CropResult),
                                C.POINTER(C.c_double)]
    L.lbf_batch_run.restype = i32
    L.lbf_batch_crop_populations.argtypes = [vp, C.c_uint32, vp]
    L.lbf_batch_crop_populations.restype = i32
    L.lbf_batch_time.argtypes = [vp]
    L.lbf_batch_time.restype = u64
    L.lbf_buffer_alloc.argtypes = [u64, C.POINTER(vp)]
    L.lbf_buffer_alloc.restype = i32
    L.lbf_buffer_free.argtypes = [vp]
    L.lbf_buffer_free.restype = None
    L.lbf_buffer_upload.argtypes = [vp, vp, u64]
    L.lbf_buffer_upload.restype = i32
    L.lbf_buffer_download.argtypes = [vp, vp, u64]
    L.lbf_buffer_download.restype = i32
    if path is None:
        _LIB = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class EngineError(RuntimeError):
    pass


class Buffer:
    """Device memory of the engine's HIP runtime (lbf_buffer_*), for entry
    points that take device pointers; shape and dtype describe its contents."""

    def __init__(self, shape, dtype, data=None):
        self.L = load_library()
        self.shape, self.dtype = tuple(shape), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        h = C.c_void_p()
        if self.L.lbf_buffer_alloc(self.nbytes, C.byref(h)):
            raise EngineError(f'cannot allocate {self.nbytes} bytes of device memory')
        self.h = h
        if data is not None:
            self.upload(data)

    @property
    def ptr(self):
        return self.h

    def upload(self, data):
        a = np.ascontiguousarray(data, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError(f'{a.nbytes} bytes for a buffer of {self.nbytes}')
        if self.L.lbf_buffer_upload(self.h, a.ctypes.data_as(C.c_void_p), a.nbytes):
            raise EngineError('device upload failed')
        return self

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        if self.L.lbf_buffer_download(out.ctypes.data_as(C.c_void_p), self.h, self.nbytes):
            raise EngineError('device download failed')
        return out

    def close(self):
        if getattr(self, 'h', None):
            self.L.lbf_buffer_free(self.h)
            self.h = None

    def __del__(self):
        self.close()


def copy_bandwidth(nbytes=1 << 30, iters=20):
    """Achievable HBM bandwidth (GB/s) of a plain copy kernel."""
    return float(load_library().lbf_copy_bandwidth(nbytes, iters))


class Engine:
    """One lattice on the GPU.

    ``case`` is the plain lattice description of ``cases.lattice``:
    D, Q, size (slowest first), model, eq, tau, density, alpha, sigma, beta,
    harmonic, labels {value: (TYPE, value, dir, phase)}.
    """

    def __init__(self, case, geometry, precision=FP64, stream_solids=True, dense=False):
        self.L = load_library()
        d = Desc()
        d.D, d.Q = case['D'], case['Q']
        d.model = 1 if case['model'] == 'color' else 0
        d.equil = 1 if case['eq'] == 'MRT' else 0
        d.precision = precision
        d.flags = (STREAM_SOLIDS if stream_solids else 0) | (LAYOUT_DENSE if dense else 0)
        d.harmonic = 1 if case.get('harmonic') else 0
        d.wall_gradient = {'legacy': 0, 'renormalized': 1, 'tanh': 2}[case.get('wall_gradient', 'legacy')]
        d.capillary = 1 if case.get('capillary') == 'csf' else 0
        d.curvature = 1 if case.get('curvature') == 'gradient' else 0
        for k, n in enumerate(case['size']):
            d.sz[k] = n
        tau = list(case['tau']) + [0.0] * (2 - len(case['tau']))
        dens = list(case['density']) + [0.0] * (2 - len(case['density']))
        d.tau[0], d.tau[1] = tau
        d.density[0], d.density[1] = dens
        d.alpha0 = case.get('alpha', 1 / 3)
        d.sigma = case.get('sigma', 0.0)
        d.beta = case.get('beta', 0.0)
        labels = (Label * 256)()
        for key, (typ, value, direction, phase) in case['labels'].items():
            labels[key].type = TYPES[typ]
            labels[key].value = value
            labels[key].dir = direction
            labels[key].phase = phase
        geom = np.ascontiguousarray(geometry, dtype=np.uint8)
        h = C.c_void_p()
        rc = self.L.lbf_create(C.byref(d), labels, geom.tobytes(), C.byref(h))
        if rc:
            raise EngineError(self.L.lbf_last_error(None).decode())
        self.h = h
        self.D, self.Q = case['D'], case['Q']
        self.NP = 2 if case['model'] == 'color' else 1
        self.N = int(self.L.lbf_sites(h))
        self.size = tuple(case['size'])

    def close(self):
        if getattr(self, 'h', None):
            self.L.lbf_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc:
            raise EngineError(self.L.lbf_last_error(self.h).decode())

    def init(self):
        self._check(self.L.lbf_init(self.h))
        return self

    def step(self, flags=0):
        self._check(self.L.lbf_step(self.h, flags))
        return self

    def advance(self, n):
        self._check(self.L.lbf_advance(self.h, n))
        return self

    def timed_advance(self, n):
        """n steps timed with HIP events on the engine's stream; returns milliseconds."""
        ms = C.c_double()
        self._check(self.L.lbf_timed_advance(self.h, n, C.byref(ms)))
        return ms.value

    def synchronize(self):
        self._check(self.L.lbf_synchronize(self.h))

    @property
    def t(self):
        return int(self.L.lbf_time(self.h))

    def status(self):
        bits = C.c_uint32()
        first = np.zeros(6, np.uint64)
        self._check(self.L.lbf_status(self.h, C.byref(bits), _ptr(first)))
        return int(bits.value), first

    def check_status(self):
        bits, first = self.status()
        if bits:
            msgs = [self.L.lbf_status_message(b).decode() for b in STATUS_BITS if bits & b]
            raise EngineError('; '.join(msgs))

    def fields(self):
        r0 = np.zeros(self.N)
        r1 = np.zeros(self.N) if self.NP == 2 else None
        u = np.zeros(self.D * self.N)
        self._check(self.L.lbf_fields(self.h, _ptr(r0), _ptr(r1), _ptr(u)))
        return r0, r1, u.reshape(self.D, self.N)

    def populations(self):
        out = np.zeros(self.NP * self.Q * self.N)
        self._check(self.L.lbf_get_populations(self.h, _ptr(out)))
        return out

    def set_populations(self, pops, t):
        pops = np.ascontiguousarray(pops, dtype=np.float64)
        self._check(self.L.lbf_set_populations(self.h, _ptr(pops), t))

    def report(self, axis, first, last, convergence=True, adjacent=True, static=False):
        """One report's device reductions; static=True uses lbf_report_static
        (no per-context buffers)."""
        req = ReportReq(axis=axis, convergence=int(convergence), adjacent=int(adjacent),
                        first=first, last=last)
        out = ReportOut()
        fn = self.L.lbf_report_static if static else self.L.lbf_report
        self._check(fn(self.h, C.byref(req), C.byref(out)))
        return out

    def links(self):
        mask = np.zeros(self.N, np.uint32)
        wall = np.zeros(self.N, np.int32)
        n = int(self.L.lbf_links(self.h, _ptr(mask), _ptr(wall), None, None, None))
        normal = np.zeros(n * self.D)
        sin = np.zeros(n)
        cos = np.zeros(n)
        self.L.lbf_links(self.h, None, None, _ptr(normal), _ptr(sin), _ptr(cos))
        return mask, wall, normal.reshape(n, self.D), sin, cos

    @property
    def step_kernel(self):
        return self.L.lbf_step_kernel(self.h).decode()

    @property
    def step_bytes(self):
        return float(self.L.lbf_step_bytes(self.h))

    @property
    def device_bytes(self):
        return int(self.L.lbf_device_bytes(self.h))


class Batch:
    """Many independent single-phase crops stepped together (lbf_batch_*).

    ``labels`` is the {value: (TYPE, value, dir, phase)} map of a lattice case;
    ``crops`` a list of uint8 label arrays (z, y, x) or (y, x) of any sizes.
    """

    def __init__(self, D, Q, eq, tau, density, labels, crops):
        self.L = load_library()
        d = BatchDesc(D=D, Q=Q, equil=1 if eq == 'MRT' else 0, ncrops=len(crops), tau=tau, density=density)
        lab = (Label * 256)()
        for key, (typ, value, direction, phase) in labels.items():
            lab[key].type = TYPES[typ]
            lab[key].value = value
            lab[key].dir = direction
            lab[key].phase = phase
        self._geoms = [np.ascontiguousarray(g, dtype=np.uint8) for g in crops]
        arr = (Crop * len(crops))()
        for i, g in enumerate(self._geoms):
            if g.ndim != D:
                raise ValueError('crop dimensionality does not match D')
            for k, n in enumerate(g.shape):
                arr[i].sz[k] = n
            arr[i].geometry = g.ctypes.data
        h = C.c_void_p()
        rc = self.L.lbf_batch_create(C.byref(d), lab, arr, C.byref(h))
        if rc:
            raise EngineError(self.L.lbf_batch_error(None).decode())
        self.h = h
        self.D, self.Q = D, Q
        self.shapes = [g.shape for g in self._geoms]
        fluid = C.c_uint64()
        self.sites = int(self.L.lbf_batch_sites(h, C.byref(fluid)))
        self.fluid = int(fluid.value)

    def close(self):
        if getattr(self, 'h', None):
            self.L.lbf_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc:
            raise EngineError(self.L.lbf_batch_error(self.h).decode())

    def tune(self, steps=10):
        per = C.c_uint32()
        ms = C.c_double()
        self._check(self.L.lbf_batch_tune(self.h, steps, C.byref(per), C.byref(ms)))
        return int(per.value), ms.value

    def shape(self):
        """Launch shape of the step: (threads per workgroup, slots per work
        unit, workgroups of the persistent grid)."""
        t, u, w = C.c_uint32(), C.c_uint32(), C.c_uint32()
        self._check(self.L.lbf_batch_shape(self.h, C.byref(t), C.byref(u), C.byref(w)))
        return int(t.value), int(u.value), int(w.value)

    def init(self):
        self._check(self.L.lbf_batch_init(self.h))
        return self

    def advance(self, n):
        self._check(self.L.lbf_batch_advance(self.h, n))
        return self

    @property
    def t(self):
        return int(self.L.lbf_batch_time(self.h))

    def run(self, axis, tol, interval, max_steps):
        """Run every crop to convergence; returns (list of result dicts, seconds)."""
        res = (CropResult * len(self.shapes))()
        sec = C.c_double()
        self._check(self.L.lbf_batch_run(self.h, axis, tol, interval, max_steps, res, C.byref(sec)))
        out = [dict(steps=int(r.steps), converged=bool(r.converged), residual=r.residual,
                    adjacent=r.adjacent_residual, K=r.k_darcy) for r in res]
        return out, sec.value

    def populations(self, c):
        n = int(np.prod(self.shapes[c]))
        out = np.zeros(self.Q * n)
        self._check(self.L.lbf_batch_crop_populations(self.h, c, _ptr(out)))
        return out
