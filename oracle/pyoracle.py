"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product package never does.
"""
import ctypes as C
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB = None

TYPES = {'VD': 0x1, 'GV': 0x2, 'PB': 0x4, 'VB': 0x8, 'AB': 0x10, 'BB': 0x20, 'HF': 0x40, 'SD': 0x80}


class Label(C.Structure):
    _fields_ = [('type', C.c_int32), ('dir', C.c_int32), ('phase', C.c_int32),
                ('pad', C.c_int32), ('value', C.c_double)]


class Desc(C.Structure):
    _fields_ = [('D', C.c_int32), ('Q', C.c_int32), ('model', C.c_int32), ('equil', C.c_int32),
                ('harmonic', C.c_int32), ('wall_gradient', C.c_int32), ('sz', C.c_uint64 * 3),
                ('tau', C.c_double * 2), ('density', C.c_double * 2), ('alpha0', C.c_double),
                ('sigma', C.c_double), ('beta', C.c_double), ('capillary', C.c_int32),
                ('curvature', C.c_int32)]


def lib():
    global _LIB
    if _LIB is None:
        path = _HERE / 'liblbf_oracle.so'
        if not path.exists():
            raise RuntimeError(f'oracle library missing: build it with make -C {_HERE}')
        L = C.CDLL(str(path))
        vp, u64, dp = C.c_void_p, C.c_uint64, C.POINTER(C.c_double)
        L.orc_new.restype = vp
        L.orc_new.argtypes = [C.POINTER(Desc), C.POINTER(Label), C.c_char_p, C.c_char_p, C.c_size_t]
        L.orc_free.argtypes = [vp]
        for name in ('orc_init', 'orc_step'):
            getattr(L, name).argtypes = [vp]
            getattr(L, name).restype = C.c_int
        L.orc_run.argtypes = [vp, u64]
        L.orc_stream_only.argtypes = [vp]
        L.orc_stream_only.restype = None
        L.orc_macro_only.argtypes = [vp]
        L.orc_run.restype = C.c_int
        L.orc_error.argtypes = [vp]
        L.orc_error.restype = C.c_char_p
        L.orc_time.argtypes = [vp]
        L.orc_time.restype = u64
        L.orc_sites.argtypes = [vp]
        L.orc_sites.restype = u64
        L.orc_fields.argtypes = [vp, vp, vp, vp]
        L.orc_get_populations.argtypes = [vp, vp]
        L.orc_adjacent.argtypes = [vp, vp, vp]
        L.orc_set_populations.argtypes = [vp, vp, u64]
        L.orc_links.argtypes = [vp, vp, vp, vp, vp, vp]
        L.orc_links.restype = u64
        L.orc_mrt_matrices.argtypes = [C.c_int, C.c_double, vp, vp, vp, vp]
        L.orc_color_gradient.argtypes = [vp, u64, vp, vp]
        L.orc_contact_angle.argtypes = [C.c_int, vp, vp, C.c_double, C.c_double, vp, C.c_int]
        L.orc_contact_angle.restype = C.c_int
        L.orc_report_sums.argtypes = [vp, C.c_int, u64, u64, vp]
        L.orc_convergence_sums.argtypes = [vp, vp, vp, vp]
        _LIB = L
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class OracleError(RuntimeError):
    pass


class Oracle:
    """One simulation in the CPU oracle.

    case: dict with D, Q, size (reference order, slowest first), model ('scsp'|'color'),
    eq ('SRT'|'MRT'), tau, density, alpha, sigma, beta, harmonic,
    labels {value: (type, value, dir, phase)}; geometry: uint8 array of the grid.
    """

    def __init__(self, case, geometry):
        L = lib()
        d = Desc()
        d.D, d.Q = case['D'], case['Q']
        d.model = 1 if case['model'] == 'color' else 0
        d.equil = 1 if case['eq'] == 'MRT' else 0
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
        geom = np.ascontiguousarray(geometry, dtype=np.uint8).tobytes()
        err = C.create_string_buffer(256)
        self.h = L.orc_new(C.byref(d), labels, geom, err, 256)
        if not self.h:
            raise OracleError(err.value.decode())
        self.D, self.Q = case['D'], case['Q']
        self.NP = 2 if case['model'] == 'color' else 1
        self.N = int(L.orc_sites(self.h))
        self.size = tuple(case['size'])

    def __del__(self):
        if getattr(self, 'h', None):
            lib().orc_free(self.h)
            self.h = None

    def _check(self, rc):
        if rc:
            raise OracleError(lib().orc_error(self.h).decode())

    def init(self):
        self._check(lib().orc_init(self.h))
        return self

    def step(self, n=1):
        self._check(lib().orc_run(self.h, n))
        return self

    @property
    def t(self):
        return int(lib().orc_time(self.h))

    def fields(self):
        r0 = np.zeros(self.N)
        r1 = np.zeros(self.N)
        u = np.zeros(self.D * self.N)
        lib().orc_fields(self.h, _ptr(r0), _ptr(r1), _ptr(u))
        return r0, (r1 if self.NP == 2 else None), u.reshape(self.D, self.N)

    def adjacent(self):
        """Adjacent report history of the last step: (u [D*N], chi [N] or None)."""
        u = np.zeros(self.N * self.D)
        chi = np.zeros(self.N) if self.NP == 2 else None
        lib().orc_adjacent(self.h, _ptr(u), _ptr(chi))
        return u, chi

    def populations(self):
        out = np.zeros(self.NP * self.Q * self.N)
        lib().orc_get_populations(self.h, _ptr(out))
        return out

    def set_populations(self, pops, t):
        pops = np.ascontiguousarray(pops, dtype=np.float64)
        lib().orc_set_populations(self.h, _ptr(pops), t)

    def macro_only(self):
        """Macro variables of the current populations (test hook)."""
        self._check(lib().orc_macro_only(self.h))
        return self

    def stream_only(self):
        """Streaming half of a step on the current populations (test hook)."""
        lib().orc_stream_only(self.h)
        return self

    def links(self):
        mask = np.zeros(self.N, np.uint32)
        wall = np.zeros(self.N, np.int32)
        n = int(lib().orc_links(self.h, _ptr(mask), _ptr(wall), None, None, None))
        normal = np.zeros(n * self.D)
        sin = np.zeros(n)
        cos = np.zeros(n)
        lib().orc_links(self.h, None, None, _ptr(normal), _ptr(sin), _ptr(cos))
        return mask, wall, normal.reshape(n, self.D), sin, cos

    def color_gradient(self, site):
        g = np.zeros(3)
        n = np.zeros(1)
        lib().orc_color_gradient(self.h, site, _ptr(g), _ptr(n))
        return g[:self.D], float(n[0])

    def report_sums(self, axis, first, last):
        out = np.zeros(9)
        lib().orc_report_sums(self.h, axis, first, last, _ptr(out))
        return out

    def convergence_sums(self, prev, adj=None):
        out = np.zeros(4)
        prev = np.ascontiguousarray(prev, dtype=np.float64)
        adj = None if adj is None else np.ascontiguousarray(adj, dtype=np.float64)
        lib().orc_convergence_sums(self.h, _ptr(prev), _ptr(adj), _ptr(out))
        return out


def mrt_matrices(Q, tau):
    A = np.zeros(Q * Q)
    B = np.zeros(Q * Q)
    M = np.zeros(Q * Q)
    MInv = np.zeros(Q * Q)
    lib().orc_mrt_matrices(Q, tau, _ptr(A), _ptr(B), _ptr(M), _ptr(MInv))
    return A.reshape(Q, Q), B.reshape(Q, Q), M.reshape(Q, Q), MInv.reshape(Q, Q)


def contact_angle(n, g, angle_deg, plant=0):
    """The oracle's geometric contact-angle rotation of one wall gradient
    (reference component order) for integer wall normal n; plant > 0 plants a
    known mistake (see orc_contact_angle)."""
    D = len(n)
    n = np.ascontiguousarray(n, dtype=np.float64)
    g = np.ascontiguousarray(g, dtype=np.float64)
    out = np.zeros(D)
    a = angle_deg / 180.0 * np.pi
    if lib().orc_contact_angle(D, _ptr(n), _ptr(g), float(np.sin(a)), float(np.cos(a)), _ptr(out), plant):
        raise ValueError('orc_contact_angle: bad arguments')
    return out
