"""ctypes binding of the C ABI declared in include/b2r_hip.h.

Loads ``libb2r_hip.so`` built in-tree next to this file.  There is no CPU
fallback: if the library is missing, importing this module raises ImportError
with the build command, and every call that fails on the device raises
:class:`B2RError` carrying the library's return code and message.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("B2R_HIP_LIB") or os.path.join(_HERE, "libb2r_hip.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build the HIP library first "
        "(python -c 'import __graft_entry__ as g; g.build()' or make -C b2r_amd/csrc)")

_lib = C.CDLL(LIB_PATH)

B2R_OK = 0
B2R_E_ARG = -1
B2R_E_GEOM = -2
B2R_E_HIP = -3
B2R_E_NOMEM = -4
B2R_E_UNSUPPORTED = -5

BOUNDARY_FROZEN = 0
BOUNDARY_PERIODIC = 1
HEAT_F32 = 0
HEAT_REF = 1


class B2RError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"b2r error {code}: {msg}")
        self.code = code


class Geom(C.Structure):
    _fields_ = [("env_x", C.c_int), ("env_y", C.c_int), ("env_z", C.c_int),
                ("block_dim_x", C.c_int), ("block_dim_y", C.c_int), ("block_dim_z", C.c_int),
                ("R", C.c_int), ("radius", C.c_int), ("boundary", C.c_int),
                ("elem_size", C.c_int)]


class Layout(C.Structure):
    _fields_ = [("B", C.c_int * 3), ("S", C.c_int * 3), ("halo", C.c_int),
                ("nblocks", C.c_int), ("x_offset", C.c_int), ("pitch", C.c_int),
                ("slice", C.c_longlong), ("block_stride", C.c_longlong),
                ("field_elems", C.c_longlong), ("elem_size", C.c_int)]


class Tuning(C.Structure):
    _fields_ = [("heat_ry", C.c_int), ("heat_zc", C.c_int), ("gol_zc", C.c_int),
                ("gol_step_zc", C.c_int), ("el_zc", C.c_int), ("heat_wy", C.c_int),
                ("gol_tile", C.c_int), ("heat_wx", C.c_int), ("el_kernel", C.c_int),
                ("el_step_zc", C.c_int), ("el_grad_mb", C.c_int), ("gol_bytes", C.c_int),
                ("gol_exchange", C.c_int)]


class HeatParams(C.Structure):
    _fields_ = [("D", C.c_int), ("mode", C.c_int), ("k", C.c_float * 3),
                ("adt", C.c_double), ("h2", C.c_double * 3)]


_vp = C.c_void_p
_i = C.c_int
_pi = C.POINTER(C.c_int)
_pf = C.POINTER(C.c_float)

_SIGS = {
    "b2r_last_error": (C.c_char_p, []),
    "b2r_version": (_i, []),
    "b2r_create": (_i, [C.POINTER(Geom), _i, C.POINTER(_vp)]),
    "b2r_destroy": (_i, [_vp]),
    "b2r_get_layout": (_i, [_vp, C.POINTER(Layout)]),
    "b2r_synchronize": (_i, [_vp]),
    "b2r_sync_flag": (_i, [_vp, _i, C.POINTER(C.c_byte)]),
    "b2r_set_frozen_faces": (_i, [_vp, _i]),
    "b2r_set_tuning": (_i, [_vp, C.POINTER(Tuning)]),
    "b2r_get_tuning": (_i, [_vp, C.POINTER(Tuning)]),
    "b2r_field_alloc": (_i, [_vp, C.POINTER(_vp)]),
    "b2r_field_free": (_i, [_vp, _vp]),
    "b2r_field_fill": (_i, [_vp, _vp, _vp]),
    "b2r_field_upload": (_i, [_vp, _vp, _vp]),
    "b2r_field_download": (_i, [_vp, _vp, _vp]),
    "b2r_block_upload": (_i, [_vp, _vp, _i, _vp]),
    "b2r_block_download": (_i, [_vp, _vp, _i, _vp]),
    "b2r_field_copy": (_i, [_vp, _vp, _vp]),
    "b2r_field_init_uniform": (_i, [_vp, _vp, C.c_uint64]),
    "b2r_field_init_bernoulli": (_i, [_vp, _vp, C.c_uint64, C.c_uint32]),
    "b2r_field_init_blobs": (_i, [_vp, _vp, _i, _vp, _vp, _vp, _vp]),
    "b2r_boundary_updating": (_i, [_vp, _i, _vp]),
    "b2r_boundary_updating_direct": (_i, [_vp, _i, C.POINTER(_vp), _i]),
    "b2r_boundary_updating_noalloc": (_i, [_vp, _i, _vp]),
    "b2r_msg_elems": (C.c_longlong, [_vp, _i]),
    "b2r_pack": (_i, [_vp, _i, _vp, _vp]),
    "b2r_unpack": (_i, [_vp, _i, _vp, _vp]),
    "b2r_heat_step": (_i, [_vp, _vp, _vp, C.POINTER(HeatParams), _i]),
    "b2r_heat_run": (_i, [_vp, _vp, _vp, C.POINTER(HeatParams), _i, _i, _pi, _pf]),
    "b2r_heat_gradient": (_i, [_vp, _vp, _vp, C.POINTER(HeatParams), _i, _i,
                               C.POINTER(C.c_double), _vp, C.POINTER(C.c_double)]),
    "b2r_heat_cpu_step": (_i, [C.POINTER(HeatParams), _vp, _vp, _i, _i, _i, _i, _i]),
    "b2r_heat_run_hybrid": (_i, [_vp, _vp, _vp, C.POINTER(HeatParams), _i, _i, _i, _pi, _pf]),
    "b2r_gol_step": (_i, [_vp, _vp, _vp, _i]),
    "b2r_gol_run": (_i, [_vp, _vp, _vp, _i, _i, _pi, _pf]),
    "b2r_stream_copy": (_i, [_vp, _vp, _vp, C.c_size_t, _i, _pf]),
    "b2r_time_steps": (_i, [_vp, _i, _vp, _vp, C.POINTER(HeatParams), _i, _pf]),
    "b2r_el_create": (_i, [_vp, C.c_float, _pf, C.POINTER(_vp)]),
    "b2r_el_destroy": (_i, [_vp]),
    "b2r_el_field": (_i, [_vp, _i, C.POINTER(_vp)]),
    "b2r_el_set_medium": (_i, [_vp, _vp, _vp, _vp]),
    "b2r_el_reset": (_i, [_vp]),
    "b2r_el_enable_displacement": (_i, [_vp]),
    "b2r_el_set_sources": (_i, [_vp, _i, _vp, _vp, _i, _vp]),
    "b2r_el_set_receivers": (_i, [_vp, _i, _vp, _i, _i]),
    "b2r_el_get_receivers": (_i, [_vp, _vp]),
    "b2r_el_get_state": (_i, [_vp, _i, _vp]),
    "b2r_el_set_state": (_i, [_vp, _i, _vp]),
    "b2r_el_forward": (_i, [_vp, _i, _i, _pf]),
    "b2r_el_forward_batch": (_i, [C.POINTER(_vp), _i, _i, _i, _pf]),
    "b2r_el_misfit": (_i, [_vp, _vp, C.POINTER(C.c_double)]),
    "b2r_el_gradient": (_i, [_vp, _vp, _i, _i, C.POINTER(C.c_double), _vp, _vp, _vp]),
    "b2r_el_gradient_ex": (_i, [_vp, _vp, _i, _i, _i, C.POINTER(C.c_double), _vp, _vp, _vp,
                                _vp]),
    "b2r_el_autotune": (_i, [_vp, C.c_char_p, C.POINTER(Tuning)]),
    "b2r_el_autotune_batch": (_i, [C.POINTER(_vp), _i, C.c_char_p]),
    "b2r_el_gradient_batch": (_i, [C.POINTER(_vp), _i, C.POINTER(_vp), _i, _i, _i,
                                   C.POINTER(C.c_double), C.POINTER(_vp), C.POINTER(_vp),
                                   C.POINTER(_vp), _pf]),
    "b2r_el_gradient_plan": (_i, [C.POINTER(_vp), _i, _i, _i, _pi, C.POINTER(C.c_size_t)]),
    "b2r_autotune": (_i, [_vp, _i, _vp, _vp, C.POINTER(HeatParams), C.c_char_p,
                          C.POINTER(Tuning)]),
    "b2r_dev_alloc": (_i, [_vp, C.c_size_t, C.POINTER(_vp)]),
    "b2r_dev_free": (_i, [_vp, _vp]),
    "b2r_memcpy_h2d": (_i, [_vp, _vp, _vp, C.c_size_t]),
    "b2r_memcpy_d2h": (_i, [_vp, _vp, _vp, C.c_size_t]),
}

for _name, (_res, _args) in _SIGS.items():
    _f = getattr(_lib, _name)
    _f.restype = _res
    _f.argtypes = _args

EXPORTED = tuple(_SIGS)


def lib():
    return _lib


def last_error() -> str:
    return _lib.b2r_last_error().decode()


def check(rc: int) -> int:
    if rc < 0:
        raise B2RError(rc, last_error())
    return rc


_DT = {1: np.uint8, 4: np.float32, 8: np.float64}


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_vp)


class Context:
    """One B2R geometry on one GPU: Bx*By*Bz padded blocks per field."""

    def __init__(self, env, blocks, R, radius, boundary=BOUNDARY_FROZEN, elem_size=4,
                 device=0):
        g = Geom(env[0], env[1], env[2], blocks[0], blocks[1], blocks[2], R, radius,
                 boundary, elem_size)
        self.geom = g
        h = _vp()
        check(_lib.b2r_create(C.byref(g), device, C.byref(h)))
        self.h = h
        lay = Layout()
        check(_lib.b2r_get_layout(h, C.byref(lay)))
        self.layout = lay
        self.env = tuple(env)
        self.dtype = _DT[elem_size]

    # -- lifetime ------------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            _lib.b2r_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tuning(self, **kw):
        t = Tuning()
        check(_lib.b2r_get_tuning(self.h, C.byref(t)))
        for k, v in kw.items():
            setattr(t, k, v)
        check(_lib.b2r_set_tuning(self.h, C.byref(t)))

    def get_tuning(self):
        t = Tuning()
        check(_lib.b2r_get_tuning(self.h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in Tuning._fields_}

    def autotune(self, kind, fa, fb, params=None, record=None):
        t = Tuning()
        check(_lib.b2r_autotune(self.h, kind, fa, fb, C.byref(params) if params else None,
                                record.encode() if record else None, C.byref(t)))
        return {k: getattr(t, k) for k, _ in Tuning._fields_}

    def set_frozen_faces(self, mask):
        check(_lib.b2r_set_frozen_faces(self.h, mask))

    def sync_flag(self, rank):
        f = (C.c_byte * 6)()
        check(_lib.b2r_sync_flag(self.h, rank, f))
        return list(f)

    # -- fields ----------------------------------------------------------------
    def field(self):
        p = _vp()
        check(_lib.b2r_field_alloc(self.h, C.byref(p)))
        return p

    def free(self, f):
        check(_lib.b2r_field_free(self.h, f))

    def fill(self, f, value):
        v = np.array([value], dtype=self.dtype)
        check(_lib.b2r_field_fill(self.h, f, _ptr(v)))

    def upload(self, f, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        gx, gy, gz = self.env
        assert a.shape == (gz, gy, gx), (a.shape, (gz, gy, gx))
        check(_lib.b2r_field_upload(self.h, f, _ptr(a)))

    def download(self, f) -> np.ndarray:
        gx, gy, gz = self.env
        a = np.empty((gz, gy, gx), dtype=self.dtype)
        check(_lib.b2r_field_download(self.h, f, _ptr(a)))
        return a

    def raw(self, f) -> np.ndarray:
        """The whole field allocation (layout order), copied to the host."""
        n = self.layout.field_elems
        a = np.empty(n, dtype=self.dtype)
        check(_lib.b2r_memcpy_d2h(self.h, _ptr(a), f, a.nbytes))
        return a

    def raw_upload(self, f, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.layout.field_elems
        check(_lib.b2r_memcpy_h2d(self.h, f, _ptr(a), a.nbytes))

    def init_uniform(self, f, seed):
        check(_lib.b2r_field_init_uniform(self.h, f, seed))

    def init_bernoulli(self, f, seed, p_q24):
        check(_lib.b2r_field_init_bernoulli(self.h, f, seed, p_q24))

    def init_blobs(self, f, amp, ex, ey, ez):
        amp = np.ascontiguousarray(amp, np.float64)
        ex, ey, ez = (np.ascontiguousarray(t, np.float64) for t in (ex, ey, ez))
        check(_lib.b2r_field_init_blobs(self.h, f, len(amp), _ptr(amp), _ptr(ex), _ptr(ey),
                                        _ptr(ez)))

    def sync(self):
        check(_lib.b2r_synchronize(self.h))

    # -- exchange ---------------------------------------------------------------
    def boundary_updating(self, f, ts=0):
        check(_lib.b2r_boundary_updating(self.h, ts, f))

    def boundary_updating_noalloc(self, f, ts=0):
        check(_lib.b2r_boundary_updating_noalloc(self.h, ts, f))

    def boundary_updating_direct(self, fields, ts=0):
        arr = (_vp * len(fields))(*fields)
        check(_lib.b2r_boundary_updating_direct(self.h, ts, arr, len(fields)))

    def msg_elems(self, dim):
        return check(_lib.b2r_msg_elems(self.h, dim))

    def dev_alloc(self, nbytes):
        p = _vp()
        check(_lib.b2r_dev_alloc(self.h, nbytes, C.byref(p)))
        return p

    def dev_free(self, p):
        check(_lib.b2r_dev_free(self.h, p))

    def pack(self, dim, f, msgs):
        check(_lib.b2r_pack(self.h, dim, f, msgs))

    def unpack(self, dim, f, msgs):
        check(_lib.b2r_unpack(self.h, dim, f, msgs))

    def d2h(self, p, nbytes, dtype):
        a = np.empty(nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        check(_lib.b2r_memcpy_d2h(self.h, _ptr(a), p, nbytes))
        return a

    def h2d(self, p, a):
        a = np.ascontiguousarray(a)
        check(_lib.b2r_memcpy_h2d(self.h, p, _ptr(a), a.nbytes))

    # -- models ---------------------------------------------------------------
    def heat_step(self, fin, fout, params: HeatParams, b2r_i):
        check(_lib.b2r_heat_step(self.h, fin, fout, C.byref(params), b2r_i))

    def heat_run(self, fa, fb, params: HeatParams, nsteps, fused=0, timed=False):
        rib = C.c_int(0)
        ms = C.c_float(0.0)
        check(_lib.b2r_heat_run(self.h, fa, fb, C.byref(params), nsteps, fused, C.byref(rib),
                                C.byref(ms) if timed else None))
        return (fb if rib.value else fa), (ms.value if timed else None)

    def heat_run_hybrid(self, fa, fb, params: HeatParams, nsteps, cpu_planes, nthreads=0,
                        timed=False):
        """b2r_heat_run with the top cpu_planes z planes stepped on the host
        cores; returns (result field, wall ms or None)."""
        rib = C.c_int(0)
        ms = C.c_float(0.0)
        check(_lib.b2r_heat_run_hybrid(self.h, fa, fb, C.byref(params), nsteps, cpu_planes,
                                       nthreads, C.byref(rib), C.byref(ms) if timed else None))
        return (fb if rib.value else fa), (ms.value if timed else None)

    def heat_gradient(self, u0, target, params: HeatParams, nsteps, ckpt, grad_u0):
        """J = 1/2 |u_N - target|^2 over owned cells; writes dJ/du0 into the
        field grad_u0 and returns (J, dJ/dk as 3 floats)."""
        J = C.c_double(0.0)
        gk = (C.c_double * 3)()
        check(_lib.b2r_heat_gradient(self.h, u0, target, C.byref(params), nsteps, ckpt,
                                     C.byref(J), grad_u0, gk))
        return J.value, np.array(list(gk), np.float64)

    def gol_step(self, fin, fout, b2r_i):
        check(_lib.b2r_gol_step(self.h, fin, fout, b2r_i))

    def gol_run(self, fa, fb, nsteps, fused=0, timed=False):
        rib = C.c_int(0)
        ms = C.c_float(0.0)
        check(_lib.b2r_gol_run(self.h, fa, fb, nsteps, fused, C.byref(rib),
                               C.byref(ms) if timed else None))
        return (fb if rib.value else fa), (ms.value if timed else None)

    def time_steps(self, kind, fa, fb, iters, params=None):
        ms = C.c_float(0.0)
        check(_lib.b2r_time_steps(self.h, kind, fa, fb, C.byref(params) if params else None,
                                  iters, C.byref(ms)))
        return ms.value

    def stream_copy(self, dst, src, nbytes, iters):
        ms = C.c_float(0.0)
        check(_lib.b2r_stream_copy(self.h, dst, src, nbytes, iters, C.byref(ms)))
        return ms.value


def heat_params(D, mode=HEAT_F32, k=(0.1, 0.1, 0.1), adt=0.1 * 0.1, h2=(1.0, 1.0, 0.25)):
    p = HeatParams()
    p.D = D
    p.mode = mode
    for a in range(3):
        p.k[a] = k[a]
        p.h2[a] = h2[a]
    p.adt = adt
    return p


def heat_cpu_step(params: HeatParams, grid: np.ndarray, b2r_i=0, nthreads=0) -> np.ndarray:
    """One inner step of delta-D on the host cores (b2r_heat_cpu_step) over a
    dense padded (z, y, x) grid; returns the new grid (cells outside the
    step's region copied from the input)."""
    a = np.ascontiguousarray(grid, np.float32)
    out = a.copy()
    nz, ny, nx = a.shape
    check(_lib.b2r_heat_cpu_step(C.byref(params), _ptr(a), _ptr(out), nx, ny, nz, b2r_i,
                                 nthreads))
    return out


# ------------------------------------------------------------------ elastic
EL_FIELDS = ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz", "lam", "mu", "b")


class ElasticShot:
    """One elastic shot on a Context of radius 4 (see b2r_el_* in b2r_hip.h)."""

    def __init__(self, ctx: Context, dt, inv_h):
        self.ctx = ctx
        ih = (C.c_float * 3)(*inv_h)
        h = _vp()
        check(_lib.b2r_el_create(ctx.h, dt, ih, C.byref(h)))
        self.h = h
        self.nrec = 0
        self.rec_nt = 0

    def close(self):
        if getattr(self, "h", None):
            _lib.b2r_el_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_medium(self, lam, mu, b):
        arrs = [np.ascontiguousarray(a, np.float32) for a in (lam, mu, b)]
        check(_lib.b2r_el_set_medium(self.h, *[_ptr(a) for a in arrs]))

    def reset(self):
        check(_lib.b2r_el_reset(self.h))

    def enable_displacement(self):
        check(_lib.b2r_el_enable_displacement(self.h))

    def set_sources(self, xyz, comp, wav):
        xyz = np.ascontiguousarray(xyz, np.int32).reshape(-1, 3)
        comp = np.ascontiguousarray(comp, np.int32).ravel()
        wav = np.ascontiguousarray(wav, np.float32)
        nt = wav.shape[0] if wav.ndim == 2 else (wav.size // max(len(comp), 1))
        check(_lib.b2r_el_set_sources(self.h, len(comp), _ptr(xyz), _ptr(comp), nt, _ptr(wav)))

    def set_receivers(self, xyz, comp, nt):
        xyz = np.ascontiguousarray(xyz, np.int32).reshape(-1, 3)
        check(_lib.b2r_el_set_receivers(self.h, len(xyz), _ptr(xyz), comp, nt))
        self.nrec, self.rec_nt = len(xyz), nt

    def receivers(self):
        out = np.empty((self.rec_nt, self.nrec), np.float32)
        check(_lib.b2r_el_get_receivers(self.h, _ptr(out)))
        return out

    def state(self, which):
        gx, gy, gz = self.ctx.env
        a = np.empty((gz, gy, gx), np.float32)
        idx = which if isinstance(which, int) else (-1 if which == "disp" else EL_FIELDS.index(which))
        check(_lib.b2r_el_get_state(self.h, idx, _ptr(a)))
        return a

    def set_state(self, which, a):
        idx = which if isinstance(which, int) else EL_FIELDS.index(which)
        a = np.ascontiguousarray(a, np.float32)
        check(_lib.b2r_el_set_state(self.h, idx, _ptr(a)))

    def forward(self, nsteps, mode=1, timed=False):
        ms = C.c_float(0.0)
        check(_lib.b2r_el_forward(self.h, nsteps, mode, C.byref(ms) if timed else None))
        return ms.value if timed else None

    def misfit(self, obs):
        obs = np.ascontiguousarray(obs, np.float32)
        m = C.c_double(0.0)
        check(_lib.b2r_el_misfit(self.h, _ptr(obs), C.byref(m)))
        return m.value

    def autotune(self, record=None):
        """Pick the elastic z chunk for this shot's size (resets the shot)."""
        t = Tuning()
        check(_lib.b2r_el_autotune(self.h, record.encode() if record else None, C.byref(t)))
        return {k: getattr(t, k) for k, _ in Tuning._fields_}

    def gradient(self, obs, nsteps, ckpt, from_current=False, adjoint_state=False, plain=False):
        """misfit and [g_lam, g_mu, g_b] (plus the 9 adjoint fields at t=0).
        plain: the per-step store-and-recompute algorithm instead of the
        reverse sweep."""
        gx, gy, gz = self.ctx.env
        obs = np.ascontiguousarray(obs, np.float32)
        g = [np.empty((gz, gy, gx), np.float32) for _ in range(3)]
        adj = np.empty((9, gz, gy, gx), np.float32) if adjoint_state else None
        m = C.c_double(0.0)
        flags = int(bool(from_current)) | (2 if plain else 0)
        check(_lib.b2r_el_gradient_ex(self.h, _ptr(obs), nsteps, ckpt, flags,
                                      C.byref(m), *[_ptr(a) for a in g],
                                      _ptr(adj) if adjoint_state else None))
        return (m.value, g, adj) if adjoint_state else (m.value, g)


def el_autotune_batch(shots, record=None):
    """Pick the elastic kernel and z chunk for a batch as a whole (resets the
    shots); returns the tuning now set in every shot's context."""
    arr = (_vp * len(shots))(*[s.h for s in shots])
    check(_lib.b2r_el_autotune_batch(arr, len(shots), record.encode() if record else None))
    return shots[0].ctx.get_tuning()


def el_forward_batch(shots, nsteps, mode=1, timed=False):
    arr = (_vp * len(shots))(*[s.h for s in shots])
    ms = C.c_float(0.0)
    check(_lib.b2r_el_forward_batch(arr, len(shots), nsteps, mode, C.byref(ms) if timed else None))
    return ms.value if timed else None


def el_gradient_batch(shots, obs, nsteps, ckpt, mode=1, timed=False):
    """Gradients of several shots: mode 0 plain per-shot loop, 1 batched
    reverse sweep, 2 batched with HIP graphs.  Returns (misfits, grads, ms)
    with grads[k] = [g_lam, g_mu, g_b] of shot k."""
    n = len(shots)
    arr = (_vp * n)(*[s.h for s in shots])
    obs = [np.ascontiguousarray(o, np.float32) for o in obs]
    grads = []
    for s in shots:
        gx, gy, gz = s.ctx.env
        grads.append([np.empty((gz, gy, gx), np.float32) for _ in range(3)])
    po = (_vp * n)(*[_ptr(o) for o in obs])
    pl, pm, pb = ((_vp * n)(*[_ptr(g[i]) for g in grads]) for i in range(3))
    mis = (C.c_double * n)()
    ms = C.c_float(0.0)
    check(_lib.b2r_el_gradient_batch(arr, n, po, nsteps, ckpt, mode, mis, pl, pm, pb,
                                     C.byref(ms) if timed else None))
    return list(mis), grads, (ms.value if timed else None)


def el_gradient_plan(shots, nsteps, ckpt):
    """(groups, device bytes of the largest group) of a batched gradient."""
    arr = (_vp * len(shots))(*[s.h for s in shots])
    n = C.c_int(0)
    big = C.c_size_t(0)
    check(_lib.b2r_el_gradient_plan(arr, len(shots), nsteps, ckpt, C.byref(n), C.byref(big)))
    return n.value, big.value
