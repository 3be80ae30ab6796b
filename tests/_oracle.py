"""ctypes wrappers of the CPU oracle (oracle/liboracle.so) and of the
reference's own CPU heat path compiled from its sources (oracle/_ref).

Checker only: used by tests, the smoke check and the bench's CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# B2R_ORACLE_LIB selects another build (the sanitizer build, for one)
ORACLE_SO = os.environ.get("B2R_ORACLE_LIB") or os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libheat3d_ref.so")

_vp = C.c_void_p
_i = C.c_int
_ll = C.c_longlong
_d = C.c_double


def _load():
    if not os.path.exists(ORACLE_SO):
        raise ImportError(f"{ORACLE_SO} missing: run make -C oracle")
    lib = C.CDLL(ORACLE_SO)
    sigs = {
        "orc_splitmix64": (C.c_uint64, [C.c_uint64]),
        "orc_gen_uniform": (None, [_vp, _ll, C.c_uint64]),
        "orc_gen_bernoulli": (None, [_vp, _ll, C.c_uint64, C.c_uint32]),
        "orc_gen_blobs": (None, [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp]),
        "orc_heat_coef_f32": (None, [_i, _vp]),
        "orc_heat_coef_f64": (None, [_i, _vp]),
        "orc_heat_f64": (None, [_vp, _i, _i, _i, _i, _vp, _i, _i]),
        "orc_heat_f32": (None, [_vp, _i, _i, _i, _i, _i, _vp, _d, _vp, _i, _i]),
        "orc_heat_ref1_block_step": (None, [_vp, _vp, _i, _i, _i, _i, _d, _d, _d, _d]),
        "orc_heat_ref1_block_run": (None, [_vp, _vp, _i, _i, _i, _i, _d, _d, _d, _d]),
        "orc_gol": (None, [_vp, _i, _i, _i, _i, _i]),
        "orc_popcount_u8": (_ll, [_vp, _ll]),
        "orc_msg_elems": (_ll, [_vp, _vp, _vp, _i, _i]),
        "orc_pack": (None, [_vp, _vp, _i, _vp, _vp, _vp, _i, _i, _i, _ll, _ll, _i, _i]),
        "orc_unpack": (None, [_vp, _vp, _i, _vp, _vp, _vp, _i, _i, _i, _ll, _ll, _i, _i]),
        "orc_exchange": (None, [_vp, _vp, _vp, _vp, _i, _i, _i, _ll, _ll, _i, _i]),
        "orc_num_threads": (_i, []),
        "orc_heat_grad_f64": (None, [_vp, _vp, _i, _i, _i, _i, _vp, _i, _i, _vp, _vp, _vp]),
        "orc_el_stress_f32": (None, [_vp, _vp, _vp, _i, _i, _i, _vp, C.c_float]),
        "orc_el_velocity_f32": (None, [_vp, _vp, _vp, _i, _i, _i, _vp, C.c_float]),
        "orc_el_forward_f32": (None, [_vp, _vp, _vp, _vp, _vp, _i, _i, _i, _vp, C.c_float, _i,
                                      _i, _vp, _vp, _vp, _i, _vp, _i, _vp]),
        "orc_el_forward_f64": (None, [_vp, _vp, _vp, _vp, _i, _i, _i, _vp, _d, _i, _i, _vp,
                                      _vp, _vp, _i, _vp, _i, _vp]),
        "orc_el_sumsq_f64": (_d, [_vp, _i, _i, _i]),
        "orc_el_adjoint_f64": (None, [_vp, _vp, _vp, _i, _i, _i, _vp, _d, _i, _i, _vp, _vp, _vp,
                                      _i, _vp, _i, _vp, _vp, _vp, _vp]),
    }
    for n, (r, a) in sigs.items():
        f = getattr(lib, n)
        f.restype = r
        f.argtypes = a
    return lib


lib = _load()


def _p(a):
    return a.ctypes.data_as(_vp)


def ref_available() -> bool:
    return os.path.exists(REF_SO)


_ref = None


def ref_lib():
    """The reference's heat-3D.cpp built by oracle/Makefile (None if absent)."""
    global _ref
    if _ref is None and ref_available():
        _ref = C.CDLL(REF_SO)
        _ref.ref_cpu_fdm_heat_diffuse1.restype = None
        _ref.ref_cpu_fdm_heat_diffuse1.argtypes = [_vp, _vp, _i, _i, _i, _i]
        _ref.ref_heat_3D_cpu_main.restype = C.c_float
        _ref.ref_heat_3D_cpu_main.argtypes = [_vp, _i, _i, _i, _i, _i]
    return _ref


# ---------------------------------------------------------------- generators
def uniform(shape, seed):
    a = np.empty(shape, np.float32)
    lib.orc_gen_uniform(_p(a), a.size, seed)
    return a


def bernoulli(shape, seed, p):
    a = np.empty(shape, np.uint8)
    lib.orc_gen_bernoulli(_p(a), a.size, seed, int(p * 16777216.0))
    return a


def blobs(env, amp, ex, ey, ez):
    gx, gy, gz = env
    out = np.empty((gz, gy, gx), np.float32)
    amp, ex, ey, ez = (np.ascontiguousarray(t, np.float64) for t in (amp, ex, ey, ez))
    lib.orc_gen_blobs(_p(out), gx, gy, gz, len(amp), _p(amp), _p(ex), _p(ey), _p(ez))
    return out


# ---------------------------------------------------------------- grids
def frame(a: np.ndarray, D: int, value=0):
    """Global owned grid (z,y,x) -> grid with a frozen frame of width D."""
    return np.pad(a, D, mode="constant", constant_values=value)


def unframe(a: np.ndarray, D: int):
    return a[D:-D, D:-D, D:-D]


def _owned(shape, D, periodic):
    """(gz, gy, gx) of an oracle grid of `shape` (z, y, x): a frozen axis
    carries a frame of D cells per side, a periodic one none (periodic: bool,
    or the per-axis mask of periodic_mask)."""
    pm = periodic_mask(periodic)
    return tuple(s if (pm >> a) & 1 else s - 2 * D for s, a in zip(shape, (2, 1, 0)))


def _cflag(periodic):
    pm = periodic_mask(periodic)
    return 1 if pm == 7 else pm


def heat_f64(u, D, k, nsteps, periodic):
    """u: framed (frozen) or plain (periodic) float64 grid, modified copy returned."""
    u = np.ascontiguousarray(u, np.float64).copy()
    gz, gy, gx = _owned(u.shape, D, periodic)
    kk = np.asarray(k, np.float64)
    lib.orc_heat_f64(_p(u), gx, gy, gz, D, _p(kk), nsteps, _cflag(periodic))
    return u


def heat_grad_f64(u0, tgt, D, k, nsteps, periodic):
    """FP64 adjoint: (J, dJ/du0 (same shape as u0, frame zero), dJ/dk[3]) for
    J = 1/2 sum_owned (u_N - tgt)^2; grids as heat_f64."""
    u0 = np.ascontiguousarray(u0, np.float64)
    tgt = np.ascontiguousarray(tgt, np.float64)
    assert u0.shape == tgt.shape
    gz, gy, gx = u0.shape if periodic else tuple(s - 2 * D for s in u0.shape)
    kk = np.asarray(k, np.float64)
    J = np.zeros(1)
    g = np.zeros_like(u0)
    gk = np.zeros(3)
    lib.orc_heat_grad_f64(_p(u0), _p(tgt), gx, gy, gz, D, _p(kk), nsteps, int(periodic),
                          _p(J), _p(g), _p(gk))
    return float(J[0]), g, gk


def heat_f32(u, D, mode, k, adt, h2, nsteps, periodic):
    u = np.ascontiguousarray(u, np.float32).copy()
    gz, gy, gx = _owned(u.shape, D, periodic)
    kk = np.asarray(k, np.float32)
    hh = np.asarray(h2, np.float64)
    lib.orc_heat_f32(_p(u), gx, gy, gz, D, mode, _p(kk), adt, _p(hh), nsteps, _cflag(periodic))
    return u


def gol(u, nsteps, periodic):
    u = np.ascontiguousarray(u, np.uint8).copy()
    gz, gy, gx = _owned(u.shape, 1, periodic)
    lib.orc_gol(_p(u), gx, gy, gz, nsteps, _cflag(periodic))
    return u


def coef_f64(D):
    c = np.zeros(5, np.float64)
    lib.orc_heat_coef_f64(D, _p(c))
    return c


def coef_f32(D):
    c = np.zeros(5, np.float32)
    lib.orc_heat_coef_f32(D, _p(c))
    return c


def ref1_block_step(h1, b2r_i, adt=0.1 * 0.1, h2=(1.0, 1.0, 0.25)):
    """Oracle restatement of cpu_fdm_heat_diffuse1 on a padded block (z,y,x)."""
    h1 = np.ascontiguousarray(h1, np.float32)
    out = h1.copy()
    nz, ny, nx = h1.shape
    lib.orc_heat_ref1_block_step(_p(h1), _p(out), nx, ny, nz, b2r_i, adt, *h2)
    return out


def reference_block_step(h1, b2r_i):
    """The reference's own cpu_fdm_heat_diffuse1 (heat-3D.cpp:48-62)."""
    r = ref_lib()
    h1 = np.ascontiguousarray(h1, np.float32)
    out = h1.copy()
    nz, ny, nx = h1.shape
    r.ref_cpu_fdm_heat_diffuse1(_p(h1), _p(out), nx, ny, nz, b2r_i)
    return out


# ---------------------------------------------------------------- exchange
def periodic_mask(periodic):
    """0/False: no axis wraps; 1/True: every axis (7); otherwise the per-axis
    mask itself (bit a: axis a wraps; a 2D grid wraps x and y only, 3), as the
    oracle's C functions read it."""
    pm = int(periodic)
    return 7 if pm == 1 else pm


def layout_args(lay, nb, esize, periodic):
    nb_ = np.asarray(nb, np.int32)
    B = np.asarray(list(lay.B), np.int32)
    S = np.asarray(list(lay.S), np.int32)
    return nb_, B, S, (lay.halo, lay.x_offset, lay.pitch, lay.slice, lay.block_stride, esize,
                       periodic_mask(periodic))


def exchange(raw, lay, nb, esize, periodic):
    raw = raw.copy()
    nb_, B, S, rest = layout_args(lay, nb, esize, periodic)
    lib.orc_exchange(_p(raw), _p(nb_), _p(B), _p(S), *rest)
    return raw


def pack(raw, lay, nb, esize, periodic, dim):
    nb_, B, S, rest = layout_args(lay, nb, esize, periodic)
    me = lib.orc_msg_elems(_p(nb_), _p(B), _p(S), lay.halo, dim)
    msgs = np.zeros(me * 2 * int(np.prod(nb)), raw.dtype)
    lib.orc_pack(_p(raw), _p(msgs), dim, _p(nb_), _p(B), _p(S), *rest)
    return msgs
