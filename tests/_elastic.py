"""Elastic helpers for the tests: framed-grid wrappers of the CPU oracle
(oracle/b2r_oracle_elastic.c) and small seeded problems.  Checker only."""
from __future__ import annotations

import numpy as np

import _oracle as O

_p = O._p


def _frame(a, dtype):
    return np.pad(np.asarray(a, dtype), 2, mode="edge")


def _unframe(a):
    return a[2:-2, 2:-2, 2:-2]


def _state(gshape, dtype, state0=None):
    st = [np.zeros(tuple(s + 4 for s in gshape), dtype) for _ in range(9)]
    if state0 is not None:
        for k in range(9):
            st[k][2:-2, 2:-2, 2:-2] = state0[k]
    return st


def _ptrs(arrs):
    import ctypes as C
    return (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


def forward(env, lam, mu, b, dt, ih, nsteps, src_xyz=(), src_comp=(), wav=None,
            rec_xyz=(), rec_comp=1, fp64=False, state0=None, disp=False):
    """Returns (owned state list of 9, receivers nt x nrec[, displacement])."""
    gx, gy, gz = env
    dt_ = np.float64 if fp64 else np.float32
    st = _state((gz, gy, gx), dt_, state0)
    L, M, B = (_frame(a, dt_) for a in (lam, mu, b))
    src_xyz = np.ascontiguousarray(np.asarray(src_xyz, np.int32).reshape(-1, 3))
    src_comp = np.ascontiguousarray(np.asarray(src_comp, np.int32).ravel())
    nsrc = len(src_comp)
    wav = np.ascontiguousarray(np.zeros((nsteps, max(nsrc, 1))) if wav is None else wav, dt_)
    rec_xyz = np.ascontiguousarray(np.asarray(rec_xyz, np.int32).reshape(-1, 3))
    nrec = len(rec_xyz)
    rec = np.zeros((nsteps, max(nrec, 1)), dt_)
    ihs = np.asarray(ih, dt_)
    if fp64:
        O.lib.orc_el_forward_f64(_ptrs(st), _p(L), _p(M), _p(B), gx, gy, gz, _p(ihs), float(dt),
                                 nsteps, nsrc, _p(src_xyz), _p(src_comp), _p(wav), nrec,
                                 _p(rec_xyz), rec_comp, _p(rec))
        d = None
    else:
        d = np.zeros_like(st[0]) if disp else None
        O.lib.orc_el_forward_f32(_ptrs(st), _p(L), _p(M), _p(B), _p(d) if disp else None, gx, gy,
                                 gz, _p(ihs), float(dt), nsteps, nsrc, _p(src_xyz), _p(src_comp),
                                 _p(wav), nrec, _p(rec_xyz), rec_comp, _p(rec))
    out = ([_unframe(a) for a in st], rec[:, :nrec])
    return out + (_unframe(d),) if disp else out


def _wrap(a):
    a[...] = np.pad(a[2:-2, 2:-2, 2:-2], 2, mode="wrap")


def forward_periodic(env, lam, mu, b, dt, ih, nsteps, src_xyz=(), src_comp=(), wav=None,
                     rec_xyz=(), rec_comp=1):
    """FP32 forward on a periodic grid: the oracle's stress and velocity stages
    (orc_el_stress_f32, orc_el_velocity_f32), each after filling the 2-cell frame
    of the fields it differentiates by periodic wrap; point forces and receivers
    as orc_el_forward_f32.  Returns (owned state list of 9, receivers nt x nrec)."""
    gx, gy, gz = env
    f32 = np.float32
    st = _state((gz, gy, gx), f32)
    L, M, B = (np.pad(np.asarray(a, f32), 2, mode="wrap") for a in (lam, mu, b))
    src_xyz = np.asarray(src_xyz, np.int32).reshape(-1, 3)
    src_comp = np.asarray(src_comp, np.int32).ravel()
    rec_xyz = np.asarray(rec_xyz, np.int32).reshape(-1, 3)
    rec = np.zeros((nsteps, len(rec_xyz)), f32)
    ihs = np.asarray(ih, f32)
    dt = f32(dt)
    for t in range(nsteps):
        for k in range(3):
            _wrap(st[k])
        O.lib.orc_el_stress_f32(_ptrs(st), _p(L), _p(M), gx, gy, gz, _p(ihs), float(dt))
        for k in range(3, 9):
            _wrap(st[k])
        O.lib.orc_el_velocity_f32(_ptrs(st), _p(B), None, gx, gy, gz, _p(ihs), float(dt))
        for i, (x, y, z) in enumerate(src_xyz):
            c = (z + 2, y + 2, x + 2)
            tb = dt * B[c]
            st[src_comp[i]][c] = st[src_comp[i]][c] + tb * f32(wav[t, i])
        for r, (x, y, z) in enumerate(rec_xyz):
            rec[t, r] = st[rec_comp][z + 2, y + 2, x + 2]
    return [_unframe(a) for a in st], rec


def adjoint(env, lam, mu, b, dt, ih, nsteps, src_xyz, src_comp, wav, rec_xyz, rec_comp, res):
    """FP64 gradients (g_lam, g_mu, g_b) of 1/2 |rec - obs|^2 given res = rec - obs."""
    gx, gy, gz = env
    L, M, B = (_frame(a, np.float64) for a in (lam, mu, b))
    g = [np.zeros_like(L) for _ in range(3)]
    src_xyz = np.ascontiguousarray(np.asarray(src_xyz, np.int32).reshape(-1, 3))
    src_comp = np.ascontiguousarray(np.asarray(src_comp, np.int32).ravel())
    wav = np.ascontiguousarray(wav, np.float64)
    rec_xyz = np.ascontiguousarray(np.asarray(rec_xyz, np.int32).reshape(-1, 3))
    res = np.ascontiguousarray(res, np.float64)
    ihs = np.asarray(ih, np.float64)
    O.lib.orc_el_adjoint_f64(_p(L), _p(M), _p(B), gx, gy, gz, _p(ihs), float(dt), nsteps,
                             len(src_comp), _p(src_xyz), _p(src_comp), _p(wav), len(rec_xyz),
                             _p(rec_xyz), rec_comp, _p(res), _p(g[0]), _p(g[1]), _p(g[2]))
    return [_unframe(a) for a in g]


def small_problem(env=(20, 18, 16), seed=3, nsteps=30):
    """A heterogeneous medium, one Ricker source in vy, a few vy receivers."""
    import sys
    import os
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from b2r_amd import workloads as W
    vp, vs, rho = W.layered_medium(env, seed, nlayers=4)
    rng = np.random.default_rng(seed)
    vp = (vp * rng.uniform(0.95, 1.05, vp.shape)).astype(np.float32)
    lam, mu, b = W.lame(vp, vs, rho)
    dx = 10.0
    dt = np.float32(W.cfl_dt(float(vp.max()), dx, 0.4))
    ih = (np.float32(1 / dx),) * 3
    src = np.array([[env[0] // 2, env[1] // 3, env[2] // 2]], np.int32)
    wav = (W.ricker(25.0, float(dt), nsteps) * np.float32(1e3))[:, None]
    rec = np.array([[x, 2, z] for x in range(2, env[0] - 2, 4) for z in range(2, env[2] - 2, 5)],
                   np.int32)
    return dict(env=env, vp=vp, vs=vs, rho=rho, lam=lam, mu=mu, b=b, dt=dt, ih=ih, src=src,
                comp=[1], wav=wav, rec=rec, nsteps=nsteps)
