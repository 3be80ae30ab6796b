"""Seeded synthetic inputs for the elastic workloads (no counterpart in the
reference, whose eqwp model is a constant medium with a force patch,
eqwp.cu:9-25, 561-566).

All arrays are (z, y, x) with x fastest, the global owned grid of a context.
y is the vertical axis, as in the reference (its source row and force act on
y, eqwp.cu:135, 565).
"""
from __future__ import annotations

import math

import numpy as np


def layered_medium(env, seed, nlayers=10, vp_range=(2000.0, 4500.0), vs_ratio=1.73,
                   rho_range=(2000.0, 2700.0)):
    """nlayers equal-thickness layers along y with vp ~ U(vp_range),
    vs = vp / vs_ratio, rho ~ U(rho_range).  Returns float32 (vp, vs, rho)."""
    gx, gy, gz = env
    rng = np.random.default_rng(seed)
    vp_l = rng.uniform(*vp_range, nlayers)
    rho_l = rng.uniform(*rho_range, nlayers)
    layer = np.minimum((np.arange(gy) * nlayers) // gy, nlayers - 1)
    vp = np.broadcast_to(vp_l[layer][None, :, None], (gz, gy, gx)).astype(np.float32)
    rho = np.broadcast_to(rho_l[layer][None, :, None], (gz, gy, gx)).astype(np.float32)
    vs = (vp.astype(np.float64) / vs_ratio).astype(np.float32)
    return vp, vs, rho


def perturbed_layers(arrays, env, seed, nlayers=10, frac=0.05):
    """The observed-data model: every array scaled per layer along y by
    (1 + frac*xi), xi ~ U(-1, 1), independently per array and layer."""
    gx, gy, gz = env
    rng = np.random.default_rng(seed)
    layer = np.minimum((np.arange(gy) * nlayers) // gy, nlayers - 1)
    out = []
    for a in arrays:
        f = 1.0 + frac * rng.uniform(-1.0, 1.0, nlayers)
        out.append((np.asarray(a, np.float64) * f[layer][None, :, None]).astype(np.float32))
    return out


def lame(vp, vs, rho):
    """lambda = rho (vp^2 - 2 vs^2), mu = rho vs^2, b = 1/rho (FP64, rounded to f32)."""
    vp, vs, rho = (np.asarray(a, np.float64) for a in (vp, vs, rho))
    lam = rho * (vp * vp - 2.0 * vs * vs)
    mu = rho * vs * vs
    return lam.astype(np.float32), mu.astype(np.float32), (1.0 / rho).astype(np.float32)


def ricker(f0, dt, nt, t0=None):
    """Ricker wavelet (1 - 2 (pi f0 (t-t0))^2) exp(-(pi f0 (t-t0))^2), t0 = 1.2/f0."""
    t0 = 1.2 / f0 if t0 is None else t0
    t = np.arange(nt, dtype=np.float64) * dt - t0
    a = (math.pi * f0 * t) ** 2
    return ((1.0 - 2.0 * a) * np.exp(-a)).astype(np.float32)


def cfl_dt(vp_max, dx, cfl=0.4):
    return cfl * dx / vp_max


def surface_receivers(env, n=64, depth=2):
    """n x n receivers on the plane y = depth, evenly spread over x and z."""
    gx, gy, gz = env
    xs = np.linspace(0, gx - 1, n).round().astype(np.int32)
    zs = np.linspace(0, gz - 1, n).round().astype(np.int32)
    X, Z = np.meshgrid(xs, zs, indexing="xy")
    return np.stack([X.ravel(), np.full(X.size, depth, np.int32), Z.ravel()], axis=1)


def shot_sizes(nshots, seed, lo=(256, 256, 256), hi=(512, 512, 256)):
    """Seeded grid sizes between lo and hi (multiples of 16)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nshots):
        out.append(tuple(int(rng.integers(lo[a] // 16, hi[a] // 16 + 1)) * 16 for a in range(3)))
    return out
