"""The seeded elastic workload generators (b2r_amd/workloads.py) against what
their definitions fix: the Ricker wavelet's closed-form peak, zero crossings,
symmetry and spectral peak; the layered medium's per-layer constancy, ranges
and vs/vp ratio; the Lame parameters inverting back to the velocities; the
5% per-layer perturbation; the CFL number; the receiver plane."""
import math

import numpy as np

from b2r_amd import workloads as W


def test_ricker_closed_form():
    f0, dt = 10.0, 1e-3
    t0 = 1.2 / f0  # 120 samples: the peak lies on a sample
    w = W.ricker(f0, dt, 400).astype(np.float64)
    k0 = int(round(t0 / dt))
    assert w.argmax() == k0 and abs(w[k0] - 1.0) < 1e-7
    # symmetric about t0
    assert np.allclose(w[k0 - 100:k0], w[k0 + 100:k0:-1], atol=1e-7)
    # zero crossings at t0 +- 1/(pi f0 sqrt 2)
    tz = 1.0 / (math.pi * f0 * math.sqrt(2.0))
    t = np.arange(400) * dt - t0
    for sgn in (-1, 1):
        k = np.where(np.sign(w[:-1]) != np.sign(w[1:]))[0]
        kk = k[np.argmin(np.abs(t[k] - sgn * tz))]
        # linear interpolation of the crossing between samples kk and kk+1
        tc = t[kk] - w[kk] * dt / (w[kk + 1] - w[kk])
        assert abs(tc - sgn * tz) < 2e-5
    # the minima 2 exp(-3/2) below zero at t0 +- sqrt(3/2)/(pi f0)
    assert abs(w.min() + 2.0 * math.exp(-1.5)) < 2e-3
    # amplitude spectrum f^2 exp(-f^2/f0^2) peaks at f0; zero mean
    n = 1 << 16
    spec = np.abs(np.fft.rfft(w, n))
    freq = np.fft.rfftfreq(n, dt)
    assert abs(freq[spec.argmax()] - f0) < 0.05
    assert abs(w.sum()) * dt < 1e-6


def test_layered_medium_structure():
    env = (24, 50, 16)
    vp, vs, rho = W.layered_medium(env, 7)
    assert vp.shape == (16, 50, 24) and vp.dtype == np.float32
    # constant along x and z
    for a in (vp, vs, rho):
        assert (a == a[:1, :, :1]).all()
    col = vp[0, :, 0]
    # 10 layers of 5 rows each, values distinct between layers
    layers = col.reshape(10, 5)
    assert (layers == layers[:, :1]).all() and len(np.unique(layers[:, 0])) == 10
    assert (col >= 2000.0).all() and (col <= 4500.0).all()
    assert (rho >= 2000.0).all() and (rho <= 2700.0).all()
    assert np.allclose(vs.astype(np.float64) * 1.73, vp, rtol=1e-6)
    # the same seed gives the same medium, another seed another one
    assert np.array_equal(W.layered_medium(env, 7)[0], vp)
    assert not np.array_equal(W.layered_medium(env, 8)[0], vp)


def test_lame_inverts_to_velocities():
    vp, vs, rho = W.layered_medium((8, 40, 8), 3)
    lam, mu, b = W.lame(vp, vs, rho)
    lam, mu, b = (a.astype(np.float64) for a in (lam, mu, b))
    assert (lam > 0).all() and (mu > 0).all()  # vs = vp/1.73 keeps lambda positive
    assert np.allclose(np.sqrt((lam + 2 * mu) * b), vp, rtol=1e-6)
    assert np.allclose(np.sqrt(mu * b), vs, rtol=1e-6)
    assert np.allclose(1.0 / b, rho, rtol=1e-6)


def test_perturbed_layers_within_five_percent():
    env = (8, 30, 8)
    base = W.layered_medium(env, 5)
    pert = W.perturbed_layers(base, env, 6)
    for a, p in zip(base, pert):
        r = p.astype(np.float64) / a
        assert (np.abs(r - 1.0) <= 0.05 + 1e-6).all()
        assert np.allclose(r[0, :, 0].reshape(10, 3), r[0, :, 0].reshape(10, 3)[:, :1], rtol=1e-6)
        assert np.allclose(r, r[:1, :, :1], rtol=1e-6)
    # independent factors per array
    ratios = [(p.astype(np.float64) / a)[0, ::3, 0] for a, p in zip(base, pert)]
    assert not np.allclose(ratios[0], ratios[2])


def test_cfl_receivers_and_sizes():
    assert abs(W.cfl_dt(4500.0, 10.0, 0.4) * 4500.0 / 10.0 - 0.4) < 1e-12
    env = (256, 200, 128)
    r = W.surface_receivers(env)
    assert r.shape == (4096, 3) and (r[:, 1] == 2).all()
    assert len({(int(x), int(z)) for x, _, z in r}) == 4096
    assert r[:, 0].min() == 0 and r[:, 0].max() == 255 and r[:, 2].max() == 127
    sizes = W.shot_sizes(16, 4)
    assert len(sizes) == 16 and sizes == W.shot_sizes(16, 4)
    for s in sizes:
        assert all(v % 16 == 0 for v in s)
        assert all(256 <= s[a] <= 512 for a in (0, 1)) and s[2] == 256
