"""Pins the elastic oracle (no GPU): FD4 exactness on polynomials, a closed-form
one-step update, energy boundedness, mirror symmetry, FP32 vs FP64 agreement,
and the FP64 adjoint gradient against central finite differences."""
import numpy as np
import pytest

import _elastic as E
from _util import rel_l2


def _const_medium(env, lam=2.0e9, mu=1.0e9, rho=2000.0):
    gx, gy, gz = env
    sh = (gz, gy, gx)
    return (np.full(sh, lam, np.float64), np.full(sh, mu, np.float64), np.full(sh, 1 / rho, np.float64))


def test_fd4_exact_on_quartic_one_stress_step():
    # v = (x^4, y^3, z^2) (grid units); FD4 is exact for degree <= 4, so one
    # stress stage gives sigma = dt*C:grad v exactly; the velocity stage sees
    # an exact divergence of a cubic stress field as well.
    env = (12, 12, 12)
    gx, gy, gz = env
    lam, mu, b = _const_medium(env, 3.0, 2.0, 1.0)
    z, y, x = np.meshgrid(*(np.arange(n, dtype=np.float64) for n in (gz, gy, gx)), indexing="ij")
    s0 = [x ** 4 / 100, y ** 3 / 10, z ** 2, 0 * x, 0 * x, 0 * x, 0 * x, 0 * x, 0 * x]
    # the oracle's frame is 2 cells of the same polynomial: pad by evaluation
    st, _ = E.forward(env, lam, mu, b, 0.5, (1.0, 1.0, 1.0), 0, fp64=True, state0=s0)
    import _oracle as O
    import ctypes as C
    full = [np.zeros((gz + 4, gy + 4, gx + 4)) for _ in range(9)]
    Z, Y, X = np.meshgrid(*(np.arange(-2, n + 2, dtype=np.float64) for n in (gz, gy, gx)),
                          indexing="ij")
    full[0][:] = X ** 4 / 100
    full[1][:] = Y ** 3 / 10
    full[2][:] = Z ** 2
    L, M, B = (np.pad(a, 2, mode="edge") for a in (lam, mu, b))
    ptrs = (C.c_void_p * 9)(*[a.ctypes.data for a in full])
    O.lib.orc_el_forward_f64(ptrs, O._p(L), O._p(M), O._p(B), gx, gy, gz,
                             O._p(np.ones(3)), 0.5, 1, 0, None, None, None, 0, None, 1, None)
    exx, eyy, ezz = 4 * X ** 3 / 100, 3 * Y ** 2 / 10, 2 * Z
    s = slice(2, -2)
    l2m = 3.0 + 2 * 2.0
    np.testing.assert_allclose(full[3][s, s, s], (0.5 * (l2m * exx + 3.0 * (eyy + ezz)))[s, s, s],
                               rtol=1e-13, atol=1e-9)
    np.testing.assert_allclose(full[4][s, s, s], (0.5 * (l2m * eyy + 3.0 * (exx + ezz)))[s, s, s],
                               rtol=1e-13, atol=1e-9)
    assert np.abs(full[6][s, s, s]).max() == 0  # shear strain of this v is zero


def test_energy_bounded_without_source():
    env = (16, 16, 16)
    lam, mu, b = _const_medium(env)
    rng = np.random.default_rng(0)
    s0 = [np.zeros((16, 16, 16)) for _ in range(9)]
    s0[1][6:10, 6:10, 6:10] = rng.uniform(-1, 1, (4, 4, 4))
    dx = 10.0
    vp = np.sqrt((2e9 + 2e9) / 2000)
    dt = 0.4 * dx / vp
    e0 = sum((a ** 2).sum() for a in s0)
    st, _ = E.forward(env, lam, mu, b, dt, (1 / dx,) * 3, 200, fp64=True, state0=s0)
    e = [float((a ** 2).sum()) for a in st]
    assert np.isfinite(sum(e)) and sum(e[:3]) < 50 * e0


def test_mirror_symmetry():
    # a vy source on the x mid-plane of a x-symmetric medium: vy(x) = vy(-x)
    env = (17, 14, 15)
    lam, mu, b = _const_medium(env)
    src = [[8, 6, 7]]
    wav = np.ones((20, 1)) * 1e3
    st, _ = E.forward(env, lam, mu, b, 1e-3, (0.1,) * 3, 20, src, [1], wav, fp64=True)
    vy = st[1]
    np.testing.assert_allclose(vy, vy[:, :, ::-1], rtol=1e-12, atol=1e-15)
    vx = st[0]
    np.testing.assert_allclose(vx, -vx[:, :, ::-1], rtol=1e-12, atol=1e-15)


def test_f32_close_to_f64():
    P = E.small_problem()
    a = E.forward(P["env"], P["lam"], P["mu"], P["b"], P["dt"], P["ih"], P["nsteps"], P["src"],
                  P["comp"], P["wav"], P["rec"], 1)
    d = E.forward(P["env"], P["lam"], P["mu"], P["b"], float(P["dt"]), P["ih"], P["nsteps"],
                  P["src"], P["comp"], P["wav"].astype(np.float64), P["rec"], 1, fp64=True)
    assert rel_l2(a[1], d[1]) < 1e-5
    for k in range(9):
        assert rel_l2(a[0][k], d[0][k]) < 1e-5


def _misfit64(P, lam, mu, b, obs):
    _, rec = E.forward(P["env"], lam, mu, b, float(P["dt"]), P["ih"], P["nsteps"], P["src"],
                       P["comp"], P["wav"].astype(np.float64), P["rec"], 1, fp64=True)
    return 0.5 * float(((rec - obs) ** 2).sum()), rec


@pytest.mark.parametrize("which", [0, 1, 2])
def test_adjoint_gradient_matches_finite_differences(which):
    P = E.small_problem(env=(14, 12, 13), nsteps=24)
    lam, mu, b = (P[k].astype(np.float64) for k in ("lam", "mu", "b"))
    obs = _misfit64(P, lam * 1.02, mu * 0.97, b, np.zeros((24, len(P["rec"]))))[1]
    J0, rec = _misfit64(P, lam, mu, b, obs)
    g = E.adjoint(P["env"], lam, mu, b, float(P["dt"]), P["ih"], 24, P["src"], P["comp"],
                  P["wav"].astype(np.float64), P["rec"], 1, rec - obs)
    rng = np.random.default_rng(which)
    m = [lam, mu, b]
    for trial in range(2):
        dirn = rng.standard_normal(m[which].shape) * m[which]
        eps = 1e-6
        mp = [a.copy() for a in m]
        mm = [a.copy() for a in m]
        mp[which] = m[which] + eps * dirn
        mm[which] = m[which] - eps * dirn
        fd = (_misfit64(P, *mp, obs)[0] - _misfit64(P, *mm, obs)[0]) / (2 * eps)
        an = float((g[which] * dirn).sum())
        assert abs(fd - an) <= 1e-6 * max(abs(fd), abs(an)) + 1e-30, (fd, an)


def _periodic_case(env, src, nsteps, seed=5):
    P = E.small_problem(env=env, seed=seed, nsteps=nsteps)
    return P, np.array([src], np.int32)


def test_periodic_forward_is_translation_equivariant():
    # a source one cell from the x face: the wave crosses the faces within the run,
    # so a frame filled any other way than by wrap breaks the equivariance
    env, nsteps = (16, 14, 12), 10
    P, src = _periodic_case(env, (1, 7, 6), nsteps)
    rec = np.array([[0, 0, 0], [15, 13, 11], [3, 7, 6]], np.int32)
    st, r = E.forward_periodic(env, P["lam"], P["mu"], P["b"], P["dt"], P["ih"], nsteps, src,
                               [1], P["wav"], rec, 1)
    assert np.abs(st[1][:, :, -3:]).max() > 0  # reached across the x face
    sh = (3, -4, 5)  # shift along z, y, x

    def roll(a):
        return np.roll(a, sh, axis=(0, 1, 2))
    s2 = (src + np.array([sh[2], sh[1], sh[0]])) % np.array(env)
    r2 = (rec + np.array([sh[2], sh[1], sh[0]])) % np.array(env)
    st2, rr2 = E.forward_periodic(env, roll(P["lam"]), roll(P["mu"]), roll(P["b"]), P["dt"],
                                  P["ih"], nsteps, s2, [1], P["wav"], r2, 1)
    for k in range(9):
        assert np.array_equal(roll(st[k]), st2[k]), k
    assert np.array_equal(r, rr2)


def test_periodic_forward_equals_frozen_away_from_faces():
    # while the wave stays two cells inside the faces, wrap and zero frame agree
    env, nsteps = (40, 40, 40), 3
    P, src = _periodic_case(env, (20, 19, 21), nsteps)
    st, _ = E.forward_periodic(env, P["lam"], P["mu"], P["b"], P["dt"], P["ih"], nsteps, src,
                               [1], P["wav"])
    fz, _ = E.forward(env, P["lam"], P["mu"], P["b"], P["dt"], P["ih"], nsteps, src, [1],
                      P["wav"])
    assert np.abs(st[1]).max() > 0
    for k in range(9):
        assert not np.any(st[k][:2]) and not np.any(st[k][-2:]), k
        assert np.array_equal(st[k], fz[k]), k


# ---------------------------------------------------------------- closed forms
class Poly:
    """A polynomial in (x, y, z) as {(i, j, k): coefficient}, for exact
    derivatives of the fields a one-step closed form needs."""

    def __init__(self, terms=None):
        self.t = {k: v for k, v in (terms or {}).items() if v != 0}

    def __add__(self, o):
        t = dict(self.t)
        for k, v in o.t.items():
            t[k] = t.get(k, 0.0) + v
        return Poly(t)

    def scale(self, a):
        return Poly({k: a * v for k, v in self.t.items()})

    def d(self, axis):
        t = {}
        for (i, j, k), v in self.t.items():
            e = [i, j, k]
            if e[axis] == 0:
                continue
            c = v * e[axis]
            e[axis] -= 1
            t[tuple(e)] = t.get(tuple(e), 0.0) + c
        return Poly(t)

    def __call__(self, x, y, z):
        out = np.zeros(np.broadcast(x, y, z).shape)
        for (i, j, k), v in self.t.items():
            out = out + v * x ** i * y ** j * z ** k
        return out


def _one_step_closed_form(lam, mu, b, dt):
    """v0, sigma0 (degree <= 3, every component different) and the exact state
    after one step: sigma' = sigma0 + dt C E v0, v' = v0 + dt b div sigma'.
    Both stages differentiate polynomials of degree <= 3, which FD4 does exactly."""
    P = Poly
    v0 = [P({(2, 1, 0): 1 / 50}),                               # vx = x^2 y / 50
          P({(0, 3, 0): 1 / 30, (1, 0, 1): 1 / 7}),             # vy = y^3/30 + xz/7
          P({(1, 0, 2): 1 / 20})]                               # vz = x z^2 / 20
    s0 = [P({(3, 0, 0): 1 / 40}), P({(0, 2, 1): 1 / 15}), P({(0, 0, 3): 1 / 25}),  # xx yy zz
          P({(2, 1, 0): 1 / 12, (0, 0, 1): 1.0}),               # sxy
          P({(1, 1, 1): 1 / 9, (0, 2, 0): -0.3}),               # sxz
          P({(0, 2, 0): 1 / 11, (1, 0, 2): 1 / 13})]            # syz
    exx, eyy, ezz = v0[0].d(0), v0[1].d(1), v0[2].d(2)
    dxy = v0[0].d(1) + v0[1].d(0)
    dxz = v0[0].d(2) + v0[2].d(0)
    dyz = v0[1].d(2) + v0[2].d(1)
    l2m = lam + 2 * mu
    s1 = [s0[0] + (exx.scale(l2m) + (eyy + ezz).scale(lam)).scale(dt),
          s0[1] + (eyy.scale(l2m) + (exx + ezz).scale(lam)).scale(dt),
          s0[2] + (ezz.scale(l2m) + (exx + eyy).scale(lam)).scale(dt),
          s0[3] + dxy.scale(mu * dt), s0[4] + dxz.scale(mu * dt), s0[5] + dyz.scale(mu * dt)]
    sxx, syy, szz, sxy, sxz, syz = s1
    ax = sxx.d(0) + sxy.d(1) + sxz.d(2)
    ay = sxy.d(0) + syy.d(1) + syz.d(2)
    az = sxz.d(0) + syz.d(1) + szz.d(2)
    v1 = [v0[0] + ax.scale(dt * b), v0[1] + ay.scale(dt * b), v0[2] + az.scale(dt * b)]
    return v0 + s0, v1 + s1


def _framed_eval(polys, env, dtype):
    gx, gy, gz = env
    Z, Y, X = np.meshgrid(*(np.arange(-2, n + 2, dtype=np.float64) for n in (gz, gy, gx)),
                          indexing="ij")
    return [np.ascontiguousarray(p(X, Y, Z), dtype) for p in polys]


@pytest.mark.parametrize("fp64", [True, False])
def test_one_step_velocity_and_stress_closed_form(fp64):
    # every one of the nine fields after one whole step (stress stage, then the
    # velocity stage with the new stresses) against the exact polynomial; a
    # component swapped in the divergence (say d/dx sxz for d/dx sxy in the y
    # velocity) changes v' by a polynomial that is not zero here
    import ctypes as C
    import _oracle as O
    env = (14, 13, 12)
    gx, gy, gz = env
    lam, mu, b, dt = 3.0, 2.0, 0.5, 0.25
    start, want = _one_step_closed_form(lam, mu, b, dt)
    dtype = np.float64 if fp64 else np.float32
    full = _framed_eval(start, env, dtype)
    ex = _framed_eval(want, env, np.float64)
    shape = (gz + 4, gy + 4, gx + 4)
    L, M, B = (np.full(shape, v, dtype) for v in (lam, mu, b))
    ptrs = (C.c_void_p * 9)(*[a.ctypes.data for a in full])
    ih = np.ones(3, dtype)
    if fp64:
        O.lib.orc_el_forward_f64(ptrs, O._p(L), O._p(M), O._p(B), gx, gy, gz, O._p(ih), dt, 1, 0,
                                 None, None, None, 0, None, 1, None)
    else:
        O.lib.orc_el_stress_f32(ptrs, O._p(L), O._p(M), gx, gy, gz, O._p(ih), dt)
        O.lib.orc_el_velocity_f32(ptrs, O._p(B), None, gx, gy, gz, O._p(ih), dt)
    # stresses: every owned cell (v is exact in the frame); velocities: owned
    # cells two or more inside, whose stencils read only updated stresses
    own, inner = slice(2, -2), slice(4, -4)
    for k in range(9):
        s = (own, own, own) if k >= 3 else (inner, inner, inner)
        got, e = full[k][s].astype(np.float64), ex[k][s]
        scale = np.abs(e).max()
        tol = 1e-12 if fp64 else 2e-6
        assert np.abs(got - e).max() <= tol * scale, (k, np.abs(got - e).max() / scale)
        # the step changed this field measurably (the check is not vacuous)
        assert np.abs(e - _framed_eval([start[k]], env, np.float64)[0][s]).max() > 1e-3 * scale


def _reciprocity_traces(env, P, s, i, r, j, fp64):
    """v_j at r from a force in component i at s (FP64 or FP32 oracle)."""
    dt = float(P["dt"]) if fp64 else P["dt"]
    wav = P["wav"].astype(np.float64 if fp64 else np.float32)
    lam, mu, b = ((P[k].astype(np.float64) for k in ("lam", "mu", "b")) if fp64 else
                  (P["lam"], P["mu"], P["b"]))
    _, rec = E.forward(env, lam, mu, b, dt, P["ih"], P["nsteps"], [s], [i], wav, [r], j,
                       fp64=fp64)
    return rec[:, 0].astype(np.float64)


@pytest.mark.parametrize("fp64,i,j", [(True, 0, 1), (True, 1, 2), (True, 2, 2), (False, 0, 1)])
def test_source_receiver_reciprocity(fp64, i, j):
    # G_ji(r, s) = G_ij(s, r) on a heterogeneous medium with a zero frame.  In
    # the energy variables u = b^-1/2 v, q = C^-1/2 sigma a step is
    # q' = q + dt K u, u' = u - dt K^T q' with K = C^1/2 E b^1/2 (the velocity
    # stage's divergence is -E^T on the zero-framed grid), so with
    # S = diag(-1, 1) the step matrix P obeys S P S = P^T and the velocity block
    # of every power of P is symmetric: the discrete scheme is exactly
    # reciprocal.  Any divergence term that is not the transpose of the strain
    # rate breaks it.
    env = (18, 16, 15)
    P = E.small_problem(env=env, seed=11, nsteps=40)
    s, r = (5, 6, 4), (12, 9, 10)
    a = _reciprocity_traces(env, P, s, i, r, j, fp64)
    c = _reciprocity_traces(env, P, r, j, s, i, fp64)
    assert np.abs(a).max() > 0
    tol = 1e-11 if fp64 else 1e-4
    assert np.abs(a - c).max() <= tol * np.abs(a).max(), np.abs(a - c).max() / np.abs(a).max()


def test_displacement_takes_the_y_force():
    # eqwp.cu:229-235: disp += Dt*Dt*inv_rho*(div sigma + f.y); from a zero state
    # the first step's divergence is zero, so the displacement after one step is
    # dt*(dt*b)*f at a y source and zero elsewhere (an x source adds nothing)
    env = (10, 9, 8)
    P = E.small_problem(env=env, seed=2, nsteps=1)
    src = np.array([[4, 4, 3], [6, 2, 5]], np.int32)
    wav = np.array([[7.5e3, -3.0e3]], np.float32)
    _, _, disp = E.forward(env, P["lam"], P["mu"], P["b"], P["dt"], P["ih"], 1, src, [1, 0],
                           wav, disp=True)
    want = np.zeros_like(disp)
    x, y, z = src[0]
    bz = P["b"][z, y, x]
    want[z, y, x] = (P["dt"] * (P["dt"] * bz)) * wav[0, 0]
    assert want[z, y, x] != 0
    assert np.array_equal(disp, want)
