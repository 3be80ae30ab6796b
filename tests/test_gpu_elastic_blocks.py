"""The reverse-sweep elastic gradient on several blocks and with periodic
faces: equal bit for bit to the one-block gradient (the blocks' exchanges
copy owned values, and the ring cells the reverse step rebuilds are the
owners' values recomputed from the same inputs), against FP64 central
differences of an independent numpy periodic forward (FD4 by np.roll), and
the dot-product identity of the periodic adjoint."""
import sys

import numpy as np
import pytest

import _elastic as E
from _util import rel_l2

pytestmark = pytest.mark.gpu

A = None


def setup_module(module):
    global A
    from b2r_amd import _abi
    A = _abi


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a, np.float32).view(np.uint32),
                          np.ascontiguousarray(b, np.float32).view(np.uint32))


def _shot(ctx, P):
    s = A.ElasticShot(ctx, float(P["dt"]), P["ih"])
    s.set_medium(P["lam"], P["mu"], P["b"])
    s.set_sources(P["src"], P["comp"], P["wav"])
    s.set_receivers(P["rec"], 1, P["nsteps"])
    return s


def _obs(P, periodic):
    m = (P["lam"] * np.float32(1.03), P["mu"] * np.float32(0.98), P["b"])
    f = E.forward_periodic if periodic else E.forward
    _, obs = f(P["env"], *m, P["dt"], P["ih"], P["nsteps"], P["src"], P["comp"], P["wav"],
               P["rec"], 1)
    return obs


def _d4(f, axis, ih):
    """Collocated FD4 first derivative with periodic wrap (array axes z, y, x)."""
    c2, c1 = -1.0 / 12.0, 2.0 / 3.0
    return (c2 * (np.roll(f, -2, axis) - np.roll(f, 2, axis)) +
            c1 * (np.roll(f, -1, axis) - np.roll(f, 1, axis))) * ih


def _rec64_periodic(P, lam, mu, b):
    """FP64 receivers of the velocity-stress scheme on the periodic grid:
    stress stage from v, velocity stage from the new stresses, point forces
    dt*b*f after it, v_y sampled after every step."""
    gx, gy, gz = P["env"]
    ihx, ihy, ihz = (float(v) for v in P["ih"])
    dt = float(P["dt"])
    vx, vy, vz = (np.zeros((gz, gy, gx)) for _ in range(3))
    sxx, syy, szz, sxy, sxz, syz = (np.zeros((gz, gy, gx)) for _ in range(6))
    dx = lambda f: _d4(f, 2, ihx)
    dy = lambda f: _d4(f, 1, ihy)
    dz = lambda f: _d4(f, 0, ihz)
    l2m = lam + 2 * mu
    rec = np.zeros((P["nsteps"], len(P["rec"])))
    wav = P["wav"].astype(np.float64)
    for t in range(P["nsteps"]):
        exx, eyy, ezz = dx(vx), dy(vy), dz(vz)
        sxx = sxx + dt * (l2m * exx + lam * (eyy + ezz))
        syy = syy + dt * (l2m * eyy + lam * (exx + ezz))
        szz = szz + dt * (l2m * ezz + lam * (exx + eyy))
        sxy = sxy + dt * mu * (dy(vx) + dx(vy))
        sxz = sxz + dt * mu * (dz(vx) + dx(vz))
        syz = syz + dt * mu * (dz(vy) + dy(vz))
        vx = vx + dt * b * (dx(sxx) + dy(sxy) + dz(sxz))
        vy = vy + dt * b * (dx(sxy) + dy(syy) + dz(syz))
        vz = vz + dt * b * (dx(sxz) + dy(syz) + dz(szz))
        v = (vx, vy, vz)
        for i, (x, y, z) in enumerate(P["src"]):
            v[P["comp"][i]][z, y, x] += dt * b[z, y, x] * wav[t, i]
        for r, (x, y, z) in enumerate(P["rec"]):
            rec[t, r] = vy[z, y, x]
    return rec


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("blocks", [(2, 1, 2), (1, 2, 1), (2, 2, 2)])
def test_gradient_block_decomposition(periodic, blocks):
    P = E.small_problem(env=(24, 20, 16), nsteps=18)
    obs = _obs(P, periodic)
    kind = A.BOUNDARY_PERIODIC if periodic else A.BOUNDARY_FROZEN
    say("elastic gradient blocks", blocks, "periodic", periodic)
    out = []
    for bl in ((1, 1, 1), blocks):
        with A.Context(P["env"], bl, 1, 4, kind, 4) as c:
            with _shot(c, P) as s:
                out.append(s.gradient(obs, P["nsteps"], 5))
    (J1, g1), (J2, g2) = out
    assert J1 == J2
    for k in range(3):
        assert bits_equal(g1[k], g2[k]), (k, rel_l2(g2[k], g1[k]))


def test_periodic_gradient_finite_differences():
    P = E.small_problem(env=(16, 14, 12), nsteps=16)
    obs = _obs(P, True)
    with A.Context(P["env"], (2, 1, 1), 1, 4, A.BOUNDARY_PERIODIC, 4) as c:
        with _shot(c, P) as s:
            J, g = s.gradient(obs, P["nsteps"], 4)
    m = [P[k].astype(np.float64) for k in ("lam", "mu", "b")]
    obs64 = obs.astype(np.float64)

    def J64(mm):
        return 0.5 * float(((_rec64_periodic(P, *mm) - obs64) ** 2).sum())

    say("periodic misfit", J, J64(m))
    assert abs(J - J64(m)) <= 1e-4 * J64(m)
    rng = np.random.default_rng(5)
    for trial in range(6):
        which = trial % 3
        dirn = rng.standard_normal(m[which].shape) * m[which]
        eps = 1e-5
        mp = [a.copy() for a in m]
        mn = [a.copy() for a in m]
        mp[which] += eps * dirn
        mn[which] -= eps * dirn
        fd = (J64(mp) - J64(mn)) / (2 * eps)
        an = float((g[which].astype(np.float64) * dirn).sum())
        say("periodic FD", which, fd, an)
        assert abs(fd - an) <= 1e-3 * max(abs(fd), abs(an))


@pytest.mark.parametrize("blocks", [(1, 1, 1), (2, 1, 2)])
def test_periodic_adjoint_dot_product(blocks):
    # P: initial state -> receiver data on the periodic grid.  With obs = 0
    # the adjoint state at t = 0 is P^T P x0, so <x0, P^T P x0> = 2 J.
    P = E.small_problem(env=(16, 18, 20), nsteps=20)
    gx, gy, gz = P["env"]
    rng = np.random.default_rng(17)
    x0 = [rng.standard_normal((gz, gy, gx)).astype(np.float32) * np.float32(1e-3)
          for _ in range(9)]
    with A.Context(P["env"], blocks, 1, 4, A.BOUNDARY_PERIODIC, 4) as c:
        s = A.ElasticShot(c, float(P["dt"]), P["ih"])
        s.set_medium(P["lam"], P["mu"], P["b"])
        s.set_receivers(P["rec"], 1, P["nsteps"])
        s.reset()
        for k in range(9):
            s.set_state(k, x0[k])
        J, _, adj0 = s.gradient(np.zeros((P["nsteps"], len(P["rec"])), np.float32),
                                P["nsteps"], 6, from_current=True, adjoint_state=True)
        s.close()
    lhs = sum(float((x0[k].astype(np.float64) * adj0[k].astype(np.float64)).sum())
              for k in range(9))
    say("periodic dot product", blocks, lhs, 2 * J)
    assert abs(lhs - 2 * J) <= 1e-4 * abs(2 * J)


def test_plain_gradient_keeps_one_frozen_block():
    # the plain per-shot loop stays on one block with frozen faces; the
    # reverse sweep takes the other geometries
    P = E.small_problem(env=(24, 20, 16), nsteps=6)
    obs = _obs(P, False)
    with A.Context(P["env"], (2, 1, 1), 1, 4, A.BOUNDARY_FROZEN, 4) as c:
        with _shot(c, P) as s:
            with pytest.raises(A.B2RError):
                s.gradient(obs, P["nsteps"], 3, plain=True)
            s.gradient(obs, P["nsteps"], 3)
    with A.Context(P["env"], (1, 1, 1), 2, 4, A.BOUNDARY_FROZEN, 4) as c:
        with _shot(c, P) as s:
            with pytest.raises(A.B2RError):
                s.gradient(obs, P["nsteps"], 3)
