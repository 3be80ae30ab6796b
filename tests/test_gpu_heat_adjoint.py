"""GPU heat adjoint (b2r_heat_gradient) against the FP64 oracle adjoint, the
dot-product identity, checkpoint and decomposition invariance, and the error
paths."""
import sys

import numpy as np
import pytest

import _oracle as O
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


def _inputs(env, seed):
    rng = np.random.default_rng(seed)
    gx, gy, gz = env
    u0 = rng.standard_normal((gz, gy, gx)).astype(np.float32)
    tgt = rng.standard_normal((gz, gy, gx)).astype(np.float32)
    return u0, tgt


def _gpu_grad(env, blocks, R, D, boundary, u0, tgt, p, nsteps, ckpt):
    with A.Context(env, blocks, R, D, boundary, 4) as c:
        fu, ft, fg = c.field(), c.field(), c.field()
        c.upload(fu, u0)
        c.upload(ft, tgt)
        J, gk = c.heat_gradient(fu, ft, p, nsteps, ckpt, fg)
        g = c.download(fg)
        # inputs are left unchanged
        assert np.array_equal(c.download(fu), u0)
    return J, g, gk


def _oracle_grad(u0, tgt, D, k, nsteps, periodic):
    u0d, tgtd = u0.astype(np.float64), tgt.astype(np.float64)
    if not periodic:
        u0d, tgtd = O.frame(u0d, D), O.frame(tgtd, D)
    J, g, gk = O.heat_grad_f64(u0d, tgtd, D, k, nsteps, periodic)
    if not periodic:
        g = O.unframe(g, D)
    return J, g, gk


@pytest.mark.parametrize("periodic", [0, 1])
@pytest.mark.parametrize("D", [1, 2, 3, 4])
def test_heat_gradient_vs_fp64_adjoint(D, periodic):
    env, nsteps = (26, 20, 18), 24
    k = (0.05 / D, 0.04 / D, 0.03 / D)
    u0, tgt = _inputs(env, D + 10 * periodic)
    p = A.heat_params(D, A.HEAT_F32, k)
    bnd = A.BOUNDARY_PERIODIC if periodic else A.BOUNDARY_FROZEN
    say("heat adjoint", D, periodic)
    J, g, gk = _gpu_grad(env, (1, 1, 1), 1, D, bnd, u0, tgt, p, nsteps, 7)
    J64, g64, gk64 = _oracle_grad(u0, tgt, D, np.array(k, np.float64).astype(np.float32)
                                  .astype(np.float64), nsteps, periodic)
    e = rel_l2(g, g64)
    ek = np.abs(gk - gk64).max() / np.abs(gk64).max()
    say("  J", J, J64, "rel-L2 g_u0", e, "g_k", ek)
    assert abs(J - J64) <= 1e-5 * J64
    assert e <= 1e-5
    assert ek <= 1e-4


def test_heat_gradient_ref_mode_d1():
    # REF mode: the same adjoint, k_a = adt / h2_a
    env, nsteps = (20, 18, 22), 16
    adt, h2 = 0.01, (1.0, 1.0, 0.25)
    u0, tgt = _inputs(env, 7)
    p = A.heat_params(1, A.HEAT_REF, adt=adt, h2=h2)
    J, g, gk = _gpu_grad(env, (1, 1, 1), 1, 1, A.BOUNDARY_FROZEN, u0, tgt, p, nsteps, 5)
    k = np.array([adt / h for h in h2])
    J64, g64, gk64 = _oracle_grad(u0, tgt, 1, k, nsteps, 0)
    assert abs(J - J64) <= 1e-5 * J64
    assert rel_l2(g, g64) <= 1e-5
    assert np.abs(gk - gk64).max() <= 1e-4 * np.abs(gk64).max()


def test_heat_gradient_checkpoint_and_decomposition_invariance():
    env, D, nsteps = (32, 24, 16), 2, 21
    u0, tgt = _inputs(env, 3)
    p = A.heat_params(D, A.HEAT_F32, (0.02, 0.03, 0.025))
    for bnd in (A.BOUNDARY_FROZEN, A.BOUNDARY_PERIODIC):
        base = _gpu_grad(env, (1, 1, 1), 1, D, bnd, u0, tgt, p, nsteps, 21)
        for blocks, R, ckpt in (((1, 1, 1), 1, 1), ((1, 1, 1), 3, 4), ((2, 2, 1), 2, 5),
                                ((2, 1, 2), 1, 8)):
            say("invariance", bnd, blocks, R, ckpt)
            J, g, gk = _gpu_grad(env, blocks, R, D, bnd, u0, tgt, p, nsteps, ckpt)
            # owned-cell states are bitwise independent of blocking, R and ckpt
            assert np.array_equal(g.view(np.uint32), base[1].view(np.uint32))
            if blocks == (1, 1, 1):
                assert J == base[0] and np.array_equal(gk, base[2])
            else:   # FP64 partial sums are taken over other tiles
                assert abs(J - base[0]) <= 1e-12 * base[0]
                assert np.abs(gk - base[2]).max() <= 1e-9 * np.abs(base[2]).max()


def test_heat_gradient_dot_product():
    # zero frame, linear forward: with target = A^N x - y the adjoint state at
    # the end is y and dJ/du0 = (A^T)^N y, so <A^N x, y> = <x, dJ/du0>
    env, D, nsteps = (24, 22, 20), 3, 30
    x, y = _inputs(env, 9)
    p = A.heat_params(D, A.HEAT_F32, (0.01, 0.02, 0.015))
    with A.Context(env, (2, 1, 1), 2, D, A.BOUNDARY_FROZEN, 4) as c:
        fa, fb, ft, fg = c.field(), c.field(), c.field(), c.field()
        c.upload(fa, x)
        res, _ = c.heat_run(fa, fb, p, nsteps)
        Ax = c.download(res)
        c.upload(ft, (Ax - y).astype(np.float32))
        c.upload(fa, x)
        _, _ = c.heat_gradient(fa, ft, p, nsteps, 8, fg)
        g = c.download(fg)
    ytrue = Ax.astype(np.float64) - (Ax - y).astype(np.float32).astype(np.float64)
    lhs = float((Ax.astype(np.float64) * ytrue).sum())
    rhs = float((x.astype(np.float64) * g.astype(np.float64)).sum())
    say("dot", lhs, rhs)
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def test_heat_gradient_errors():
    env = (16, 16, 16)
    p = A.heat_params(1)
    with A.Context(env, (1, 1, 1), 1, 1, A.BOUNDARY_FROZEN, 4) as c:
        fu, ft, fg = c.field(), c.field(), c.field()
        with pytest.raises(A.B2RError):
            c.heat_gradient(fu, ft, p, 0, 1, fg)            # nsteps
        with pytest.raises(A.B2RError):
            c.heat_gradient(fu, ft, p, 4, 0, fg)            # ckpt
        with pytest.raises(A.B2RError):
            c.heat_gradient(fu, ft, A.heat_params(2), 4, 2, fg)   # D != radius
        c.set_frozen_faces(0)                                # open faces
        with pytest.raises(A.B2RError):
            c.heat_gradient(fu, ft, p, 4, 2, fg)


def test_heat_autograd_wrapper():
    # torch wrapper: gradients w.r.t. u0 and k against FP64 central differences
    import torch
    from b2r_amd.autograd import heat_misfit
    env, D, nsteps = (18, 16, 14), 2, 12
    u0, tgt = _inputs(env, 21)
    k0 = np.array([0.02, 0.025, 0.03])
    with A.Context(env, (1, 1, 1), 1, D, A.BOUNDARY_FROZEN, 4) as c:
        ft = c.field()
        c.upload(ft, tgt)
        u = torch.tensor(u0, dtype=torch.float64, requires_grad=True)
        k = torch.tensor(k0.astype(np.float32).astype(np.float64), requires_grad=True)
        J = heat_misfit(u, k, c, ft, D, nsteps, 5)
        J.backward()

    def J64(uu, kk):
        v = O.heat_f64(O.frame(uu.astype(np.float64), D), D, kk, nsteps, 0)
        r = O.unframe(v, D) - tgt.astype(np.float64)
        return 0.5 * float((r * r).sum())

    kd = k.detach().numpy()
    eps = 1e-6
    for a in range(3):
        kp, km = kd.copy(), kd.copy()
        kp[a] += eps
        km[a] -= eps
        fd = (J64(u0, kp) - J64(u0, km)) / (2 * eps)
        assert abs(fd - float(k.grad[a])) <= 1e-4 * abs(fd), (a, fd, float(k.grad[a]))
    d = np.random.default_rng(2).standard_normal(u0.shape)
    fd = (J64(u0 + eps * d, kd) - J64(u0 - eps * d, kd)) / (2 * eps)
    an = float((u.grad.numpy() * d).sum())
    say("autograd heat", fd, an)
    assert abs(fd - an) <= 1e-4 * abs(fd)
