"""Pins the FP64 heat adjoint of the oracle (no reference counterpart) by
mathematics the adjoint must satisfy: central finite differences of the FP64
forward (w.r.t. the diffusion numbers and along random directions of the
initial field, with a nonzero frozen frame) and the dot-product identity
<A^N x, y> = <x, (A^T)^N y>, for all four stencil radii, frozen and periodic."""
import numpy as np
import pytest

import _oracle as O


def _case(D, periodic, shape=(9, 8, 7), seed=0, frame_value=0.0):
    rng = np.random.default_rng(seed + 10 * D + periodic)
    gz, gy, gx = shape
    u0 = rng.standard_normal(shape)
    tgt = rng.standard_normal(shape)
    if not periodic:
        u0 = O.frame(u0, D, frame_value)
        tgt = O.frame(tgt, D, 0.0)
    k = np.array([0.05, 0.04, 0.03]) / D
    return u0, tgt, k


def _J(u0, tgt, D, k, n, periodic):
    u = O.heat_f64(u0, D, k, n, periodic)
    r = u - tgt
    if not periodic:
        r = O.unframe(r, D)
    return 0.5 * float((r * r).sum())


@pytest.mark.parametrize("periodic", [0, 1])
@pytest.mark.parametrize("D", [1, 2, 3, 4])
def test_heat_adjoint_matches_finite_differences(D, periodic):
    n = 6
    u0, tgt, k = _case(D, periodic, frame_value=0.7)
    J, g, gk = O.heat_grad_f64(u0, tgt, D, k, n, periodic)
    assert abs(J - _J(u0, tgt, D, k, n, periodic)) <= 1e-12 * J
    eps = 1e-6
    for a in range(3):
        kp, km = k.copy(), k.copy()
        kp[a] += eps
        km[a] -= eps
        fd = (_J(u0, tgt, D, kp, n, periodic) - _J(u0, tgt, D, km, n, periodic)) / (2 * eps)
        assert abs(fd - gk[a]) <= 1e-6 * max(1.0, abs(fd)), (a, fd, gk[a])
    rng = np.random.default_rng(5)
    d = rng.standard_normal(u0.shape)
    if not periodic:
        d = O.frame(O.unframe(d, D), D, 0.0)     # the frame is not a variable
    fd = (_J(u0 + eps * d, tgt, D, k, n, periodic) - _J(u0 - eps * d, tgt, D, k, n, periodic)) / (2 * eps)
    an = float((g * d).sum())
    assert abs(fd - an) <= 1e-6 * max(1.0, abs(fd)), (fd, an)
    if not periodic:
        assert not np.any(g - O.frame(O.unframe(g, D), D, 0.0))


@pytest.mark.parametrize("periodic", [0, 1])
@pytest.mark.parametrize("D", [1, 4])
def test_heat_adjoint_dot_product(D, periodic):
    # zero frame: u_N = A^N u0 is linear; with tgt = A^N x - y the adjoint
    # state at the end is y, so dJ/du0 = (A^T)^N y
    n = 9
    x, y, k = _case(D, periodic, seed=3)
    Ax = O.heat_f64(x, D, k, n, periodic)
    _, g, _ = O.heat_grad_f64(x, Ax - y, D, k, n, periodic)
    if not periodic:
        Ax, x, y, g = (O.unframe(a, D) for a in (Ax, x, y, g))
    lhs = float((Ax * y).sum())
    rhs = float((x * g).sum())
    assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs)), (lhs, rhs)


def test_heat_adjoint_planted_transpose_error_is_caught():
    # the dot-product check above fails for an operator that is not the
    # transpose: one extra forward step on the adjoint side breaks it
    D, periodic, n = 2, 0, 5
    x, y, k = _case(D, periodic, seed=4)
    Ax = O.heat_f64(x, D, k, n, periodic)
    _, g, _ = O.heat_grad_f64(x, Ax - y, D, k, n + 1, periodic)    # wrong N
    Ax1 = O.heat_f64(x, D, k, n + 1, periodic)
    _, g_ok, _ = O.heat_grad_f64(x, Ax1 - y, D, k, n + 1, periodic)
    lhs = float((O.unframe(Ax1, D) * O.unframe(y, D)).sum())
    assert abs(lhs - float((O.unframe(x, D) * O.unframe(g_ok, D)).sum())) <= 1e-12 * abs(lhs)
    bad = float((O.unframe(x, D) * O.unframe(g, D)).sum())
    assert abs(lhs - bad) > 1e-6 * abs(lhs)
