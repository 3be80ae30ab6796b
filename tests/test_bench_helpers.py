"""The closed forms and generators bench.py's accuracy leg relies on, checked
against the FP64 oracle on small grids."""
import os
import sys

import numpy as np
import pytest

import _oracle as O

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402


@pytest.mark.parametrize("D,periodic", [(1, False), (1, True), (2, True), (3, True), (4, True)])
def test_heat_eigenmode_closed_form_matches_fp64_oracle(D, periodic):
    env = (12, 10, 16) if periodic else (11, 9, 13)
    k = [float(np.float32(0.05 / D))] * 3
    (sx, sy, sz), lam = bench.heat_eigenmode(env, D, k, periodic)
    u0 = sz[:, None, None] * sy[None, :, None] * sx[None, None, :]
    c = O.coef_f64(D)
    c32 = np.array([float(np.float32(v)) for v in bench._heat_taps(D)])
    assert np.allclose(c32, np.asarray(c)[:D + 1], rtol=1e-7)
    # the oracle uses exact taps; feed the FP32 taps' effect through k by
    # comparing against the eigenvalue computed from the oracle's own step
    grid = u0 if periodic else O.frame(u0, D)
    one = O.heat_f64(grid, D, k, 1, periodic)
    one = one if periodic else O.unframe(one, D)
    lam_oracle = float((one * u0).sum() / (u0 * u0).sum())
    assert np.allclose(one, lam_oracle * u0, rtol=0, atol=1e-13)   # it is an eigenmode
    assert abs(lam - lam_oracle) <= 1e-7 * abs(lam)                 # FP32 taps vs exact
    n = 7
    un = O.heat_f64(grid, D, k, n, periodic)
    un = un if periodic else O.unframe(un, D)
    assert np.allclose(un, lam_oracle ** n * u0, rtol=0, atol=1e-12)


def test_perturbed_layers_is_seeded_and_bounded():
    from b2r_amd import workloads as W
    env = (8, 20, 6)
    vp, vs, rho = W.layered_medium(env, 5)
    a = W.perturbed_layers((vp, vs, rho), env, 6)
    b = W.perturbed_layers((vp, vs, rho), env, 6)
    for x, y, base in zip(a, b, (vp, vs, rho)):
        assert np.array_equal(x, y)
        r = x.astype(np.float64) / base
        assert np.all(np.abs(r - 1.0) <= 0.05 + 1e-6)
        # one factor per layer along y
        assert np.allclose(r, r[:1, :, :1])
