"""Pins the CPU oracle (no GPU): against the reference's own CPU heat path
compiled from its sources, the golden fixtures made from it, closed forms,
invariants and independent brute-force checks.  A plausible mistake in the
oracle (a wrong tap, sign, index or loop bound) fails at least one of these."""
import math
import os

from fractions import Fraction

import numpy as np
import pytest

import _oracle as O
from _util import Lay, block_view, expected_halo_ids, expected_layout, id_field, rel_l2

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ------------------------------------------------------------ reference heat
@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference absent)")
@pytest.mark.parametrize("shape,b2r_i", [((9, 10, 12), 0), ((11, 13, 17), 1), ((8, 8, 8), 2)])
def test_ref1_step_bitwise_vs_reference_heat3d(shape, b2r_i):
    h1 = O.uniform(shape, seed=7) * 3 - 1
    got = O.ref1_block_step(h1, b2r_i)
    want = O.reference_block_step(h1, b2r_i)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference absent)")
def test_ref_mode_global_equals_reference_r1_loop():
    # R=1 with a zero frame: the reference's region [1, N-2] is the owned grid,
    # so its step loop and the oracle's frozen-frame REF mode coincide.
    owned = O.uniform((9, 10, 12), seed=3)
    g = O.frame(owned, 1)
    ref = g.copy()
    for _ in range(6):
        ref = O.reference_block_step(ref, 0)
    got = O.heat_f32(g, 1, 1, (0, 0, 0), 0.1 * 0.1, (1.0, 1.0, 0.25), 6, periodic=False)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_golden_fixtures_reproduced():
    files = sorted(f for f in os.listdir(GOLDEN) if f.startswith("heat_delta1_ref") and f.endswith(".npz"))
    assert files, "golden fixtures missing"
    for f in files:
        z = np.load(os.path.join(GOLDEN, f))
        out = z["input"].copy()
        for i in range(int(z["steps"])):
            out = O.ref1_block_step(out, 0)
        assert np.array_equal(out.view(np.uint32), z["output"].view(np.uint32)), f


# ------------------------------------------------------------ heat closed forms
def test_dirichlet_sine_eigenmode_delta1():
    N = (11, 9, 13)  # x, y, z owned
    m = (2, 3, 1)
    k = (0.07, 0.05, 0.11)
    ix = [np.arange(1, n + 1) for n in N]
    sx, sy, sz = (np.sin(math.pi * m[a] * ix[a] / (N[a] + 1)) for a in range(3))
    u0 = sz[:, None, None] * sy[None, :, None] * sx[None, None, :]
    lam = 1 + sum(k[a] * (2 * math.cos(math.pi * m[a] / (N[a] + 1)) - 2) for a in range(3))
    n = 17
    got = O.unframe(O.heat_f64(O.frame(u0, 1), 1, k, n, periodic=False), 1)
    np.testing.assert_allclose(got, lam ** n * u0, rtol=0, atol=1e-13)


@pytest.mark.parametrize("D", [1, 2, 3, 4])
def test_periodic_fourier_mode(D):
    N = (16, 12, 20)
    m = (3, 1, 2)
    k = (0.03, 0.04, 0.02)
    c = O.coef_f64(D)
    th = [2 * math.pi * m[a] / N[a] for a in range(3)]
    z, y, x = np.meshgrid(np.arange(N[2]), np.arange(N[1]), np.arange(N[0]), indexing="ij")
    u0 = np.cos(th[0] * x + th[1] * y + th[2] * z + 0.3)
    sym = [c[0] + 2 * sum(c[j] * math.cos(j * th[a]) for j in range(1, D + 1)) for a in range(3)]
    lam = 1 + sum(k[a] * sym[a] for a in range(3))
    n = 9
    got = O.heat_f64(u0, D, k, n, periodic=True)
    np.testing.assert_allclose(got, lam ** n * u0, rtol=0, atol=1e-12)


@pytest.mark.parametrize("D", [1, 2, 3, 4])
def test_taps_match_taylor_weights(D):
    # the 2D+1 taps must be the unique order-2D central second-difference weights
    c = O.coef_f64(D)
    j = np.arange(1, D + 1)
    # sum over taps of c_j * j^(2p) for p=0..D: 2*sum c_j j^0 + c0 = 0, 2*sum c_j j^2 = 2,
    # and 2*sum c_j j^(2p) = 0 for p = 2..D
    assert abs(c[0] + 2 * c[1:D + 1].sum()) < 1e-14
    assert abs(2 * (c[1:D + 1] * j ** 2).sum() - 2.0) < 1e-13
    for p in range(2, D + 1):
        assert abs(2 * (c[1:D + 1] * j ** (2 * p)).sum()) < 1e-10


@pytest.mark.parametrize("D", [1, 2, 3, 4])
def test_polynomial_exactness(D):
    # one step on u = x^(2D+1) + y^(2D): the stencil is exact up to degree 2D+1
    N = 3 * D + 6
    k = (0.01, 0.02, 0.03)
    xs = np.arange(-D, N + D, dtype=np.float64) / N
    z, y, x = np.meshgrid(xs, xs, xs, indexing="ij")
    u = x ** (2 * D + 1) + y ** (2 * D)
    got = O.heat_f64(u, D, k, 1, periodic=False)
    h = 1.0 / N
    p1, p2 = 2 * D + 1, 2 * D
    lap_x = p1 * (p1 - 1) * x ** (p1 - 2) * h * h
    lap_y = p2 * (p2 - 1) * y ** (p2 - 2) * h * h
    want = u + k[0] * lap_x + k[1] * lap_y
    s = slice(D, -D)
    np.testing.assert_allclose(got[s, s, s], want[s, s, s], rtol=0, atol=1e-12)


@pytest.mark.parametrize("D", [1, 2, 3, 4])
def test_periodic_conservation(D):
    u = O.uniform((10, 12, 14), seed=5).astype(np.float64)
    got = O.heat_f64(u, D, (0.02, 0.03, 0.01), 11, periodic=True)
    assert abs(got.sum() - u.sum()) < 1e-9


def test_max_principle_delta1():
    u = O.frame(O.uniform((12, 12, 12), seed=9).astype(np.float64), 1)
    got = O.heat_f64(u, 1, (0.05, 0.05, 0.05), 20, periodic=False)
    assert got.min() >= 0.0 and got.max() <= u.max()


@pytest.mark.parametrize("D", [1, 2, 3, 4])
@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("periodic", [False, True])
def test_f32_order_close_to_f64(D, mode, periodic):
    owned = O.uniform((10, 11, 13), seed=D)
    g = owned if periodic else O.frame(owned, D)
    k = (0.02, 0.03, 0.025)
    h2 = (1.0, 0.8, 1.25)
    adt = 0.05
    n = 12
    f32 = O.heat_f32(g, D, mode, k, adt, h2, n, periodic)
    kk = k if mode == 0 else tuple(adt / v for v in h2)
    f64 = O.heat_f64(g.astype(np.float64), D, kk, n, periodic)
    assert rel_l2(f32, f64) < 2e-6


def _round_f32(q):
    """The float32 nearest to the rational q, ties to even (normal range)."""
    if q == 0:
        return np.float32(0.0)
    sign = -1 if q < 0 else 1
    q = abs(q)
    e = q.numerator.bit_length() - q.denominator.bit_length()
    if Fraction(2) ** e > q:
        e -= 1  # 2**e <= q < 2**(e+1)
    m = q / Fraction(2) ** (e - 23)  # in [2**23, 2**24)
    n, r = divmod(m.numerator, m.denominator)
    if 2 * r > m.denominator or (2 * r == m.denominator and n & 1):
        n += 1
    return np.float32(sign * n * 2.0 ** (e - 23))


def _fmaf(a, b, c):
    """a*b + c rounded once to float32, in exact rational arithmetic."""
    return _round_f32(Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c)))


def test_exact_fmaf_helper():
    # one rounding, not two: 24929 * 673 = 2^24 + 1, a midpoint between float32
    # neighbours 2^24 and 2^24 + 2; adding 2^-30 puts the exact sum just above
    # it (rounds up), while rounding to double first lands on the midpoint and
    # ties to even (down)
    a, b, c = np.float32(24929), np.float32(673), np.float32(2.0 ** -30)
    assert _fmaf(a, b, c) == np.float32(2 ** 24 + 2)
    assert np.float32(float(a) * float(b) + float(c)) == np.float32(2 ** 24)
    assert _fmaf(a, b, -c) == np.float32(2 ** 24)
    assert _fmaf(np.float32(3), np.float32(5), np.float32(-15)) == 0
    assert _fmaf(np.float32(-1.5), np.float32(2), np.float32(0.25)) == np.float32(-2.75)


def test_f32_mode_matches_exact_one_step():
    # an independent restatement of the F32 evaluation order: float32 adds and
    # products in numpy, each fused multiply-add in exact rational arithmetic
    D = 3
    owned = O.uniform((6, 7, 5), seed=11)
    g = O.frame(owned, D)
    k = np.float32([0.021, 0.033, 0.017])
    got = O.heat_f32(g, D, 0, k, 0.0, (1, 1, 1), 1, False)
    c = O.coef_f32(D)
    u = g
    want = u.copy()
    nz, ny, nx = u.shape
    for z in range(D, nz - D):
        for y in range(D, ny - D):
            for x in range(D, nx - D):
                def pair(a, j):
                    if a == 0:
                        return np.float32(u[z, y, x - j] + u[z, y, x + j])
                    if a == 1:
                        return np.float32(u[z, y - j, x] + u[z, y + j, x])
                    return np.float32(u[z - j, y, x] + u[z + j, y, x])

                def lap(a):
                    acc = np.float32(c[D] * pair(a, D))
                    for j in range(D - 1, 0, -1):
                        acc = _fmaf(c[j], pair(a, j), acc)
                    return _fmaf(c[0], u[z, y, x], acc)
                t = np.float32(k[0] * lap(0))
                t = _fmaf(k[1], lap(1), t)
                t = _fmaf(k[2], lap(2), t)
                want[z, y, x] = np.float32(u[z, y, x] + t)
    s = slice(D, -D)
    assert np.array_equal(got[s, s, s].view(np.uint32), want[s, s, s].view(np.uint32))


# ------------------------------------------------------------ Game of Life
def _count_np(u):
    n = np.zeros(u.shape, np.int32)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                if dx or dy or dz:
                    n += np.roll(u, (dz, dy, dx), axis=(0, 1, 2))
    return n


def _gol_np(u):
    n = _count_np(u.astype(np.int32))
    out = u.copy()
    out[(u != 0) & (n <= 1)] = 0
    out[(u == 0) & (n == 5)] = 1
    out[(u != 0) & (n >= 8)] = 0
    return out


def test_gol_matches_numpy_rule_periodic():
    u = O.bernoulli((10, 12, 14), seed=1, p=0.3)
    want = u
    for _ in range(5):
        want = _gol_np(want)
    got = O.gol(u, 5, periodic=True)
    assert np.array_equal(got, want)


def test_gol_still_life_cube_and_lone_cell():
    g = np.zeros((8, 8, 8), np.uint8)
    g[3:5, 3:5, 3:5] = 1           # 2x2x2 cube: each cell has 7 neighbours
    g[1, 1, 6] = 1                 # lone cell: dies
    out = O.gol(O.frame(g, 1), 3, periodic=False)
    want = np.zeros_like(g)
    want[3:5, 3:5, 3:5] = 1
    assert np.array_equal(O.unframe(out, 1), want)


def test_gol_empty_and_full():
    z = np.zeros((6, 7, 8), np.uint8)
    assert not O.gol(z, 4, periodic=True).any()
    f = np.ones((6, 7, 8), np.uint8)  # 26 neighbours each -> overcrowded
    assert not O.gol(f, 1, periodic=True).any()


def test_gol_birth_at_five():
    g = np.zeros((7, 7, 7), np.uint8)
    g[2, 2, 2] = g[2, 2, 3] = g[2, 3, 2] = g[3, 2, 2] = g[4, 4, 4] = 1  # (3,3,3) sees 5
    out = O.gol(g, 1, periodic=True)
    assert out[3, 3, 3] == 1


def test_gol_translation_equivariance():
    u = O.bernoulli((12, 10, 16), seed=4, p=0.3)
    a = O.gol(np.roll(u, (3, -2, 5), axis=(0, 1, 2)), 6, periodic=True)
    b = np.roll(O.gol(u, 6, periodic=True), (3, -2, 5), axis=(0, 1, 2))
    assert np.array_equal(a, b)


# ------------------------------------------------------------ generators
def test_splitmix64_known_answers():
    # SplitMix64 of state 0 after one increment: the published first output
    assert O.lib.orc_splitmix64(0) == 0xE220A8397B1DCDAF
    assert O.lib.orc_splitmix64(0x9E3779B97F4A7C15) == 0x6E789E6AA1B965F4


def test_uniform_bernoulli_moments():
    u = O.uniform((64, 64, 64), seed=0)
    assert u.min() >= 0 and u.max() < 1
    assert abs(u.mean() - 0.5) < 2e-3 and abs(u.var() - 1 / 12) < 2e-3
    b = O.bernoulli((64, 64, 64), seed=1, p=0.3)
    assert set(np.unique(b)) <= {0, 1} and abs(b.mean() - 0.3) < 3e-3


# ------------------------------------------------------------ halo exchange
CASES = [
    ((8, 6, 10), (2, 1, 2), 1, 1, 4, False),
    ((8, 6, 10), (2, 1, 2), 2, 1, 4, True),
    ((12, 12, 12), (3, 2, 1), 1, 2, 8, False),
    ((12, 16, 16), (1, 2, 2), 2, 2, 8, True),
    ((16, 8, 12), (2, 2, 3), 1, 1, 1, True),
    ((16, 8, 12), (2, 1, 1), 3, 1, 1, False),
    ((8, 8, 8), (1, 1, 1), 2, 2, 4, True),
]


@pytest.mark.parametrize("env,blocks,R,radius,esize,periodic", CASES)
def test_exchange_index_map(env, blocks, R, radius, esize, periodic):
    lay = Lay(expected_layout(env, blocks, R, radius, esize))
    dt = {1: np.uint8, 4: np.float32, 8: np.float64}[esize]
    poison = 255 if esize == 1 else -1
    raw = id_field(lay, blocks, dt if esize > 1 else np.float64, poison)
    if esize == 1:  # ids mod 251 fit a byte; poison 255 stays distinct
        raw = np.where(raw == poison, 255, raw % 251).astype(np.uint8)
    out = O.exchange(raw, lay, blocks, esize, periodic)
    want = expected_halo_ids(lay, blocks, periodic, poison)
    for b, w in enumerate(want):
        got = block_view(out, lay, b).astype(np.float64)
        if esize == 1:
            w = np.where(w == poison, 255, w % 251)
        assert np.array_equal(got, w), f"block {b}"
