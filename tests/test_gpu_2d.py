"""The 2D environment (_ENV_3D_ = 0, b2r_config.h:13; env_z = 1) on the GPU:
the z check and the Z exchange phase are skipped (b2r.cpp:53-67,
b2r_halo_sync.cu:170, 419, 489), the one plane keeps a zero z frame, and heat
(k_z = 0), Game of Life and the elastic step run as their 3D kernels; each
against the oracle bit for bit, periodic grids wrapping x and y only."""
import sys

import numpy as np
import pytest

import _elastic as E
import _oracle as O
from _util import block_view, expected_halo_ids, expected_layout, id_field

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
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    if a.dtype == np.float32:
        return np.array_equal(a.view(np.uint32), b.view(np.uint32))
    return np.array_equal(a, b)


EXCH2D = [
    ((24, 16, 1), (2, 2, 1), 1, 2, 4, False),
    ((24, 16, 1), (3, 2, 1), 2, 1, 4, True),
    ((64, 32, 1), (2, 2, 1), 2, 1, 1, True),
    ((40, 24, 1), (2, 3, 1), 1, 3, 8, False),
]


@pytest.mark.parametrize("env,blocks,R,radius,esize,periodic", EXCH2D)
def test_2d_exchange_index_map(env, blocks, R, radius, esize, periodic):
    say("2d exchange", env, blocks, R, radius, esize, periodic)
    pm = 3 if periodic else 0
    with A.Context(env, blocks, R, radius, int(periodic), esize) as c:
        lay = c.layout
        assert list(lay.S) == expected_layout(env, blocks, R, radius, esize)["S"]
        raw = id_field(lay, blocks, np.float64, -1)
        if esize == 1:
            raw = np.where(raw == -1, 255, raw % 251).astype(np.uint8)
        else:
            raw = raw.astype({4: np.float32, 8: np.float64}[esize])
        want = O.exchange(raw, lay, blocks, esize, pm)
        ids = expected_halo_ids(lay, blocks, pm, -1)
        for b, w in enumerate(ids):  # the oracle itself against the index map
            got = block_view(want, lay, b).astype(np.float64)
            if esize == 1:
                w = np.where(w == -1, 255, w % 251)
            assert np.array_equal(got, w), b
        for direct in (False, True):
            f = c.field()
            c.raw_upload(f, raw)
            if direct:
                c.boundary_updating_direct([f], 0)
            else:
                c.boundary_updating(f, 0)
            assert np.array_equal(c.raw(f), want), direct
            c.free(f)


@pytest.mark.parametrize("fused", [0, 1])
@pytest.mark.parametrize("env,blocks,R,D,periodic", [
    ((40, 32, 1), (2, 2, 1), 2, 1, False), ((36, 30, 1), (1, 3, 1), 1, 2, True),
    ((48, 40, 1), (2, 1, 1), 1, 4, False), ((40, 40, 1), (2, 2, 1), 2, 2, True)])
def test_2d_heat_matches_oracle(env, blocks, R, D, periodic, fused):
    say("2d heat", env, blocks, R, D, periodic, fused)
    k = (0.031, 0.027, 0.0)
    p = A.heat_params(D, A.HEAT_F32, k)
    with A.Context(env, blocks, R, D, int(periodic), 4) as c:
        a, b = c.field(), c.field()
        c.init_uniform(a, 5)
        res, _ = c.heat_run(a, b, p, 7, fused=fused)
        got = c.download(res)
    u = O.uniform((1, env[1], env[0]), 5)
    # oracle grid: x and y wrap or carry a D frame, z one plane in a zero frame
    pm = 3 if periodic else 0
    g = np.pad(u, ((D, D), (0, 0), (0, 0)) if periodic else D)
    want = O.heat_f32(g, D, A.HEAT_F32, k, 0.0, (1, 1, 1), 7, pm)
    want = want[D:-D] if periodic else O.unframe(want, D)
    assert bits_equal(got, want)


def test_2d_heat_rejects_a_z_term():
    with A.Context((32, 32, 1), (1, 1, 1), 1, 1, 0, 4) as c:
        a, b = c.field(), c.field()
        for p in (A.heat_params(1, A.HEAT_F32, (0.1, 0.1, 0.1)), A.heat_params(1, A.HEAT_REF)):
            with pytest.raises(A.B2RError) as e:
                c.heat_run(a, b, p, 2)
            assert e.value.code == A.B2R_E_UNSUPPORTED


@pytest.mark.parametrize("fused", [0, 1])
@pytest.mark.parametrize("env,blocks,R,periodic", [
    ((64, 48, 1), (2, 2, 1), 2, True), ((64, 40, 1), (2, 1, 1), 4, False),
    ((50, 36, 1), (1, 2, 1), 1, True), ((128, 64, 1), (2, 2, 1), 3, True)])
def test_2d_gol_matches_oracle(env, blocks, R, periodic, fused):
    say("2d gol", env, blocks, R, periodic, fused)
    with A.Context(env, blocks, R, 1, int(periodic), 1) as c:
        a, b = c.field(), c.field()
        c.init_bernoulli(a, 3, int(0.4 * 16777216))
        res, _ = c.gol_run(a, b, 9, fused=fused)
        got = c.download(res)
    u = O.bernoulli((1, env[1], env[0]), 3, 0.4)
    pm = 3 if periodic else 0
    g = np.pad(u, ((1, 1), (0, 0), (0, 0))) if periodic else O.frame(u, 1)
    want = O.gol(g, 9, pm)
    want = want[1:-1] if periodic else want[1:-1, 1:-1, 1:-1]
    assert np.any(want) and np.array_equal(got, want)


@pytest.mark.parametrize("tile", [13, 14])
@pytest.mark.parametrize("R", [2, 4])
@pytest.mark.parametrize("blocks", [(2, 2, 1), (1, 1, 1)])
def test_2d_gol_whole_rows(blocks, R, tile):
    # the whole-row tiles on a 2D grid periodic in x over 512 cells: the z
    # faces stay frozen, so every generation clips its update box along z
    env = (512, 48, 1)
    say("2d gol whole rows", blocks, R, tile)
    with A.Context(env, blocks, R, 1, 1, 1) as c:
        c.set_tuning(gol_tile=tile, gol_zc=0, gol_bytes=0)
        a, b = c.field(), c.field()
        c.init_bernoulli(a, 5, int(0.4 * 16777216))
        res, _ = c.gol_run(a, b, 2 * R + 1, fused=1)
        got = c.download(res)
    u = O.bernoulli((1, env[1], env[0]), 5, 0.4)
    want = O.gol(np.pad(u, ((1, 1), (0, 0), (0, 0))), 2 * R + 1, 3)[1:-1]
    assert np.any(want) and np.array_equal(got, want)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_2d_elastic_forward_matches_oracle(mode):
    # one plane with a zero z frame: every z derivative is zero there, the
    # plane-strain limit; the FP32 oracle with gz = 1 (its 2-cell zero frame)
    env = (40, 36, 1)
    say("2d elastic", env, mode)
    P = E.small_problem(env=(40, 36, 1), nsteps=25)
    src = np.array([[20, 12, 0]], np.int32)
    rec = np.array([[x, 3, 0] for x in range(4, 36, 5)], np.int32)
    with A.Context(env, (2, 1, 1), 1, 4, A.BOUNDARY_FROZEN, 4) as c:
        s = A.ElasticShot(c, float(P["dt"]), P["ih"])
        s.set_medium(P["lam"], P["mu"], P["b"])
        s.set_sources(src, [1], P["wav"])
        s.set_receivers(rec, 1, P["nsteps"])
        s.forward(P["nsteps"], mode)
        got = [s.state(k) for k in range(9)]
        grec = s.receivers()
        s.close()
    st, r = E.forward(env, P["lam"], P["mu"], P["b"], P["dt"], P["ih"], P["nsteps"], src, [1],
                      P["wav"], rec, 1)
    assert np.any(st[1]) and not np.any(st[2])  # vz never driven in plane strain
    for k in range(9):
        assert bits_equal(got[k], st[k]), k
    assert bits_equal(grec, r)


def test_2d_geometry_checks():
    g = A.Geom(32, 32, 1, 1, 1, 2, 1, 1, 0, 4)
    import ctypes as C
    h = C.c_void_p()
    assert A.lib().b2r_create(C.byref(g), 0, C.byref(h)) == A.B2R_E_GEOM
    # the z check is skipped in 2D: R*radius = 8 <= min(Bx, By)/2 = 8
    with A.Context((16, 16, 1), (1, 1, 1), 2, 4, 0, 4) as c:
        assert c.layout.S[2] == 1 + 16
        assert c.sync_flag(0)[4:] == [0, 0]
