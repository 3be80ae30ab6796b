"""GPU parity through the C ABI: layout, generators, halo exchange, heat delta1-4
and Game of Life against the CPU oracle (bit-exact) and the reference's golden
fixtures; R- and decomposition-invariance."""
import os
import sys

import numpy as np
import pytest

import _oracle as O
from _util import block_view, expected_layout, id_field, rel_l2

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
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    if a.dtype == np.float32:
        return np.array_equal(a.view(np.uint32), b.view(np.uint32))
    return np.array_equal(a, b)


# ------------------------------------------------------------------ layout
@pytest.mark.parametrize("env,blocks,R,radius,esize", [
    ((16, 8, 12), (2, 1, 2), 2, 1, 1), ((20, 20, 20), (2, 2, 2), 2, 2, 4),
    ((24, 16, 8), (1, 2, 1), 1, 4, 8), ((64, 32, 32), (2, 2, 2), 4, 1, 1)])
def test_layout_matches_rule(env, blocks, R, radius, esize):
    say("layout", env, blocks, R, radius, esize)
    with A.Context(env, blocks, R, radius, 0, esize) as c:
        lay = c.layout
        want = expected_layout(env, blocks, R, radius, esize)
        assert list(lay.B) == want["B"] and list(lay.S) == want["S"]
        for k in ("halo", "nblocks", "x_offset", "pitch", "slice", "block_stride", "field_elems"):
            assert getattr(lay, k) == want[k], k


# ------------------------------------------------------------------ generators
def test_generators_bitwise():
    env = (24, 10, 18)
    say("generators", env)
    with A.Context(env, (2, 1, 3), 1, 1, 0, 4) as c:
        f = c.field()
        c.init_uniform(f, 0)
        assert bits_equal(c.download(f), O.uniform((18, 10, 24), 0))
        rng = np.random.default_rng(0)
        nb = 3
        amp = rng.uniform(0.5, 2, nb)
        ex, ey, ez = (rng.uniform(0, 1, nb * n) for n in env)
        c.init_blobs(f, amp, ex, ey, ez)
        assert bits_equal(c.download(f), O.blobs(env, amp, ex, ey, ez))
    with A.Context(env, (1, 2, 3), 1, 1, 1, 1) as c:
        f = c.field()
        c.init_bernoulli(f, 1, int(0.3 * 16777216))
        assert np.array_equal(c.download(f), O.bernoulli((18, 10, 24), 1, 0.3))


def test_upload_download_roundtrip_and_errors():
    env = (12, 8, 10)
    with A.Context(env, (2, 2, 1), 1, 1, 0, 4) as c:
        f = c.field()
        u = O.uniform((10, 8, 12), 5)
        c.upload(f, u)
        assert bits_equal(c.download(f), u)
        with pytest.raises(A.B2RError):
            c.heat_step(f, f, A.heat_params(1), 0)          # in == out
        with pytest.raises(A.B2RError):
            c.heat_step(f, c.field(), A.heat_params(2), 0)  # D != radius
        with pytest.raises(A.B2RError):
            c.gol_step(f, c.field(), 0)                     # wrong cell type
        for bad in ({"gol_tile": 16}, {"gol_tile": -1}, {"heat_wy": 5}, {"heat_wy": 3, "heat_wx": 2},
                    {"heat_ry": 3}, {"gol_zc": -1}, {"el_kernel": 3}):
            with pytest.raises(A.B2RError):
                c.set_tuning(**bad)                         # no such launch shape
        c.set_tuning(gol_tile=15, heat_wy=3, heat_wx=3)     # the largest built ones
        assert c.get_tuning()["gol_tile"] == 15


# ------------------------------------------------------------------ exchange
EXCH = [
    ((8, 6, 10), (2, 1, 2), 1, 1, 4, False),
    ((8, 6, 10), (2, 1, 2), 2, 1, 4, True),
    ((12, 12, 12), (3, 2, 1), 1, 2, 8, False),
    ((12, 16, 16), (1, 2, 2), 2, 2, 8, True),
    ((16, 8, 12), (2, 2, 3), 1, 1, 1, True),
    ((16, 8, 12), (2, 1, 1), 3, 1, 1, False),
    ((8, 8, 8), (1, 1, 1), 2, 2, 4, True),
    ((40, 24, 32), (2, 2, 2), 2, 2, 4, True),
]


def _id_raw(lay, blocks, esize):
    dt = {1: np.uint8, 4: np.float32, 8: np.float64}[esize]
    raw = id_field(lay, blocks, np.float64, -1)
    if esize == 1:
        return np.where(raw == -1, 255, raw % 251).astype(np.uint8)
    return raw.astype(dt)


@pytest.mark.parametrize("env,blocks,R,radius,esize,periodic", EXCH)
def test_exchange_matches_oracle(env, blocks, R, radius, esize, periodic):
    say("exchange", env, blocks, R, radius, esize, periodic)
    with A.Context(env, blocks, R, radius, int(periodic), esize) as c:
        lay = c.layout
        raw = _id_raw(lay, blocks, esize)
        want = O.exchange(raw, lay, blocks, esize, periodic)
        f = c.field()
        c.raw_upload(f, raw)
        c.boundary_updating(f, 0)
        assert np.array_equal(c.raw(f), want), "pack/unpack exchange"
        g = c.field()
        c.raw_upload(g, raw)
        c.boundary_updating_direct([g], 0)
        assert np.array_equal(c.raw(g), want), "direct exchange"
        # message contents in the reference's message layout
        for dim in range(3):
            me = c.msg_elems(dim)
            msgs = c.dev_alloc(me * 2 * lay.nblocks * esize)
            c.raw_upload(f, raw)
            c.pack(dim, f, msgs)
            got = c.d2h(msgs, me * 2 * lay.nblocks * esize, raw.dtype)
            assert np.array_equal(got, O.pack(raw, lay, blocks, esize, periodic, dim)), dim
            c.dev_free(msgs)


# ------------------------------------------------------------------ heat
HEAT = [
    # env, blocks, R, D, periodic, nsteps
    ((20, 18, 16), (1, 1, 1), 1, 1, False, 9),
    ((20, 18, 16), (2, 1, 2), 2, 1, False, 9),
    ((24, 16, 20), (2, 2, 1), 3, 1, True, 10),
    ((32, 24, 32), (2, 1, 2), 2, 2, False, 8),
    ((24, 24, 24), (2, 2, 2), 1, 2, True, 7),
    ((36, 24, 30), (2, 1, 1), 2, 3, False, 6),
    ((30, 24, 36), (1, 2, 2), 1, 3, True, 5),
    ((32, 24, 32), (2, 1, 2), 2, 4, False, 6),
    ((40, 32, 48), (1, 1, 2), 1, 4, True, 5),
    ((130, 66, 70), (2, 1, 1), 2, 1, False, 6),
    ((96, 80, 72), (2, 2, 2), 2, 4, False, 4),
    # deep windows, R*D up to 8 (the fused kernel's inner 48 x 16 columns)
    ((40, 40, 40), (1, 1, 1), 6, 1, False, 13),
    ((40, 36, 32), (1, 1, 1), 8, 1, True, 17),
    ((48, 40, 44), (2, 1, 1), 4, 2, True, 9),
    ((54, 40, 48), (1, 2, 2), 3, 2, False, 7),
]


def _heat_case(env, blocks, R, D, periodic, nsteps, mode, k=(0.021, 0.033, 0.017),
               adt=0.1 * 0.1, h2=(1.0, 1.0, 0.25), fused=0):
    p = A.heat_params(D, mode, k, adt, h2)
    with A.Context(env, blocks, R, D, int(periodic), 4) as c:
        a, b = c.field(), c.field()
        c.init_uniform(a, 17)
        res, _ = c.heat_run(a, b, p, nsteps, fused=fused)
        return c.download(res)


@pytest.mark.parametrize("fused", [0, 1])
@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("env,blocks,R,D,periodic,nsteps", HEAT)
def test_heat_bitwise_vs_oracle(env, blocks, R, D, periodic, nsteps, mode, fused):
    say("heat", env, blocks, "R", R, "D", D, "periodic", periodic, "mode", mode, "fused", fused)
    k = (0.021, 0.033, 0.017)
    got = _heat_case(env, blocks, R, D, periodic, nsteps, mode, k, fused=fused)
    owned = O.uniform((env[2], env[1], env[0]), 17)
    g = owned if periodic else O.frame(owned, D)
    want = O.heat_f32(g, D, mode, k, 0.1 * 0.1, (1.0, 1.0, 0.25), nsteps, periodic)
    if not periodic:
        want = O.unframe(want, D)
    assert bits_equal(got, want)
    kk = k if mode == 0 else (0.01, 0.01, 0.04)
    f64 = O.heat_f64(g.astype(np.float64), D, kk, nsteps, periodic)
    if not periodic:
        f64 = O.unframe(f64, D)
    assert rel_l2(got, f64) < 1e-6


@pytest.mark.parametrize("fname", sorted(f for f in os.listdir(os.path.join(os.path.dirname(__file__), "golden"))
                                         if f.startswith("heat_delta1_ref")))
def test_heat_ref_mode_matches_reference_golden(fname):
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", fname))
    inp, out, steps = z["input"], z["output"], int(z["steps"])
    owned = O.unframe(inp, 1)
    env = (owned.shape[2], owned.shape[1], owned.shape[0])
    say("golden", fname, env)
    with A.Context(env, (1, 1, 1), 1, 1, 0, 4) as c:
        a, b = c.field(), c.field()
        c.upload(a, owned)
        res, _ = c.heat_run(a, b, A.heat_params(1, A.HEAT_REF), steps)
        assert bits_equal(c.download(res), O.unframe(out, 1))


def test_heat_R_and_decomposition_invariance():
    env = (96, 64, 80)
    outs = []
    for blocks, R, fused in [((1, 1, 1), 1, 0), ((2, 2, 2), 1, 1), ((2, 1, 2), 3, 0),
                             ((1, 2, 1), 4, 0), ((2, 2, 1), 2, 1)]:
        say("heat invariance", blocks, R, fused)
        outs.append(_heat_case(env, blocks, R, 2, False, 12, 0, fused=fused))
    for o in outs[1:]:
        assert bits_equal(o, outs[0])


def test_heat_step_single_inner_region():
    # one inner step b2r_i=1 with R=2 updates [2D, S-1-2D] only, the rest of
    # the output keeps its previous contents
    env, D, R = (16, 16, 16), 1, 2
    with A.Context(env, (1, 1, 1), R, D, 1, 4) as c:
        a, b = c.field(), c.field()
        c.init_uniform(a, 3)
        c.boundary_updating(a, 0)
        c.fill(b, -7.0)
        c.heat_step(a, b, A.heat_params(D), 1)
        lay = c.layout
        blk_in = block_view(c.raw(a), lay, 0)
        blk = block_view(c.raw(b), lay, 0)
        S = lay.S[0]
        lo, hi = D * 2, S - 1 - D * 2
        inner = blk[lo:hi + 1, lo:hi + 1, lo:hi + 1]
        assert not np.any(inner == -7.0)
        mask = np.ones_like(blk, bool)
        mask[lo:hi + 1, lo:hi + 1, lo:hi + 1] = False
        assert np.all(blk[mask] == -7.0)
        want = O.heat_f32(blk_in, D, 0, (0.1, 0.1, 0.1), 0, (1, 1, 1), 1, periodic=False)
        # the frozen-frame oracle on the padded block equals the interior update
        w = want[lo:hi + 1, lo:hi + 1, lo:hi + 1]
        assert bits_equal(inner, w)


# ------------------------------------------------------------------ Game of Life
GOL = [
    ((64, 40, 48), (2, 2, 2), 1, True, 12),
    ((64, 40, 48), (2, 2, 2), 2, True, 12),
    ((64, 40, 48), (2, 2, 2), 4, True, 12),
    ((64, 40, 48), (2, 2, 2), 4, False, 13),
    ((50, 30, 20), (1, 3, 2), 3, True, 7),
    ((36, 22, 26), (1, 1, 1), 2, False, 9),
    ((100, 36, 36), (4, 1, 2), 1, True, 5),
]


@pytest.mark.parametrize("fused", [0, 1])
@pytest.mark.parametrize("env,blocks,R,periodic,nsteps", GOL)
def test_gol_bitwise_vs_oracle(env, blocks, R, periodic, nsteps, fused):
    say("gol", env, blocks, "R", R, "periodic", periodic, "steps", nsteps, "fused", fused)
    with A.Context(env, blocks, R, 1, int(periodic), 1) as c:
        a, b = c.field(), c.field()
        c.init_bernoulli(a, 1, int(0.3 * 16777216))
        res, _ = c.gol_run(a, b, nsteps, fused=fused)
        got = c.download(res)
    u = O.bernoulli((env[2], env[1], env[0]), 1, 0.3)
    g = u if periodic else O.frame(u, 1)
    want = O.gol(g, nsteps, periodic)
    if not periodic:
        want = O.unframe(want, 1)
    assert np.array_equal(got, want)


def test_gol_known_patterns():
    env = (32, 32, 32)
    u = np.zeros((32, 32, 32), np.uint8)
    u[10:12, 10:12, 10:12] = 1   # still life
    u[20, 5, 7] = 1              # dies
    u[2, 2, 2] = u[2, 2, 3] = u[2, 3, 2] = u[3, 2, 2] = u[4, 4, 4] = 1
    with A.Context(env, (2, 2, 2), 1, 1, 1, 1) as c:
        a, b = c.field(), c.field()
        c.upload(a, u)
        res, _ = c.gol_run(a, b, 1)
        got = c.download(res)
    assert np.array_equal(got, O.gol(u, 1, periodic=True))
    assert got[10:12, 10:12, 10:12].all() and got[20, 5, 7] == 0 and got[3, 3, 3] == 1


def test_gol_full_size_R_invariance_and_population():
    # 512^3, 2x2x2 periodic: R = 1, 2, 4 agree bit for bit; population checked
    # against the oracle for the first generations
    env = (512, 512, 512)
    outs = {}
    for R in (1, 2, 4):
        for fused, tile in ((0, 0), (1, 1), (1, 5), (1, 11), (1, 13), (1, 14)):
            say("gol 512^3 R", R, "fused", fused, "tile", tile)
            with A.Context(env, (2, 2, 2), R, 1, 1, 1) as c:
                c.set_tuning(gol_tile=tile)
                a, b = c.field(), c.field()
                c.init_bernoulli(a, 1, int(0.3 * 16777216))
                res, _ = c.gol_run(a, b, 8, fused=fused)
                outs[(R, fused, tile)] = c.download(res)
    u = O.bernoulli((512, 512, 512), 1, 0.3)
    want = O.gol(u, 8, periodic=True)
    for key, o in outs.items():
        assert np.array_equal(o, want), key


@pytest.mark.parametrize("tile", [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12])
@pytest.mark.parametrize("zc", [4, 9, 64])
def test_gol_fused_chunking(zc, tile):
    # z-chunk lengths that do not divide B, and a single chunk, for every tile shape
    env = (96, 72, 64)
    with A.Context(env, (1, 2, 1), 3, 1, 0, 1) as c:
        c.set_tuning(gol_zc=zc, gol_step_zc=zc, gol_tile=tile)
        a, b = c.field(), c.field()
        c.init_bernoulli(a, 5, int(0.3 * 16777216))
        res, _ = c.gol_run(a, b, 9, fused=1)
        got = c.download(res)
    want = O.unframe(O.gol(O.frame(O.bernoulli((64, 72, 96), 5, 0.3), 1), 9, False), 1)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("gol_bytes", [0, 1])
@pytest.mark.parametrize("tile", [6, 7, 8, 9, 10, 11, 12, 3])
@pytest.mark.parametrize("R", [1, 2, 4])
@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("env,blocks", [((64, 48, 40), (2, 3, 2)), ((64, 30, 28), (1, 2, 2)),
                                        ((96, 26, 24), (3, 1, 1)), ((128, 20, 36), (2, 2, 3))])
def test_gol_bits_halo_from_neighbours(env, blocks, periodic, R, tile, gol_bytes):
    # block x ranges of whole 32-cell words: the bit-sliced kernel (tile 6)
    # reads every halo word from the owning neighbour, resolved z, y, x as the
    # ordered exchange fills it, and no exchange runs between its launches;
    # the byte-lane kernel (tile 3) runs with the exchange.  Single-block axes
    # wrap onto themselves when periodic; frozen faces keep their own halo.
    # gol_bytes 0 runs the bit-sliced windows on bit-packed fields, the last
    # generation (3R + 1 is not a whole number of windows) as a one-generation
    # launch of the same kernel, still packed; gol_bytes 1 runs that
    # generation step by step on the byte field after an exchange.
    if tile == 3 and gol_bytes:
        pytest.skip("the byte-lane kernel has no packed form")
    gens = 3 * R + 1
    say("gol pull", env, blocks, "R", R, "periodic", periodic, "tile", tile, "bytes", gol_bytes)
    with A.Context(env, blocks, R, 1, 1 if periodic else 0, 1) as c:
        c.set_tuning(gol_tile=tile, gol_zc=0, gol_bytes=gol_bytes)
        a, b = c.field(), c.field()
        c.init_bernoulli(a, 8, int(0.3 * 16777216))
        res, _ = c.gol_run(a, b, gens, fused=1)
        got = c.download(res)
    u = O.bernoulli((env[2], env[1], env[0]), 8, 0.3)
    if periodic:
        want = O.gol(u, gens, True)
    else:
        want = O.unframe(O.gol(O.frame(u, 1), gens, False), 1)
    assert np.array_equal(got, want)


WHOLE_ROWS = [((512, 40, 24), (2, 1, 1)), ((512, 36, 28), (1, 2, 2)),
              ((512, 24, 20), (16, 1, 1)), ((512, 48, 24), (4, 2, 1)),
              ((512, 20, 36), (8, 1, 2))]


@pytest.mark.parametrize("gol_bytes,zc", [(0, 0), (0, 7), (1, 0)])
@pytest.mark.parametrize("tile", [13, 14, 15])
@pytest.mark.parametrize("R", [1, 2, 3, 4])
@pytest.mark.parametrize("env,blocks", WHOLE_ROWS)
def test_gol_whole_rows(env, blocks, R, tile, gol_bytes, zc):
    # the whole-row tiles: a grid periodic in x over 512 cells, 1 to 16 blocks
    # across it (a lane's word then lies in block q / (B_x / 32)), partial y
    # tiles, z chunks that do not divide B_z, and 3R + 1 generations (the last
    # as a one-generation launch); packed and byte fields
    gens = 3 * R + 1
    say("gol whole rows", env, blocks, "R", R, "tile", tile, "bytes", gol_bytes, "zc", zc)
    with A.Context(env, blocks, R, 1, 1, 1) as c:
        c.set_tuning(gol_tile=tile, gol_zc=zc, gol_bytes=gol_bytes)
        a, b = c.field(), c.field()
        c.init_bernoulli(a, 17, int(0.3 * 16777216))
        res, _ = c.gol_run(a, b, gens, fused=1)
        got = c.download(res)
    u = O.bernoulli((env[2], env[1], env[0]), 17, 0.3)
    assert np.array_equal(got, O.gol(u, gens, True))


@pytest.mark.parametrize("tile", [13, 14])
@pytest.mark.parametrize("env,blocks,periodic", [((512, 40, 24), (2, 1, 1), False),
                                                 ((256, 40, 24), (2, 1, 1), True),
                                                 ((544, 40, 24), (1, 1, 1), True)])
def test_gol_whole_rows_elsewhere(env, blocks, periodic, tile):
    # where the grid is not periodic in x over 512 cells the whole-row tiles
    # run as tile 11
    R, gens = 4, 9
    say("gol whole rows fallback", env, blocks, "periodic", periodic, "tile", tile)
    with A.Context(env, blocks, R, 1, 1 if periodic else 0, 1) as c:
        c.set_tuning(gol_tile=tile, gol_zc=0, gol_bytes=0)
        a, b = c.field(), c.field()
        c.init_bernoulli(a, 19, int(0.3 * 16777216))
        res, _ = c.gol_run(a, b, gens, fused=1)
        got = c.download(res)
    u = O.bernoulli((env[2], env[1], env[0]), 19, 0.3)
    if periodic:
        want = O.gol(u, gens, True)
    else:
        want = O.unframe(O.gol(O.frame(u, 1), gens, False), 1)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("tile", [8, 11, 14])
@pytest.mark.parametrize("periodic", [False, True])
def test_gol_packed_runs_in_sequence(periodic, tile):
    # Two runs on the same fields and the fields of another context in
    # between: each run packs its input again (the packed pair is the
    # context's), whole windows only, then windows plus single steps, and the
    # timing helper's packed launches leave the byte fields untouched.
    # (tile 14 on a 512-cell periodic x extent: whole grid rows)
    env, blocks, R = ((512, 40, 36) if tile >= 13 else (128, 40, 36)), (2, 2, 1), 4
    u = O.bernoulli((env[2], env[1], env[0]), 12, 0.3)
    with A.Context(env, blocks, R, 1, 1 if periodic else 0, 1) as c, \
            A.Context((64, 32, 32), (1, 1, 1), 2, 1, 1, 1) as c2:
        c.set_tuning(gol_tile=tile, gol_zc=0, gol_bytes=0)
        a, b = c.field(), c.field()
        c.upload(a, u)
        say("gol packed run 1, periodic", periodic, "tile", tile)
        res, _ = c.gol_run(a, b, 8, fused=1)
        x, y = c2.field(), c2.field()
        c2.init_bernoulli(x, 3, int(0.3 * 16777216))
        c2.gol_run(x, y, 6, fused=1)
        dst = a if res == b else b
        say("gol packed run 2")
        res2, _ = c.gol_run(res, dst, 11, fused=1)
        got = c.download(res2)
        before = c.download(a), c.download(b)
        c.time_steps(2, a, b, 3)
        after = c.download(a), c.download(b)
    if periodic:
        want = O.gol(u, 19, True)
    else:
        want = O.unframe(O.gol(O.frame(u, 1), 19, False), 1)
    assert np.array_equal(got, want)
    assert all(np.array_equal(p, q) for p, q in zip(before, after))


@pytest.mark.parametrize("tile", [8, 11, 14])
@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("R,gens", [(4, 1), (4, 3), (4, 4), (4, 5), (2, 3), (1, 1), (1, 2)])
def test_gol_packed_first_and_last_launch(periodic, R, gens, tile):
    # Runs of one and two launches on packed fields: the last launch unpacks
    # as it stores; with no frozen face the first packs as it loads (a single
    # launch then goes from byte field to byte field), with frozen faces a
    # pack pass fills the packed frame first.
    env, blocks = ((512, 40, 36) if tile >= 13 else (128, 40, 36)), (2, 2, 1)
    u = O.bernoulli((env[2], env[1], env[0]), 21, 0.3)
    say("gol packed first/last launch, R", R, "gens", gens, "periodic", periodic, "tile", tile)
    with A.Context(env, blocks, R, 1, 1 if periodic else 0, 1) as c:
        c.set_tuning(gol_tile=tile, gol_zc=0, gol_bytes=0)
        a, b = c.field(), c.field()
        c.upload(a, u)
        res, _ = c.gol_run(a, b, gens, fused=1)
        got = c.download(res)
    if periodic:
        want = O.gol(u, gens, True)
    else:
        want = O.unframe(O.gol(O.frame(u, 1), gens, False), 1)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("wy,wx", [(4, 1), (4, 2), (4, 4), (8, 1), (8, 2), (16, 1), (4, 3),
                                   (2, 6), (3, 3)])
@pytest.mark.parametrize("ry", [1, 2, 4, 8])
@pytest.mark.parametrize("zc", [0, 5])
def test_heat_tuning_variants(wy, wx, ry, zc):
    env, D = (70, 40, 36), 3
    k = (0.021, 0.033, 0.017)
    p = A.heat_params(D, 0, k)
    with A.Context(env, (1, 2, 1), 2, D, 0, 4) as c:
        c.set_tuning(heat_wy=wy, heat_wx=wx, heat_ry=ry, heat_zc=zc)
        a, b = c.field(), c.field()
        c.init_uniform(a, 9)
        res, _ = c.heat_run(a, b, p, 6)
        got = c.download(res)
    g = O.frame(O.uniform((36, 40, 70), 9), D)
    want = O.unframe(O.heat_f32(g, D, 0, k, 0, (1, 1, 1), 6, False), D)
    assert bits_equal(got, want)


@pytest.mark.parametrize("tile", [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12])
@pytest.mark.parametrize("R", [1, 2, 3, 4])
@pytest.mark.parametrize("periodic", [False, True])
def test_gol_fused_tile_shapes(tile, R, periodic):
    # every tile shape of the fused kernels (byte-lane 1-3, bit-sliced 4-6),
    # blocks narrower and wider than a tile
    env, gens = (288, 40, 36), 2 * R + 1
    say("gol fused tile", tile, "R", R, "periodic", periodic)
    with A.Context(env, (2, 1, 1), R, 1, 1 if periodic else 0, 1) as c:
        c.set_tuning(gol_tile=tile, gol_zc=11)
        a, b = c.field(), c.field()
        c.init_bernoulli(a, 4, int(0.3 * 16777216))
        res, _ = c.gol_run(a, b, gens, fused=1)
        got = c.download(res)
    u = O.bernoulli((env[2], env[1], env[0]), 4, 0.3)
    if periodic:
        want = O.gol(u, gens, True)
    else:
        want = O.unframe(O.gol(O.frame(u, 1), gens, False), 1)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("env,blocks,R", [((512, 40, 24), (2, 1, 1), 2), ((96, 40, 24), (1, 1, 1), 4)])
def test_gol_autotune_keeps_an_explicit_choice(env, blocks, R, tmp_path):
    # the fused Game of Life tuner keeps a tile that runs as itself on the
    # geometry (whole-row tiles only on a 512-cell periodic x extent) with an
    # explicit z chunk, records it, and the run with it stays bit-exact
    rec = str(tmp_path / "rec.txt")
    say("gol autotune", env, blocks, R)
    with A.Context(env, blocks, R, 1, 1, 1) as c:
        a, b = c.field(), c.field()
        c.init_bernoulli(a, 23, int(0.3 * 16777216))
        t = c.autotune(2, a, b, None, rec)
        assert t == {k: v for k, v in c.get_tuning().items()}
        assert 1 <= t["gol_tile"] <= 15 and t["gol_zc"] > 0
        if env[0] != 512:
            assert t["gol_tile"] < 13
        c.init_bernoulli(a, 23, int(0.3 * 16777216))
        res, _ = c.gol_run(a, b, 2 * R + 1, fused=1)
        got = c.download(res)
    line = open(rec).read().strip().splitlines()[-1]
    assert "kind=gol_fused" in line and f"best=gol_tile:{t['gol_tile']},gol_zc:{t['gol_zc']}" in line
    u = O.bernoulli((env[2], env[1], env[0]), 23, 0.3)
    assert np.array_equal(got, O.gol(u, 2 * R + 1, True))


def test_heat_autotune_keeps_a_valid_choice(tmp_path):
    # the heat step tuner's choice is one of the built shapes, recorded, and
    # the run with it equals the oracle bit for bit
    env, D, k = (70, 40, 36), 4, (0.021, 0.033, 0.017)
    rec = str(tmp_path / "rec.txt")
    p = A.heat_params(D, 0, k)
    with A.Context(env, (1, 2, 1), 1, D, 0, 4) as c:
        a, b = c.field(), c.field()
        c.init_uniform(a, 9)
        t = c.autotune(0, a, b, p, rec)
        assert (t["heat_wy"], t["heat_wx"]) in {(4, 1), (4, 2), (4, 3), (4, 4), (8, 1), (8, 2),
                                                (16, 1), (2, 6), (3, 3), (0, 0)}
        assert t["heat_ry"] in (0, 1, 2, 4, 8)
        c.init_uniform(a, 9)
        res, _ = c.heat_run(a, b, p, 5)
        got = c.download(res)
    assert "kind=heat_step" in open(rec).read()
    g = O.frame(O.uniform((36, 40, 70), 9), D)
    want = O.unframe(O.heat_f32(g, D, 0, k, 0, (1, 1, 1), 5, False), D)
    assert bits_equal(got, want)
