"""The reference's host interface on the GPU: the b2r.cpp-shaped driver
(b2r.cpp:125-250 without MPI) built against a b2r_config.h and linked to the
shim library, heat_3D_gpu_main / gol_3D_gpu_main on padded host blocks with
every edge-flag pattern (heat-3D.cu:583-642, gol-3D.cu:119-155), and the
neighbour flags of b2r_sync_flag through the C ABI (b2r_halo_sync.cu:140-183);
all against the oracle."""
import ctypes as C
import itertools
import os
import subprocess
import sys

import numpy as np
import pytest

import _eqwp as Q
import _oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.environ.get("B2R_HOST_LIB") or os.path.join(ROOT, "b2r_amd", "libb2r_host.so")
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


# ------------------------------------------------------------------ drop-in driver
DROPIN = [
    # config dir, env of its b2r_config.h, B2R_D there, periodic, R, blocks, N_ITER
    ("frozen_d2", (48, 40, 32), 2, False, 1, (2, 1, 1), 10),
    ("frozen_d2", (48, 40, 32), 2, False, 2, (2, 1, 2), 12),
    ("periodic_d4", (40, 48, 40), 4, True, 2, (1, 2, 1), 12),
    ("periodic_d4", (40, 48, 40), 4, True, 1, (2, 2, 2), 9),
]


@pytest.mark.parametrize("cfg,env,dcfg,periodic,R,blocks,iters", DROPIN)
def test_dropin_driver_matches_oracle(tmp_path, cfg, env, dcfg, periodic, R, blocks, iters):
    exe = os.path.join(ROOT, "tests", "drop_in", "_bin", cfg, "b2r_dropin")
    out = str(tmp_path / "vy.bin")
    say("b2r_dropin", cfg, R, blocks, iters)
    r = subprocess.run([exe, str(R), *map(str, blocks), str(iters), out], capture_output=True,
                       text=True, timeout=120, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    # rank 0's neighbour flags (b2r_sync_flag before simu_env_init, b2r.cpp:169)
    flags = [int(v) for v in r.stdout.split("halo flags:")[1].split()]
    want_flags = []
    for a in range(3):
        want_flags += [-1 if periodic else 0, 1 if blocks[a] > 1 else (-1 if periodic else 0)]
    assert flags == want_flags
    got = np.fromfile(out, np.float32).reshape(env[2], env[1], env[0])
    st = Q.forward(env, blocks, R, dcfg, iters, periodic=periodic)[0]
    assert np.any(st[1])
    assert bits_equal(got, st[1])


# ------------------------------------------------------------------ model entries
def host_lib():
    return C.CDLL(HOST)


def _ref_block_steps(blk, R, D, frozen_lo, frozen_hi, heat):
    """R inner steps of the reference's plain kernel on one padded block: step
    i updates [D(i+1), N-1-D(i+1)] per axis, clipped to the owned range
    [R*D, N-1-R*D] on the faces marked frozen; each step is the oracle's step
    on the box with a frame of D cells around it."""
    h = R * D
    # the two device buffers of heat-3D.cu:587-628 / gol-3D.cu:124-145: step i
    # reads one and writes its region into the other, whose cells outside the
    # region still hold the state of two steps before
    buf = [np.ascontiguousarray(blk).copy(), np.ascontiguousarray(blk).copy()]
    s = 0
    for i in range(R):
        cur = buf[s]
        lo, hi = [], []
        for a in range(3):  # axes x, y, z
            n = blk.shape[2 - a]
            l_, h_ = D * (i + 1), n - 1 - D * (i + 1)
            if frozen_lo[a]:
                l_ = max(l_, h)
            if frozen_hi[a]:
                h_ = min(h_, n - 1 - h)
            lo.append(l_)
            hi.append(h_)
        sub = cur[lo[2] - D:hi[2] + D + 1, lo[1] - D:hi[1] + D + 1, lo[0] - D:hi[0] + D + 1]
        if heat:
            new = O.heat_f32(sub, D, A.HEAT_REF, (0, 0, 0), 0.1 * 0.1, (1.0, 1.0, 0.25), 1, False)
        else:
            new = O.gol(sub, 1, False)
        buf[1 - s][lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = new[D:-D, D:-D, D:-D]
        s = 1 - s
    return buf[s]


EF = list(itertools.product((-1, 0, 1), repeat=3))


@pytest.mark.parametrize("verify", [0, 1])
@pytest.mark.parametrize("R,D", [(1, 1), (2, 1), (2, 2), (1, 4)])
def test_heat_3D_gpu_main_every_edge_flag(R, D, verify):
    # heat-3D.cu:583-642: the active kernel ignores ef (verify 0); the verify
    # semantics (heat-3D.cu:57-158, B2R_HEAT_VERIFY 1) freeze the faces ef
    # marks.  Only the owned z slab [R*D, Ngz-R*D) is copied back.
    lib = host_lib()
    C.c_int.in_dll(lib, "B2R_D").value = D
    C.c_int.in_dll(lib, "B2R_HEAT_VERIFY").value = verify
    f = lib._Z16heat_3D_gpu_mainPfiPciii
    f.restype = C.c_double
    h = R * D
    shape = (10 + 2 * h, 12 + 2 * h, 14 + 2 * h)  # z, y, x
    rng = np.random.default_rng(R * 10 + D)
    try:
        for ef in (EF if verify else EF[:3]):
            say("heat_3D_gpu_main", R, D, verify, ef)
            blk = rng.uniform(0, 1, shape).astype(np.float32)
            got = blk.copy()
            efa = (C.c_char * 3)(*[bytes([v & 0xFF]) for v in ef])
            f(got.ctypes.data_as(C.c_void_p), R, efa, shape[2], shape[1], shape[0])
            lo = [verify and v == -1 for v in ef]
            hi = [verify and v == 1 for v in ef]
            res = _ref_block_steps(blk, R, D, lo, hi, True)
            want = blk.copy()
            want[h:shape[0] - h] = res[h:shape[0] - h]
            assert bits_equal(got, want), ef
    finally:
        C.c_int.in_dll(lib, "B2R_HEAT_VERIFY").value = 0
        C.c_int.in_dll(lib, "B2R_D").value = 1


@pytest.mark.parametrize("R", [1, 2, 3])
def test_gol_3D_gpu_main(R):
    # gol-3D.cu:119-155: R generations of the plain kernel on the block, ef
    # unused, the whole block copied back
    lib = host_lib()
    f = lib._Z15gol_3D_gpu_mainPhiPciii
    f.restype = C.c_double
    shape = (12 + 2 * R, 9 + 2 * R, 20 + 2 * R)
    blk = (np.random.default_rng(R).uniform(0, 1, shape) < 0.3).astype(np.uint8)
    for ef in ((0, 0, 0), (-1, 1, 0)):
        say("gol_3D_gpu_main", R, ef)
        got = blk.copy()
        efa = (C.c_char * 3)(*[bytes([v & 0xFF]) for v in ef])
        f(got.ctypes.data_as(C.c_void_p), R, efa, shape[2], shape[1], shape[0])
        assert np.array_equal(got, _ref_block_steps(blk, R, 1, [0] * 3, [0] * 3, False)), ef


# ------------------------------------------------------------------ sync flags
@pytest.mark.parametrize("periodic", [0, 1])
def test_sync_flag_abi_3x2x2(periodic):
    with A.Context((24, 16, 16), (3, 2, 2), 1, 1, periodic, 4) as c:
        for rank in range(12):
            bx, by, bz = rank % 3, (rank // 3) % 2, rank // 6
            want = []
            for cc, n in ((bx, 3), (by, 2), (bz, 2)):
                want.append(1 if cc > 0 else (-1 if periodic else 0))
                want.append(1 if cc < n - 1 else (-1 if periodic else 0))
            assert c.sync_flag(rank) == want, rank


# ------------------------------------------------------------------ GoL with the exchange
@pytest.mark.parametrize("R,periodic", [(1, True), (2, True), (4, True), (4, False), (3, True)])
def test_gol_fused_with_pack_unpack_exchange(R, periodic):
    # gol_exchange = 1: boundary_updating (pack/unpack) before every
    # R-generation launch, the bit-sliced kernel reading its own halo
    env = (64, 40, 48)
    say("gol exchange", R, periodic)
    with A.Context(env, (2, 2, 2), R, 1, int(periodic), 1) as c:
        c.set_tuning(gol_exchange=1)
        a, b = c.field(), c.field()
        c.init_bernoulli(a, 1, int(0.3 * 16777216))
        res, _ = c.gol_run(a, b, 12, fused=1)
        got = c.download(res)
    u = O.bernoulli((env[2], env[1], env[0]), 1, 0.3)
    want = O.gol(u, 12, True) if periodic else O.gol(O.frame(u, 1), 12, False)[1:-1, 1:-1, 1:-1]
    assert np.array_equal(got, want)


@pytest.mark.parametrize("tile", [13, 14])
@pytest.mark.parametrize("R", [1, 2, 4])
def test_gol_whole_rows_with_pack_unpack_exchange(R, tile):
    # the whole-row tiles after the pack/unpack exchange: y and z halos read
    # from the block's own (exchanged) halo, x neighbours from the grid row
    env = (512, 40, 48)
    say("gol exchange whole rows", R, tile)
    with A.Context(env, (2, 2, 2), R, 1, 1, 1) as c:
        c.set_tuning(gol_exchange=1, gol_tile=tile, gol_zc=0)
        a, b = c.field(), c.field()
        c.init_bernoulli(a, 2, int(0.3 * 16777216))
        res, _ = c.gol_run(a, b, 3 * R + 1, fused=1)
        got = c.download(res)
    u = O.bernoulli((env[2], env[1], env[0]), 2, 0.3)
    assert np.array_equal(got, O.gol(u, 3 * R + 1, True))


def test_el_gradient_plan_groups():
    # the groups of the batched gradient follow the memory budget
    env = (24, 20, 16)
    ctxs, shots = [], []
    try:
        for _ in range(4):
            c = A.Context(env, (1, 1, 1), 1, 4, A.BOUNDARY_FROZEN, 4)
            s = A.ElasticShot(c, 1e-3, (0.1, 0.1, 0.1))
            s.set_receivers([[3, 3, 3]], 1, 200)
            ctxs.append(c)
            shots.append(s)
        n, big = A.el_gradient_plan(shots, 200, 10)
        assert n == 1 and big > 0
        per_shot = big // 4
        for c in ctxs:
            c.set_tuning(el_grad_mb=max(1, (2 * per_shot + (1 << 20) - 1) >> 20))
        n2, big2 = A.el_gradient_plan(shots, 200, 10)
        assert n2 >= 2 and big2 <= (max(1, (2 * per_shot + (1 << 20) - 1) >> 20) << 20)
    finally:
        for s in shots:
            s.close()
        for c in ctxs:
            c.close()
