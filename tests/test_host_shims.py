"""Host-side logic of the reference-named shims and the host heat step, run
without a GPU: the neighbour flags of b2r_sync_flag(rank) (b2r_halo_sync.cu:
140-183), the host step b2r_heat_cpu_step against the oracle and the
reference's own cpu_fdm_heat_diffuse1, heat_3D_cpu_main (heat-3D.cpp:136-184)
against the reference's, and the b2r.cpp-shaped driver reading its own globals
through the shims (ELF interposition)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import _oracle as O
from conftest import has_gpu
from b2r_amd import _abi as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.environ.get("B2R_HOST_LIB") or os.path.join(ROOT, "b2r_amd", "libb2r_host.so")


def host_lib():
    return C.CDLL(HOST)


def _int(lib, name):
    return C.c_int.in_dll(lib, name)


def _expected_flags(rank, bd, periodic, three_d=True):
    bx = (rank % (bd[0] * bd[1])) % bd[0]
    by = (rank % (bd[0] * bd[1])) // bd[0]
    bz = rank // (bd[0] * bd[1])
    c = (bx, by, bz)
    out = []
    for a in range(3 if three_d else 2):
        lo = 1 if c[a] >= 1 else (-1 if periodic else 0)
        hi = 1 if c[a] < bd[a] - 1 else (-1 if periodic else 0)
        out += [lo, hi]
    return out


@pytest.mark.parametrize("periodic", [0, 1])
@pytest.mark.parametrize("bd", [(3, 2, 2), (1, 1, 1), (2, 3, 1)])
def test_sync_flag_shim(bd, periodic):
    lib = host_lib()
    for name, v in zip(("BLOCK_DIM_X", "BLOCK_DIM_Y", "BLOCK_DIM_Z"), bd):
        _int(lib, name).value = v
    _int(lib, "B2R_PERIODIC").value = periodic
    _int(lib, "B2R_ENV_3D").value = 1
    flags = (C.c_byte * 6).in_dll(lib, "b2r_halo_update")
    for rank in range(bd[0] * bd[1] * bd[2]):
        lib._Z13b2r_sync_flagi(rank)
        assert list(flags) == _expected_flags(rank, bd, periodic), rank


def test_sync_flag_shim_2d_writes_four_entries():
    # a 2D build's b2r_halo_update has 4 entries (b2r.h:38-42): 4, 5 untouched
    lib = host_lib()
    for name, v in zip(("BLOCK_DIM_X", "BLOCK_DIM_Y", "BLOCK_DIM_Z"), (2, 2, 1)):
        _int(lib, name).value = v
    _int(lib, "B2R_PERIODIC").value = 0
    _int(lib, "B2R_ENV_3D").value = 0
    flags = (C.c_byte * 6).in_dll(lib, "b2r_halo_update")
    flags[4] = flags[5] = 77
    try:
        lib._Z13b2r_sync_flagi(3)
        assert list(flags)[:4] == _expected_flags(3, (2, 2, 1), 0, False)
        assert flags[4] == 77 and flags[5] == 77
    finally:
        _int(lib, "B2R_ENV_3D").value = 1


@pytest.mark.parametrize("mode", [A.HEAT_F32, A.HEAT_REF])
@pytest.mark.parametrize("D", [1, 2, 3, 4])
def test_cpu_step_matches_oracle(D, mode):
    # one inner step b2r_i = 0 on a grid with a D-wide frame updates exactly
    # the owned cells, as the oracle's frozen step
    rng = np.random.default_rng(D + 10 * mode)
    owned = rng.uniform(0, 1, (9, 11, 13)).astype(np.float32)
    g = O.frame(owned, D, 0.25)
    k = (0.021, 0.033, 0.017)
    p = A.heat_params(D, mode, k, 0.1 * 0.1, (1.0, 1.0, 0.25))
    got = g
    for _ in range(3):
        got = A.heat_cpu_step(p, got, 0, nthreads=3)
    want = O.heat_f32(g, D, mode, k, 0.1 * 0.1, (1.0, 1.0, 0.25), 3, False)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_cpu_step_thread_counts_and_fork():
    # the host step's worker pool: any thread count (more threads than rows
    # included) gives the one-thread result bit for bit, and a process forked
    # after the pool started runs the step with workers of its own
    rng = np.random.default_rng(3)
    g = rng.uniform(0, 1, (7, 9, 21)).astype(np.float32)
    p = A.heat_params(2, A.HEAT_F32, (0.05, 0.07, 0.03))
    want = A.heat_cpu_step(p, g, 0, nthreads=1)
    for nt in (2, 3, 8, 64, 0):
        assert np.array_equal(A.heat_cpu_step(p, g, 0, nthreads=nt).view(np.uint32),
                              want.view(np.uint32)), nt
    pid = os.fork()
    if pid == 0:
        ok = np.array_equal(A.heat_cpu_step(p, g, 0, nthreads=4), want)
        os._exit(0 if ok else 3)
    _, status = os.waitpid(pid, 0)
    assert os.WIFEXITED(status) and os.WEXITSTATUS(status) == 0
    assert np.array_equal(A.heat_cpu_step(p, g, 0, nthreads=4), want)


@pytest.mark.skipif(not O.ref_available(), reason="the reference's heat-3D.cpp is not built here")
@pytest.mark.parametrize("b2r_i", [0, 1, 2])
def test_cpu_step_matches_reference_step(b2r_i):
    # REF mode, D = 1: the reference's own cpu_fdm_heat_diffuse1 (heat-3D.cpp:48-62)
    # with its region [i+1, N-2-i], bit for bit
    rng = np.random.default_rng(7 + b2r_i)
    blk = rng.uniform(0, 1, (14, 12, 17)).astype(np.float32)
    p = A.heat_params(1, A.HEAT_REF, adt=0.1 * 0.1, h2=(1.0, 1.0, 0.25))
    got = A.heat_cpu_step(p, blk, b2r_i)
    want = O.reference_block_step(blk, b2r_i)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def _heat_cpu_main(lib, blk, R, D):
    a = np.ascontiguousarray(blk, np.float32).copy()
    nz, ny, nx = a.shape
    f = lib._Z16heat_3D_cpu_mainPfiiiii
    f.restype = C.c_float
    f(a.ctypes.data_as(C.c_void_p), R, D, nx, ny, nz)
    return a


@pytest.mark.skipif(not O.ref_available(), reason="the reference's heat-3D.cpp is not built here")
@pytest.mark.parametrize("R", [1, 2, 3, 4])
def test_heat_3D_cpu_main_matches_reference(R):
    # heat-3D.cpp:136-184 at D = 1: the R-step ping-pong over the reference's
    # own step.  With R even the result is h_in itself in both; with R odd the
    # reference copies back the z slab [R*B2R_D, Ngz - R*B2R_D) with the
    # config's B2R_D (2) in place of the step's D (a defect), the shim the
    # slab [R*D, Ngz - R*D): checked against the reference's steps with that copy
    rng = np.random.default_rng(R)
    blk = rng.uniform(0, 1, (16, 13, 15)).astype(np.float32)
    got = _heat_cpu_main(host_lib(), blk, R, 1)
    # the two buffers of heat-3D.cpp:143-166: h_in itself and a copy, alternating
    buf = [blk.copy(), blk.copy()]
    s = 0
    for i in range(R):
        buf[1 - s] = _ref_step_into(buf[s], buf[1 - s], i)
        s = 1 - s
    want = buf[0]
    if s == 1:  # the result is in the copy: its owned z slab goes back
        want[R:-R] = buf[1][R:-R]
    else:
        ref = blk.copy()
        O.ref_lib().ref_heat_3D_cpu_main(ref.ctypes.data_as(C.c_void_p), R, 1, 15, 13, 16)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def _ref_step_into(src, dst, i):
    """cpu_fdm_heat_diffuse1 writing into an existing buffer (the ping-pong keeps
    the other buffer's cells outside the region)."""
    out = np.ascontiguousarray(dst, np.float32).copy()
    nz, ny, nx = src.shape
    O.ref_lib().ref_cpu_fdm_heat_diffuse1(src.ctypes.data_as(C.c_void_p),
                                         out.ctypes.data_as(C.c_void_p), nx, ny, nz, i)
    return out


@pytest.mark.skipif(has_gpu(), reason="checks the no-device path; the GPU tests run the driver")
def test_dropin_driver_reads_its_own_globals():
    # the b2r.cpp-shaped driver defines the globals of b2r.h itself; the shims
    # must read its simu_env_init results (B2R_B_*, B2R_BLOCK_SIZE_*): without a
    # device the run then stops at the first device call, not at the geometry
    exe = os.path.join(ROOT, "tests", "drop_in", "_bin", "frozen_d2", "b2r_dropin")
    r = subprocess.run([exe, "1", "2", "1", "1", "4", "/dev/null"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode != 0
    assert "must be set first" not in r.stderr, r.stderr
    assert "hipGetDeviceCount" in r.stderr or "device" in r.stderr, r.stderr
