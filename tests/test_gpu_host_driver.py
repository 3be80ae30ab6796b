"""The C++ driver b2r_main (b2r.cpp's time loop over the reference-named shims:
b2r_sync_init/flag, boundary_updating, heat_3D_gpu_steps, the GoL step and
eqwp_gpu_main) run end to end on the GPU, its final grid against the oracle bit
for bit."""
import os
import subprocess
import sys

import numpy as np
import pytest

import _eqwp as Q
import _oracle as O
from b2r_amd import output

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.environ.get("B2R_MAIN") or os.path.join(ROOT, "b2r_amd", "b2r_main")


def run(tmp_path, *args):
    out = str(tmp_path / "dump.bin")
    print("b2r_main", *args, flush=True)
    sys.stdout.flush()
    # the reference's output files go to the working directory
    r = subprocess.run([EXE, *map(str, args), "--dump", out], capture_output=True, text=True,
                       timeout=120, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Iterations with B2R_R" in r.stdout and "GLUP/s" in r.stdout
    return out


@pytest.mark.parametrize("R,blocks,D", [(2, (2, 2, 1), 2), (1, (1, 2, 2), 1), (3, (2, 1, 1), 1)])
def test_driver_heat_matches_oracle(tmp_path, R, blocks, D):
    env, iters = (32, 32, 24), 10
    f = run(tmp_path, R, *blocks, "--model", "heat", "--env", *env, "--D", D, "--iters", iters)
    got = np.fromfile(f, np.float32).reshape(env[2], env[1], env[0])
    u = O.frame(O.uniform((env[2], env[1], env[0]), 0), D)
    want = O.unframe(O.heat_f32(u, D, 1, (0, 0, 0), 0.1 * 0.1, (1.0, 1.0, 0.25), iters, False), D)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("periodic", [False, True])
def test_driver_gol_matches_oracle(tmp_path, periodic):
    env, iters = (32, 32, 32), 7
    args = [2, 2, 2, 2, "--model", "gol", "--env", *env, "--iters", iters]
    if periodic:
        args.append("--periodic")
    got = np.fromfile(run(tmp_path, *args), np.uint8).reshape(env[2], env[1], env[0])
    u = O.bernoulli((env[2], env[1], env[0]), 1, 0.3)
    if periodic:
        want = O.gol(u, iters, True)
    else:
        want = O.gol(O.frame(u, 1), iters, False)[1:-1, 1:-1, 1:-1]
    assert np.array_equal(got, want)


def test_driver_eqwp_matches_oracle(tmp_path):
    # the reference model of eqwp.cu: constant medium, the y-force patch of
    # every block (its padded coordinates, B2R_D = 4 here), forced while the
    # window start is below 50
    env, iters = (40, 40, 40), 12
    got = np.fromfile(run(tmp_path, 1, 2, 1, 1, "--model", "eqwp", "--env", *env, "--iters",
                          iters), np.float32).reshape(env[2], env[1], env[0])
    st, _ = Q.forward(env, (2, 1, 1), 1, 4, iters)
    assert np.any(st[1])
    assert np.array_equal(got.view(np.uint32), st[1].view(np.uint32))


def test_driver_eqwp_periodic_matches_oracle(tmp_path):
    # _PERIODIC_ for eqwp (b2r_config.h:12): the reference model on a periodic grid
    env, iters = (40, 40, 40), 14
    got = np.fromfile(run(tmp_path, 2, 1, 2, 1, "--model", "eqwp", "--env", *env, "--iters",
                          iters, "--periodic"), np.float32).reshape(env[2], env[1], env[0])
    st, _ = Q.forward(env, (1, 2, 1), 2, 4, iters, periodic=True)
    assert np.any(st[1][:, :4])  # the wave crossed the y face
    assert np.array_equal(got.view(np.uint32), st[1].view(np.uint32))


def test_driver_eqwp_writes_displacement(tmp_path):
    # eqwp.cu:689-709 with the write switched on: displacement-<ts>.bin and .txt
    # after the last step group (ts = its first step), against the oracle; the
    # force enters the displacement at the patch (eqwp.cu:229-235)
    env, R, iters = (32, 32, 32), 2, 10
    run(tmp_path, R, 1, 1, 1, "--model", "eqwp", "--env", *env, "--iters", iters, "--output",
        "both")
    got = output.read_displacement_bin(tmp_path / ("displacement-%d.bin" % (iters - R)))
    _, _, disp = Q.forward(env, (1, 1, 1), R, 4, iters, disp=True)
    assert np.any(disp)
    assert np.array_equal(got.view(np.uint32), disp.view(np.uint32))
    txt = output.read_displacement_txt(tmp_path / ("displacement-%d.txt" % (iters - R)), env)
    # six significant digits: a relative rounding error of at most 5e-6 (compared in
    # float64: rtol times a float32 subnormal would underflow in float32)
    assert np.allclose(txt, got.astype(np.float64), rtol=1e-5, atol=0)


def test_driver_eqwp_force_schedule_follows_window_starts(tmp_path):
    # eqwp.cu:135-139 forces every inner step of a window that starts before
    # FRC_TIME = 50: with R = 4 steps 48..51 are forced too (52 steps)
    env, R, iters = (32, 32, 32), 4, 56
    assert Q.forced_steps(R) == 52
    got = np.fromfile(run(tmp_path, R, 1, 1, 1, "--model", "eqwp", "--env", *env, "--iters",
                          iters), np.float32).reshape(env[2], env[1], env[0])
    st, _ = Q.forward(env, (1, 1, 1), R, 4, iters)
    assert np.array_equal(got.view(np.uint32), st[1].view(np.uint32))


def test_driver_heat_writes_block_dumps(tmp_path):
    # heat-3D.hpp:105-127: one padded-block text dump per block, "%0.2f"
    env, R, blocks, D, iters = (32, 32, 24), 2, (2, 2, 1), 2, 6
    run(tmp_path, R, *blocks, "--model", "heat", "--env", *env, "--D", D, "--iters", iters,
        "--output", "txt")
    u = O.frame(O.uniform((env[2], env[1], env[0]), 0), D)
    want = O.unframe(O.heat_f32(u, D, 1, (0, 0, 0), 0.1 * 0.1, (1.0, 1.0, 0.25), iters, False), D)
    h = R * D
    B = [env[a] // blocks[a] for a in range(3)]
    for r in range(blocks[0] * blocks[1] * blocks[2]):
        cx, cy, cz = r % blocks[0], r // blocks[0] % blocks[1], r // (blocks[0] * blocks[1])
        blk = output.read_block_step_txt(tmp_path / ("R-%d-rank-%d-i-%0.1f.txt" % (R, r, iters)))
        assert blk.shape == (B[2] + 2 * h, B[1] + 2 * h, B[0] + 2 * h)
        own = blk[h:h + B[2], h:h + B[1], h:h + B[0]]
        ref = want[cz * B[2]:(cz + 1) * B[2], cy * B[1]:(cy + 1) * B[1], cx * B[0]:(cx + 1) * B[0]]
        assert np.allclose(own, ref, rtol=0, atol=0.005 + 1e-6)


@pytest.mark.parametrize("model,periodic", [("heat", False), ("gol", True), ("gol", False)])
def test_driver_2d_matches_oracle(tmp_path, model, periodic):
    # ENV_DIM_Z = 1: the 2D environment (_ENV_3D_ = 0, b2r.cpp:63-67, 142-146);
    # heat runs F32 with k = Alpha*Dt/h^2 on x and y and no z term
    env, R, iters = (32, 32, 1), 2, 7
    args = [R, 2, 2, 1, "--model", model, "--env", *env, "--iters", iters]
    if periodic:
        args.append("--periodic")
    if model == "heat":
        args += ["--D", 1]
    raw = np.fromfile(run(tmp_path, *args), np.float32 if model == "heat" else np.uint8)
    got = raw.reshape(1, env[1], env[0])
    if model == "heat":
        u = O.frame(O.uniform((1, env[1], env[0]), 0), 1)
        k = (np.float32(0.01), np.float32(0.01), 0.0)
        want = O.unframe(O.heat_f32(u, 1, 0, k, 0.0, (1, 1, 1), iters, False), 1)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    else:
        u = O.bernoulli((1, env[1], env[0]), 1, 0.3)
        if periodic:
            want = O.gol(np.pad(u, ((1, 1), (0, 0), (0, 0))), iters, 3)[1:-1]
        else:
            want = O.gol(O.frame(u, 1), iters, False)[1:-1, 1:-1, 1:-1]
        assert np.array_equal(got, want)
