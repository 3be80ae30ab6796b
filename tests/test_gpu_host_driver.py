"""The C++ driver b2r_main (b2r.cpp's time loop over the reference-named shims:
b2r_sync_init/flag, boundary_updating, heat_3D_gpu_steps, the GoL step and
eqwp_gpu_main) run end to end on the GPU, its final grid against the oracle bit
for bit."""
import os
import subprocess
import sys

import numpy as np
import pytest

import _elastic as E
import _oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "b2r_amd", "b2r_main")


def run(tmp_path, *args):
    out = str(tmp_path / "dump.bin")
    print("b2r_main", *args, flush=True)
    sys.stdout.flush()
    r = subprocess.run([EXE, *map(str, args), "--dump", out], capture_output=True, text=True,
                       timeout=120)
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
    # the reference model of eqwp.cu: constant medium, y-force patch for 50 steps
    env, iters = (40, 40, 40), 12
    got = np.fromfile(run(tmp_path, 1, 2, 1, 1, "--model", "eqwp", "--env", *env, "--iters",
                          iters), np.float32).reshape(env[2], env[1], env[0])
    gx, gy, gz = env
    shape = (gz, gy, gx)
    lam = np.full(shape, 32.0, np.float32)
    mu = np.full(shape, 174.87, np.float32)
    b = np.full(shape, np.float32(1.0 / 1800.0), np.float32)
    src = np.array([[x, gy - 7, z] for x in range(gx // 2, gx // 2 + gx // 10)
                    for z in range(gz // 2, gz // 2 + gz // 10)], np.int32)
    wav = np.full((iters, len(src)), -10000.0, np.float32)
    st, _ = E.forward(env, lam, mu, b, np.float32(0.5), (np.float32(0.1),) * 3, iters, src,
                      [1] * len(src), wav, np.array([[1, 1, 1]], np.int32), 1)
    assert np.any(st[1])
    assert np.array_equal(got.view(np.uint32), st[1].view(np.uint32))
