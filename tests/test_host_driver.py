"""The C++ host driver (b2r_main) keeps b2r.cpp's CLI and its argument checks
(b2r.cpp:132-167, 54-59); these cases fail before any device call."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.environ.get("B2R_MAIN") or os.path.join(ROOT, "b2r_amd", "b2r_main")


def run(*args):
    return subprocess.run([EXE, *map(str, args)], capture_output=True, text=True, timeout=60)


def test_too_few_arguments():
    r = run(1, 2)
    assert r.returncode == 1 and "not sufficient parameter" in r.stdout


def test_indivisible_environment():
    r = run(1, 3, 1, 1, "--env", 10, 10, 10)
    assert r.returncode == 1 and "perfectly divisible" in r.stdout


def test_halo_deeper_than_half_block():
    r = run(2, 2, 2, 2, "--model", "eqwp", "--env", 16, 16, 16)
    assert r.returncode != 0 and "min(B)/2" in r.stderr


def test_unknown_model():
    r = run(1, 1, 1, 1, "--model", "lbm")
    assert r.returncode == 1 and "unknown model" in r.stdout


def test_unknown_output():
    r = run(1, 1, 1, 1, "--output", "hdf5")
    assert r.returncode == 1 and "unknown output" in r.stdout
