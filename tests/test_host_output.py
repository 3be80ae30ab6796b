"""The on-disk writers of the host library (b2r_amd/host/b2r_output.cpp), called
on host arrays through ctypes: file names and layouts of eqwp.cu:616-650 and
heat-3D.hpp:105-127, read back with b2r_amd.output."""
import ctypes as C
import os

import numpy as np
import pytest

from b2r_amd import output

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST_SO = os.environ.get("B2R_HOST_LIB") or os.path.join(ROOT, "b2r_amd", "libb2r_host.so")


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(HOST_SO):
        pytest.fail("libb2r_host.so not built (make -C b2r_amd/host)")
    return C.CDLL(HOST_SO)


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def test_displacement_bin_layout(lib, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    a = np.random.default_rng(3).standard_normal((5, 4, 3)).astype(np.float32)  # z, y, x
    lib._Z20write_block_data_biniPfiii(17, _fp(a), 3, 4, 5)
    raw = (tmp_path / "displacement-17.bin").read_bytes()
    assert np.frombuffer(raw[:12], np.int32).tolist() == [3, 4, 5]
    assert len(raw) == 12 + a.nbytes and raw[12:] == a.tobytes()
    assert np.array_equal(output.read_displacement_bin(tmp_path / "displacement-17.bin"), a)


def test_displacement_txt_layout(lib, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    a = np.array([[[1.5, -2.0, 3.25e-7], [0.0, 123456.78, 1.0 / 3.0]]], np.float32)
    lib._Z16write_block_dataiPfiii(4, _fp(a), 3, 2, 1)
    text = (tmp_path / "displacement-4.txt").read_text()
    # default stream formatting: 6 significant digits, every value followed by ','
    assert text == "1.5,-2,3.25e-07,\n0,123457,0.333333,\n"
    got = output.read_displacement_txt(tmp_path / "displacement-4.txt", (3, 2, 1))
    assert np.allclose(got, a.astype(np.float64), rtol=1e-5, atol=0)
    # subnormal and tiny values keep six significant digits too
    b = np.array([[[1.4e-45, -3.6041e-42, 5.2708e-41, 1e-38, -7.5e30, 3.4e38]]], np.float32)
    lib._Z16write_block_dataiPfiii(5, _fp(b), 6, 1, 1)
    got = output.read_displacement_txt(tmp_path / "displacement-5.txt", (6, 1, 1))
    assert np.allclose(got, b.astype(np.float64), rtol=1e-5, atol=0)


def test_block_step_txt_layout(lib, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    for name, v in (("B2R_R", 2), ("B2R_BLOCK_SIZE_X", 3), ("B2R_BLOCK_SIZE_Y", 2),
                    ("B2R_BLOCK_SIZE_Z", 2)):
        C.c_int.in_dll(lib, name).value = v
    blk = np.arange(12, dtype=np.float32).reshape(2, 2, 3) / 8
    lib._Z21print_block_data_stepifPKf(5, C.c_float(10.0), _fp(blk))
    text = (tmp_path / "R-2-rank-5-i-10.0.txt").read_text()
    assert text.splitlines()[:3] == ["slice: [0, :, :]", "0.00,0.12,0.25,", "0.38,0.50,0.62,"]
    got = output.read_block_step_txt(tmp_path / "R-2-rank-5-i-10.0.txt")
    assert got.shape == (2, 2, 3) and np.allclose(got, blk, atol=0.005 + 1e-9)
