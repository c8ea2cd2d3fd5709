"""The reference's binary field files (PETSc binary Vec format) on the CPU:
byte layout checked against numpy's big-endian view of an independently
written file, round trips, .info files and error returns.  No device is used
(cles_petsc_vec_write / cles_petsc_vec_read need no context)."""
import os

import numpy as np
import pytest

import chimera_comp_les_amd as cl

CLASSID = 1211214  # VEC_FILE_CLASSID


def test_write_layout_is_petsc_binary(tmp_path):
    x = np.random.default_rng(1).standard_normal(37) * 1e3
    x[3] = -0.0
    x[5] = np.inf
    path = str(tmp_path / "Qfield00007_0.dat")
    cl.petsc_vec_write(path, x, block_size=5 if x.size % 5 == 0 else 1)
    raw = open(path, "rb").read()
    head = np.frombuffer(raw[:8], dtype=">i4")
    assert list(head) == [CLASSID, x.size]
    body = np.frombuffer(raw[8:], dtype=">f8")
    assert body.size == x.size and np.array_equal(body.view(">u8").astype(np.uint64), x.view(np.uint64))
    assert len(raw) == 8 + 8 * x.size


def test_info_file_for_blocked_vectors(tmp_path):
    path = str(tmp_path / "sigma00001_0.dat")
    cl.petsc_vec_write(path, np.ones(9), block_size=3)
    assert open(path + ".info").read() == "-vecload_block_size 3\n"
    path1 = str(tmp_path / "pfield00001_0.dat")
    cl.petsc_vec_write(path1, np.ones(9), block_size=1)
    assert not os.path.exists(path1 + ".info")


def test_read_back_and_read_foreign_file(tmp_path):
    x = np.random.default_rng(2).standard_normal(1000)
    path = str(tmp_path / "a.dat")
    cl.petsc_vec_write(path, x)
    assert np.array_equal(cl.petsc_vec_read(path, x.size), x)
    # a file written by numpy in the same format reads back bit for bit
    y = np.random.default_rng(3).standard_normal(300000)   # crosses the 64 Ki-value buffer
    foreign = str(tmp_path / "b.dat")
    with open(foreign, "wb") as f:
        f.write(np.array([CLASSID, y.size], dtype=">i4").tobytes())
        f.write(y.astype(">f8").tobytes())
    assert np.array_equal(cl.petsc_vec_read(foreign, y.size), y)


def test_errors(tmp_path):
    path = str(tmp_path / "c.dat")
    cl.petsc_vec_write(path, np.ones(10))
    with pytest.raises(cl.CLESError, match="expected 11"):
        cl.petsc_vec_read(path, 11)
    bad = str(tmp_path / "d.dat")
    with open(bad, "wb") as f:
        f.write(np.array([1211216, 10], dtype=">i4").tobytes() + np.ones(10, ">f8").tobytes())
    with pytest.raises(cl.CLESError, match="not a PETSc binary Vec"):
        cl.petsc_vec_read(bad, 10)
    short = str(tmp_path / "e.dat")
    with open(short, "wb") as f:
        f.write(np.array([CLASSID, 10], dtype=">i4").tobytes() + np.ones(4, ">f8").tobytes())
    with pytest.raises(cl.CLESError, match="truncated"):
        cl.petsc_vec_read(short, 10)
    with pytest.raises(cl.CLESError, match="cannot open"):
        cl.petsc_vec_read(str(tmp_path / "missing.dat"), 10)
