"""NPY writer of the library against the reference's own writer (golden
files from tests/golden/make_npy_golden.py) and numpy's reader."""
import os

import numpy as np
import pytest

from nimbus_amd import lib as L

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "npy")


@pytest.mark.parametrize("name", sorted(os.listdir(GOLD)))
def test_npy_bytes_equal_reference_writer(tmp_path, name):
    parts = name[:-4].split("_")
    dtype = np.float32 if parts[0] == "f4" else np.float64
    shape = tuple(int(x) for x in parts[1:])
    data = (0.25 * np.arange(int(np.prod(shape)), dtype=np.float64) - 1.5).astype(dtype).reshape(shape)
    out = tmp_path / name
    L.npy_write(str(out), data)
    assert out.read_bytes() == open(os.path.join(GOLD, name), "rb").read()
    back = np.load(str(out))
    assert back.dtype == dtype and back.shape == shape and np.array_equal(back, data)


@pytest.mark.parametrize("shape", [(1,), (17,), (5, 1000), (3, 3, 3)])
def test_npy_round_trips_through_numpy(tmp_path, shape):
    rng = np.random.default_rng(2)
    for dtype in (np.float32, np.float64):
        a = rng.standard_normal(shape).astype(dtype)
        p = tmp_path / f"a_{len(shape)}_{np.dtype(dtype).itemsize}.npy"
        L.npy_write(str(p), a)
        raw = p.read_bytes()
        assert raw[:8] == b"\x93NUMPY\x01\x00"
        hlen = raw[8] | (raw[9] << 8)
        assert (10 + hlen) % 16 == 0 and raw[10 + hlen - 1:10 + hlen] == b"\n"
        assert np.array_equal(np.load(str(p)), a)


def test_npy_refusals(tmp_path):
    with pytest.raises(L.NimbusError):
        L.npy_write(str(tmp_path / "x.npy"), np.zeros((2, 2, 2, 2), np.float32))
    with pytest.raises(L.NimbusError):
        L.npy_write(str(tmp_path / "no_such_dir" / "x.npy"), np.zeros(3, np.float32))
    with pytest.raises(L.NimbusError):
        L.npy_write(str(tmp_path / "x.npy"), np.zeros(3, np.int32))
