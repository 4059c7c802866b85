"""The BTKF fixture/result format: Python writer and reader round trip, and
the layout of the header and sections (barytreek_amd/cpp/barytreek.hpp)."""
import struct

import numpy as np
import pytest

from barytreek_amd import fileio


def test_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    secs = {"x": rng.standard_normal(17), "pairs": np.arange(12, dtype=np.int32),
            "panels": np.arange(5, dtype=np.uint8), "big": np.arange(3, dtype=np.int64)}
    p = tmp_path / "f.btk"
    fileio.write(str(p), secs)
    back = fileio.read(str(p))
    assert sorted(back) == sorted(secs)
    for k, v in secs.items():
        assert back[k].dtype == v.dtype and np.array_equal(back[k], v)


def test_layout(tmp_path):
    p = tmp_path / "f.btk"
    fileio.write(str(p), {"q": np.array([1.5, -2.0])})
    raw = p.read_bytes()
    assert raw[:4] == b"BTKF" and struct.unpack("<3I", raw[4:16]) == (1, 1, 0)
    assert raw[16:40].rstrip(b"\0") == b"q"
    assert struct.unpack("<IIQ", raw[40:56]) == (0, 0, 2)
    assert np.frombuffer(raw[56:72], np.float64).tolist() == [1.5, -2.0]
    assert len(raw) == 72


def test_rejects_other_files(tmp_path):
    p = tmp_path / "f.btk"
    p.write_bytes(b"NOPE" + bytes(12))
    with pytest.raises(ValueError):
        fileio.read(str(p))
