"""The "BTKF" v1 fixture/result format (see barytreek_amd/cpp/barytreek.hpp):
a 16-byte header, then named sections of f64 / i32 / u8 / i64 arrays."""
from __future__ import annotations

import struct

import numpy as np

MAGIC, VERSION = 0x464B5442, 1
DTYPES = {0: np.float64, 1: np.int32, 2: np.uint8, 3: np.int64}
CODES = {np.dtype(v): k for k, v in DTYPES.items()}


def write(path: str, sections: dict) -> None:
    with open(path, "wb") as f:
        f.write(struct.pack("<4I", MAGIC, VERSION, len(sections), 0))
        for name in sorted(sections):
            a = np.ascontiguousarray(sections[name])
            if a.dtype not in CODES:
                a = a.view(np.uint8)
            raw = a.tobytes()
            nb = name.encode()
            if len(nb) >= 24:
                raise ValueError(f"section name too long: {name}")
            f.write(nb.ljust(24, b"\0"))
            f.write(struct.pack("<IIQ", CODES[a.dtype], 0, a.size))
            f.write(raw)
            f.write(b"\0" * ((8 - len(raw) % 8) % 8))


def read(path: str) -> dict:
    out = {}
    with open(path, "rb") as f:
        magic, version, n, _ = struct.unpack("<4I", f.read(16))
        if magic != MAGIC or version != VERSION:
            raise ValueError(f"not a BTKF v1 file: {path}")
        for _ in range(n):
            name = f.read(24).rstrip(b"\0").decode()
            dt, _, count = struct.unpack("<IIQ", f.read(16))
            if dt not in DTYPES:
                raise ValueError(f"corrupt section {name!r}")
            a = np.frombuffer(f.read(count * np.dtype(DTYPES[dt]).itemsize), DTYPES[dt])
            if a.size != count:
                raise ValueError(f"truncated section {name!r}")
            f.read((8 - a.nbytes % 8) % 8)
            out[name] = a.copy()
    return out
