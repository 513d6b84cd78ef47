"""Tiny named-array record file shared by the tests and the reference harness.

Layout: b"CFDGREC1", then records of
  u32 name length | name bytes | u8 dtype code | u64 element count | raw little-endian data
dtype codes: 0 float64, 1 int32, 2 uint8, 3 int64.
"""
from __future__ import annotations

import struct

import numpy as np

MAGIC = b"CFDGREC1"
_CODES = {np.dtype(np.float64): 0, np.dtype(np.int32): 1, np.dtype(np.uint8): 2, np.dtype(np.int64): 3}
_TYPES = {v: k for k, v in _CODES.items()}


def write_records(path, arrays):
    with open(path, "wb") as f:
        f.write(MAGIC)
        for name, a in arrays.items():
            a = np.ascontiguousarray(a)
            if a.dtype not in _CODES:
                raise TypeError(f"{name}: unsupported dtype {a.dtype}")
            nb = name.encode()
            f.write(struct.pack("<I", len(nb)))
            f.write(nb)
            f.write(struct.pack("<BQ", _CODES[a.dtype], a.size))
            f.write(a.tobytes())


def read_records(path):
    out = {}
    with open(path, "rb") as f:
        if f.read(8) != MAGIC:
            raise ValueError(f"{path}: not a record file")
        while True:
            hdr = f.read(4)
            if not hdr:
                break
            (ln,) = struct.unpack("<I", hdr)
            name = f.read(ln).decode()
            code, count = struct.unpack("<BQ", f.read(9))
            dt = _TYPES[code]
            out[name] = np.frombuffer(f.read(count * dt.itemsize), dtype=dt).copy()
    return out
