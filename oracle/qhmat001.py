"""QHMAT001 reader/writer in numpy -- TEST INFRASTRUCTURE ONLY.

Independent restatement of the reference file format (src/qhmat/io.cpp:127-257):
little-endian; "QHMAT001", u64 version=1, u64 order, u64 block count,
u64 permutation[order]; per block u64 row0,row1,col0,col1, u8 admissible,
u8 dense, then either f64[m*n] or u64 rank, u64 menu, f64 truncation error,
f64 quantization bound, u8 widths[rank], f64 left scales[rank], f64 right
scales[rank], u32 left words, u32 right words (component j of a factor with R
rows has ceil(R*b_j/32) words); trailer u64 FNV-1a-64 of all preceding bytes.
"""

from __future__ import annotations

import struct

import numpy as np

FNV_BASIS = 14695981039346656037
FNV_PRIME = 1099511628211
MASK64 = (1 << 64) - 1


def fnv1a64(data: bytes) -> int:
    h = FNV_BASIS
    for b in data:
        h = ((h ^ b) * FNV_PRIME) & MASK64
    return h


def words_for(rows: int, bits: int) -> int:
    return (rows * bits + 31) // 32


class _Cursor:
    def __init__(self, data: bytes):
        self.data = data
        self.at = 0

    def take(self, n: int) -> bytes:
        if self.at + n > len(self.data):
            raise ValueError("truncated")
        out = self.data[self.at:self.at + n]
        self.at += n
        return out

    def u8(self) -> int:
        return self.take(1)[0]

    def u64(self) -> int:
        return struct.unpack("<Q", self.take(8))[0]

    def f64(self) -> float:
        return struct.unpack("<d", self.take(8))[0]

    def array(self, dtype, count: int) -> np.ndarray:
        dt = np.dtype(dtype).newbyteorder("<")
        return np.frombuffer(self.take(dt.itemsize * count), dtype=dt).astype(dt.newbyteorder("="))


def parse(data: bytes, verify_checksum: bool = True) -> dict:
    """Returns {'order', 'perm', 'blocks': [...]}; each block a dict."""
    if len(data) < 24:
        raise ValueError("truncated")
    payload, trailer = data[:-8], data[-8:]
    if verify_checksum and fnv1a64(payload) != struct.unpack("<Q", trailer)[0]:
        raise ValueError("checksum")
    c = _Cursor(payload)
    if c.take(8) != b"QHMAT001":
        raise ValueError("magic")
    if c.u64() != 1:
        raise ValueError("version")
    order = c.u64()
    count = c.u64()
    perm = c.array(np.uint64, order).astype(np.int64)
    blocks = []
    for _ in range(count):
        r0, r1, c0, c1 = (c.u64() for _ in range(4))
        admissible, dense = c.u8(), c.u8()
        m, n = r1 - r0, c1 - c0
        blk = {"rect": (r0, r1, c0, c1), "admissible": bool(admissible), "dense": bool(dense)}
        if dense:
            blk["values"] = c.array(np.float64, m * n).reshape(m, n)
        else:
            rank = c.u64()
            blk["menu_size"] = c.u64()
            blk["truncation_error"] = c.f64()
            blk["quantization_bound"] = c.f64()
            widths = c.array(np.uint8, rank).astype(np.int64)
            blk["widths"] = widths
            blk["lscale"] = c.array(np.float64, rank)
            blk["rscale"] = c.array(np.float64, rank)
            lw = sum(words_for(m, int(b)) for b in widths)
            rw = sum(words_for(n, int(b)) for b in widths)
            blk["lwords"] = c.array(np.uint32, lw)
            blk["rwords"] = c.array(np.uint32, rw)
        blocks.append(blk)
    if c.at != len(payload):
        raise ValueError("trailing")
    return {"order": order, "perm": perm, "blocks": blocks}


def component_offsets(rows: int, widths) -> np.ndarray:
    """First word of each component in a factor of `rows` rows."""
    sizes = [words_for(rows, int(b)) for b in widths]
    return np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64) if sizes else np.zeros(0, np.int64)


def write(store: dict) -> bytes:
    """Serialises a dict of the shape parse() returns."""
    out = bytearray(b"QHMAT001")
    out += struct.pack("<QQQ", 1, store["order"], len(store["blocks"]))
    out += np.asarray(store["perm"], dtype="<u8").tobytes()
    for blk in store["blocks"]:
        out += struct.pack("<QQQQBB", *blk["rect"], int(blk["admissible"]), int(blk["dense"]))
        if blk["dense"]:
            out += np.asarray(blk["values"], dtype="<f8").tobytes()
            continue
        out += struct.pack("<QQdd", len(blk["widths"]), blk["menu_size"],
                           blk["truncation_error"], blk["quantization_bound"])
        out += np.asarray(blk["widths"], dtype=np.uint8).tobytes()
        out += np.asarray(blk["lscale"], dtype="<f8").tobytes()
        out += np.asarray(blk["rscale"], dtype="<f8").tobytes()
        out += np.asarray(blk["lwords"], dtype="<u4").tobytes()
        out += np.asarray(blk["rwords"], dtype="<u4").tobytes()
    out += struct.pack("<Q", fnv1a64(bytes(out)))
    return bytes(out)
