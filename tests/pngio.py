"""Minimal PNG reader for the test suite: 8-bit RGBA, non-interlaced, every
filter type (PNG specification, section 9).  Checks the signature and the
CRC of every chunk."""
from __future__ import annotations

import struct
import zlib

import numpy as np

SIGNATURE = b"\x89PNG\r\n\x1a\n"


def chunks(data: bytes):
    assert data[:8] == SIGNATURE, "not a PNG file"
    pos = 8
    while pos < len(data):
        (n,) = struct.unpack(">I", data[pos:pos + 4])
        kind = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        (crc,) = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert zlib.crc32(kind + body) & 0xFFFFFFFF == crc, f"bad CRC in {kind!r}"
        yield kind, body
        pos += 12 + n


def _paeth(a: int, b: int, c: int) -> int:
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    if pa <= pb and pa <= pc:
        return a
    return b if pb <= pc else c


def read_rgba(path) -> np.ndarray:
    """uint8 [height][width][4]."""
    data = open(path, "rb").read()
    idat = b""
    width = height = None
    kinds = []
    for kind, body in chunks(data):
        kinds.append(kind)
        if kind == b"IHDR":
            width, height, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", body)
            assert (depth, ctype, comp, filt, inter) == (8, 6, 0, 0, 0)
        elif kind == b"IDAT":
            idat += body
    assert kinds[0] == b"IHDR" and kinds[-1] == b"IEND"
    raw = zlib.decompress(idat)
    stride = width * 4
    assert len(raw) == height * (stride + 1)
    out = np.zeros((height, stride), np.uint8)
    prev = np.zeros(stride, np.int64)
    for y in range(height):
        f = raw[y * (stride + 1)]
        line = np.frombuffer(raw, np.uint8, stride, y * (stride + 1) + 1).astype(np.int64)
        cur = np.zeros(stride, np.int64)
        if f == 0:
            cur = line
        elif f == 2:
            cur = (line + prev) & 255
        else:
            for i in range(stride):
                a = cur[i - 4] if i >= 4 else 0
                c = prev[i - 4] if i >= 4 else 0
                pred = {1: a, 3: (a + prev[i]) // 2, 4: _paeth(a, prev[i], c)}[f]
                cur[i] = (line[i] + pred) & 255
        out[y] = cur
        prev = cur
    return out.reshape(height, width, 4)
