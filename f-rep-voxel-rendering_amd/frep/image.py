"""Image files of rendered frames: PNG (8-bit RGBA, no filtering, zlib) and
binary PPM (RGB).  Rows are written top to bottom from row height-1 of the
render, whose row 0 is the bottom of the image (dy grows with the pixel row,
vulkan/shaders/raytracer/main.comp:161-184)."""
from __future__ import annotations

import struct
import zlib

import numpy as np


def _rows_top_down(rgba: np.ndarray) -> np.ndarray:
    a = np.asarray(rgba, dtype=np.uint8)
    if a.ndim != 3 or a.shape[2] != 4:
        raise ValueError("expected an (height, width, 4) uint8 image")
    return a[::-1]


def png_bytes(rgba: np.ndarray) -> bytes:
    img = _rows_top_down(rgba)
    h, w, _ = img.shape
    raw = b"".join(b"\x00" + img[y].tobytes() for y in range(h))

    def chunk(kind: bytes, data: bytes) -> bytes:
        return struct.pack(">I", len(data)) + kind + data + \
            struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF)
    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0))
            + chunk(b"IDAT", zlib.compress(raw, 6)) + chunk(b"IEND", b""))


def write_png(path, rgba: np.ndarray) -> None:
    with open(path, "wb") as f:
        f.write(png_bytes(rgba))


def read_png(path) -> np.ndarray:
    """Reads back what write_png writes (RGBA8, filter 0 rows); returns the
    image in render row order."""
    data = open(path, "rb").read()
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError("not a PNG file")
    pos, idat, w, h = 8, b"", 0, 0
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        kind, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        if kind == b"IHDR":
            w, h, depth, ctype = struct.unpack(">IIBB", body[:10])
            if depth != 8 or ctype != 6:
                raise ValueError("only 8-bit RGBA")
        elif kind == b"IDAT":
            idat += body
        pos += 12 + n
    raw = zlib.decompress(idat)
    rows = np.frombuffer(raw, dtype=np.uint8).reshape(h, 1 + 4 * w)
    if np.any(rows[:, 0] != 0):
        raise ValueError("only unfiltered rows")
    return rows[:, 1:].reshape(h, w, 4)[::-1].copy()


def write_ppm(path, rgba: np.ndarray) -> None:
    img = _rows_top_down(rgba)
    h, w, _ = img.shape
    with open(path, "wb") as f:
        f.write(f"P6 {w} {h} 255\n".encode())
        f.write(np.ascontiguousarray(img[:, :, :3]).tobytes())
