"""Image output with the reference CLI's tone mapping (main.cpp:150-165):
clamp to [0, 1] (PostProcessing.cpp:30-33), then fastPow(x, 1/2.2) * 255
truncated, for all four channels."""
from __future__ import annotations

import struct
import zlib

import numpy as np


def fast_pow(a: np.ndarray, b: float) -> np.ndarray:
    """Ankerl's approximate pow (2012): high word <- b*(hi - 1072632447) + 1072632447, low word <- 0."""
    a = np.ascontiguousarray(a, np.float64)
    bits = a.view(np.uint64)
    hi = (bits >> np.uint64(32)).astype(np.int64).astype(np.int32).astype(np.float64)
    nhi = (b * (hi - 1072632447) + 1072632447)
    nhi = np.trunc(nhi).astype(np.int64).astype(np.int32).view(np.uint32).astype(np.uint64)
    return (nhi << np.uint64(32)).view(np.float64)


def to_ldr(rgba: np.ndarray) -> np.ndarray:
    x = np.clip(rgba.astype(np.float32), 0, 1).astype(np.float64)
    return (fast_pow(x, 1.0 / 2.2) * 255).astype(np.uint8)


def write_png(path, rgba8: np.ndarray):
    h, w, c = rgba8.shape
    raw = b"".join(b"\x00" + rgba8[y].tobytes() for y in range(h))

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)

    color = {4: 6, 3: 2}[c]
    png = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, color, 0, 0, 0))
    png += chunk(b"IDAT", zlib.compress(raw, 6)) + chunk(b"IEND", b"")
    with open(path, "wb") as f:
        f.write(png)
