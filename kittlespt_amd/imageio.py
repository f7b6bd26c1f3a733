"""Image files from readbacks (the reference shows its frames through GL
textures, renderer.cu:197-205; here frames are read back and written out).

* `write_png(path, rgba)`: the display buffer (`Context.readback(4)`, already
  tone-mapped and sRGB-encoded, values in [0, 1]) as 8-bit RGBA PNG;
* `write_pfm(path, rgb)`: linear float data (`readback(0)`/`estimate()`) as a
  little-endian colour PFM;
* `write_exr(path, rgba)`: the same linear data as an uncompressed scanline
  OpenEXR (FLOAT A/B/G/R channels);
* `read_png_rgba8(path)`: the inverse of `write_png` for 8-bit RGBA files it
  wrote (non-interlaced, filter type 0), used by the tests.

The frame buffers keep the reference's GL-texture convention: row 0 is the
bottom row (camera rays map pixel row py to NDC y = 2(py + 0.5)/H - 1,
camera.cuh:16-26), so `write_png` flips to PNG's top-down rows by default and
`write_pfm` writes rows as they are (PFM is bottom-up).
"""
from __future__ import annotations

import struct
import zlib

import numpy as np


def _chunk(tag: bytes, data: bytes) -> bytes:
    return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)


def to_rgba8(rgba) -> np.ndarray:
    """[H, W, 3|4] floats in [0, 1] -> uint8 RGBA (round to nearest, clamped;
    alpha 1 when absent)."""
    a = np.asarray(rgba, np.float32)
    if a.ndim != 3 or a.shape[2] not in (3, 4):
        raise ValueError(f"expected [H, W, 3|4], got {a.shape}")
    if a.shape[2] == 3:
        a = np.concatenate([a, np.ones(a.shape[:2] + (1,), np.float32)], axis=2)
    a = np.nan_to_num(a, nan=0.0, posinf=1.0, neginf=0.0)
    return np.clip(np.floor(a * 255.0 + 0.5), 0, 255).astype(np.uint8)


def write_png(path: str, rgba, flip_y: bool = True) -> None:
    px = to_rgba8(rgba)
    if flip_y:
        px = px[::-1]
    h, w = px.shape[:2]
    raw = b"".join(b"\x00" + px[y].tobytes() for y in range(h))
    ihdr = struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)
    with open(path, "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n")
        f.write(_chunk(b"IHDR", ihdr))
        f.write(_chunk(b"IDAT", zlib.compress(raw, 6)))
        f.write(_chunk(b"IEND", b""))


def read_png_rgba8(path: str) -> np.ndarray:
    with open(path, "rb") as f:
        data = f.read()
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError("not a PNG file")
    pos, idat, w, h = 8, b"", 0, 0
    while pos < len(data):
        (n,) = struct.unpack(">I", data[pos:pos + 4])
        tag, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        pos += 12 + n
        if tag == b"IHDR":
            w, h, depth, ctype, _, _, interlace = struct.unpack(">IIBBBBB", body)
            if (depth, ctype, interlace) != (8, 6, 0):
                raise ValueError("only 8-bit RGBA non-interlaced PNG is read")
        elif tag == b"IDAT":
            idat += body
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + 4 * w)
    if np.any(raw[:, 0] != 0):
        raise ValueError("only filter type 0 is read")
    return raw[:, 1:].reshape(h, w, 4).copy()


def write_exr(path: str, rgba, flip_y: bool = True) -> None:
    """Linear float data ([H, W, 3|4]) as an OpenEXR file: single part,
    scanline, no compression, FLOAT channels A, B, G, R (stored in
    alphabetical order, as the format requires), increasing-Y lines.  Rows
    are flipped to the format's top-down order by default (the buffers keep
    the GL bottom-up convention)."""
    a = np.asarray(rgba, np.float32)
    if a.ndim != 3 or a.shape[2] not in (3, 4):
        raise ValueError(f"expected [H, W, 3|4], got {a.shape}")
    if a.shape[2] == 3:
        a = np.concatenate([a, np.ones(a.shape[:2] + (1,), np.float32)], axis=2)
    if flip_y:
        a = a[::-1]
    h, w = a.shape[:2]

    def attr(name, typ, data):
        return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(data)) + data

    chans = b""
    for c in ("A", "B", "G", "R"):
        chans += c.encode() + b"\0" + struct.pack("<iB3xii", 2, 0, 1, 1)   # FLOAT, pLinear 0, x/y sampling 1
    chans += b"\0"
    box = struct.pack("<iiii", 0, 0, w - 1, h - 1)
    header = (struct.pack("<ii", 20000630, 2) +
              attr("channels", "chlist", chans) +
              attr("compression", "compression", b"\0") +
              attr("dataWindow", "box2i", box) +
              attr("displayWindow", "box2i", box) +
              attr("lineOrder", "lineOrder", b"\0") +
              attr("pixelAspectRatio", "float", struct.pack("<f", 1.0)) +
              attr("screenWindowCenter", "v2f", struct.pack("<ff", 0.0, 0.0)) +
              attr("screenWindowWidth", "float", struct.pack("<f", 1.0)) + b"\0")
    line_bytes = 4 * w * 4
    first = len(header) + 8 * h
    offsets = struct.pack(f"<{h}Q", *[first + y * (8 + line_bytes) for y in range(h)])
    planes = np.ascontiguousarray(a[..., [3, 2, 1, 0]].transpose(0, 2, 1), "<f4")   # [H, channel, W]
    with open(path, "wb") as f:
        f.write(header)
        f.write(offsets)
        for y in range(h):
            f.write(struct.pack("<ii", y, line_bytes))
            f.write(planes[y].tobytes())


def write_pfm(path: str, rgb) -> None:
    """Colour PFM, little-endian (scale -1); the buffer's row 0 (the bottom
    row) is written first, as the format orders rows."""
    a = np.asarray(rgb, np.float32)
    if a.ndim != 3 or a.shape[2] < 3:
        raise ValueError(f"expected [H, W, >=3], got {a.shape}")
    a = np.ascontiguousarray(a[:, :, :3], "<f4")
    h, w = a.shape[:2]
    with open(path, "wb") as f:
        f.write(f"PF\n{w} {h}\n-1.0\n".encode("ascii"))
        f.write(a.tobytes())
