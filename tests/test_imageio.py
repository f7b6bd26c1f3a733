"""Image writers (kittlespt_amd/imageio.py): PNG round trip through an
independent decoder path (zlib + the PNG chunk layout), CRCs, rounding and
clamping; PFM header and row order."""
import struct
import zlib

import numpy as np

from kittlespt_amd import imageio


def test_png_round_trip_and_chunks(tmp_path):
    rng = np.random.default_rng(3)
    img = rng.uniform(-0.2, 1.2, (7, 5, 4)).astype(np.float32)
    img[0, 0, 0] = np.nan
    p = str(tmp_path / "a.png")
    imageio.write_png(p, img)
    got = imageio.read_png_rgba8(p)
    want = np.clip(np.floor(np.nan_to_num(img, nan=0.0) * 255 + 0.5), 0, 255).astype(np.uint8)[::-1]
    assert np.array_equal(got, want)
    data = open(p, "rb").read()
    pos = 8
    tags = []
    while pos < len(data):
        (n,) = struct.unpack(">I", data[pos:pos + 4])
        tag, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        (crc,) = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(tag + body) & 0xFFFFFFFF
        tags.append(tag)
        pos += 12 + n
    assert tags == [b"IHDR", b"IDAT", b"IEND"]
    w, h = struct.unpack(">II", data[16:24])
    assert (w, h) == (5, 7)


def test_png_rgb_gets_opaque_alpha(tmp_path):
    p = str(tmp_path / "b.png")
    imageio.write_png(p, np.full((2, 3, 3), 0.5, np.float32), flip_y=False)
    got = imageio.read_png_rgba8(p)
    assert np.all(got[..., 3] == 255) and np.all(got[..., :3] == 128)


def test_pfm_header_and_rows(tmp_path):
    img = np.arange(2 * 3 * 4, dtype=np.float32).reshape(2, 3, 4)
    p = str(tmp_path / "c.pfm")
    imageio.write_pfm(p, img)
    data = open(p, "rb").read()
    assert data.startswith(b"PF\n3 2\n-1.0\n")
    body = np.frombuffer(data[len(b"PF\n3 2\n-1.0\n"):], "<f4").reshape(2, 3, 3)
    assert np.array_equal(body, img[:, :, :3])
