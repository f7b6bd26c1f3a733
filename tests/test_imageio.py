"""Image writers (kittlespt_amd/imageio.py): PNG round trip through an
independent decoder path (zlib + the PNG chunk layout), CRCs, rounding and
clamping; PFM header and row order."""
import struct
import zlib

import numpy as np
import pytest

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


def _read_exr(path):
    """Minimal OpenEXR reader for single-part, uncompressed scanline files,
    written from the file layout (magic, version, attributes, offset table,
    lines of [y, size, channel planes in chlist order])."""
    import struct
    data = open(path, "rb").read()
    assert struct.unpack("<ii", data[:8]) == (20000630, 2)
    pos, attrs = 8, {}
    while data[pos] != 0:
        e = data.index(b"\0", pos)
        name = data[pos:e].decode()
        e2 = data.index(b"\0", e + 1)
        typ = data[e + 1:e2].decode()
        (n,) = struct.unpack("<i", data[e2 + 1:e2 + 5])
        attrs[name] = (typ, data[e2 + 5:e2 + 5 + n])
        pos = e2 + 5 + n
    pos += 1
    chl, names = attrs["channels"][1], []
    q = 0
    while chl[q] != 0:
        e = chl.index(b"\0", q)
        names.append(chl[q:e].decode())
        assert struct.unpack("<i", chl[e + 1:e + 5])[0] == 2     # FLOAT
        q = e + 1 + 16
    assert attrs["compression"][1] == b"\0"
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    offs = struct.unpack(f"<{h}Q", data[pos:pos + 8 * h])
    img = np.zeros((h, w, len(names)), np.float32)
    for y in range(h):
        ly, size = struct.unpack("<ii", data[offs[y]:offs[y] + 8])
        assert ly == y and size == 4 * w * len(names)
        planes = np.frombuffer(data, "<f4", w * len(names), offs[y] + 8).reshape(len(names), w)
        img[y] = planes.T
    return names, img


def test_exr_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    rgba = (rng.random((6, 9, 4)) * 100.0 - 5.0).astype(np.float32)
    p = tmp_path / "a.exr"
    imageio.write_exr(str(p), rgba)
    names, img = _read_exr(str(p))
    assert names == ["A", "B", "G", "R"]
    got = img[..., [3, 2, 1, 0]][::-1]                  # back to RGBA, bottom-up rows
    assert np.array_equal(got.view(np.uint32), rgba.view(np.uint32))


# ---------------------------------------------------------------- the C-ABI writers (C++ callers)
def _c_write(fn, path, a, *extra):
    from kittlespt_amd import abi
    a = np.ascontiguousarray(a, np.float32)
    h, w, ch = a.shape
    return getattr(abi.lib(), fn)(str(path).encode(), a.ctypes.data, w, h, ch, *extra)


def _frame(seed, ch, h=7, w=9):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-0.2, 1.3, (h, w, ch)).astype(np.float32)
    a[0, 0, 0], a[1, 2, 1], a[2, 3, 2], a[3, 4, 0] = np.nan, np.inf, -np.inf, 0.5 / 255.0
    return a


@pytest.mark.parametrize("ch", [3, 4])
def test_c_writers_equal_python(tmp_path, ch):
    """kpt_write_exr and kpt_write_pfm write the Python writers' bytes;
    kpt_write_png_rgba8 the same pixels (NaN, +-inf and a value on a rounding
    boundary included) in a PNG the reader takes."""
    from kittlespt_amd import abi
    a = _frame(ch, ch)
    for flip in (1, 0):
        imageio.write_exr(str(tmp_path / "p.exr"), a, flip_y=bool(flip))
        assert _c_write("kpt_write_exr", tmp_path / "c.exr", a, flip) == abi.KPT_OK
        assert (tmp_path / "p.exr").read_bytes() == (tmp_path / "c.exr").read_bytes()
        imageio.write_png(str(tmp_path / "p.png"), a, flip_y=bool(flip))
        assert _c_write("kpt_write_png_rgba8", tmp_path / "c.png", a, flip) == abi.KPT_OK
        assert np.array_equal(imageio.read_png_rgba8(str(tmp_path / "p.png")),
                              imageio.read_png_rgba8(str(tmp_path / "c.png")))
    imageio.write_pfm(str(tmp_path / "p.pfm"), a)
    assert _c_write("kpt_write_pfm", tmp_path / "c.pfm", a) == abi.KPT_OK
    assert (tmp_path / "p.pfm").read_bytes() == (tmp_path / "c.pfm").read_bytes()


def test_c_writers_reject_bad_arguments(tmp_path):
    from kittlespt_amd import abi
    a = np.zeros((4, 5, 2), np.float32)
    assert _c_write("kpt_write_png_rgba8", tmp_path / "x.png", a, 1) == abi.KPT_ERR_INVALID_ARG   # 2 channels
    b = _frame(2, 3)
    assert _c_write("kpt_write_exr", tmp_path / "no" / "such" / "dir.exr", b, 1) == abi.KPT_ERR_INVALID_ARG
