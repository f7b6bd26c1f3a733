"""glTF 2.0 import (kittlespt_amd/gltf.py, the SampleApp importer
model_importer.hpp:51-456): PNG decoding checked against an encoder written
from the PNG specification's filter definitions, matrices against closed
forms, and whole files (.gltf with an embedded buffer, .glb) against the
arrays they were built from.  No reference-side assets exist here; the files
are synthetic.
"""
import base64
import json
import math
import struct
import zlib

import numpy as np
import pytest

from kittlespt_amd import gltf


# ---------------------------------------------------------------- PNG
def _encode_png(px, ctype, filters, palette=None):
    """Encoder: each row filtered with filters[y % len(filters)] (PNG spec 9.2:
    the filtered byte is raw - predictor(left, up, upper-left))."""
    h, w = px.shape[:2]
    chans = px.shape[2] if px.ndim == 3 else 1
    rows = px.reshape(h, w * chans).astype(np.int32)
    out = b""
    prev = np.zeros(w * chans, np.int32)
    for y in range(h):
        f = filters[y % len(filters)]
        cur = rows[y]
        enc = np.zeros_like(cur)
        for x in range(len(cur)):
            a = int(cur[x - chans]) if x >= chans else 0
            b = int(prev[x])
            c = int(prev[x - chans]) if x >= chans else 0
            if f == 0:
                p = 0
            elif f == 1:
                p = a
            elif f == 2:
                p = b
            elif f == 3:
                p = (a + b) // 2
            else:
                pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
                p = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
            enc[x] = (int(cur[x]) - p) % 256
        out += bytes([f]) + enc.astype(np.uint8).tobytes()
        prev = cur

    def chunk(tag, data):
        return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)
    png = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, ctype, 0, 0, 0))
    if palette is not None:
        png += chunk(b"PLTE", palette.astype(np.uint8).tobytes())
    return png + chunk(b"IDAT", zlib.compress(out)) + chunk(b"IEND", b"")


@pytest.mark.parametrize("ctype,chans", [(2, 3), (6, 4), (0, 1), (4, 2)])
def test_png_all_filters_round_trip(ctype, chans):
    rng = np.random.default_rng(ctype)
    px = rng.integers(0, 256, (7, 5, chans), dtype=np.uint8)
    got = gltf.decode_png(_encode_png(px, ctype, [0, 1, 2, 3, 4, 4, 3]))
    want = {3: px, 4: px[..., :3], 1: np.repeat(px, 3, 2), 2: np.repeat(px[..., :1], 3, 2)}[chans]
    assert got.shape == (7, 5, 3) and np.array_equal(got, want)


def test_png_palette():
    pal = np.array([[255, 0, 0], [0, 255, 0], [0, 0, 255]], np.uint8)
    idx = np.array([[0, 1, 2], [2, 1, 0]], np.uint8)[..., None]
    got = gltf.decode_png(_encode_png(idx, 3, [1, 2], palette=pal))
    assert np.array_equal(got, pal[idx[..., 0]])


def test_png_rejects_other_formats():
    with pytest.raises(ValueError):
        gltf.decode_png(b"\xff\xd8\xff\xe0JFIF")


# ---------------------------------------------------------------- matrices
def _apply(m, p):           # m stored [col][row]
    v = np.array([p[0], p[1], p[2], 1.0])
    return (m.astype(np.float64).T @ v)[:3]


def test_quaternion_closed_forms():
    s = math.sqrt(0.5)
    r = gltf.quat_to_mat4(0.0, s, 0.0, s)                 # 90 degrees about +y
    assert np.allclose(_apply(r, (1, 0, 0)), (0, 0, -1), atol=1e-6)
    assert np.allclose(_apply(r, (0, 0, 1)), (1, 0, 0), atol=1e-6)
    assert np.array_equal(gltf.quat_to_mat4(0, 0, 0, 1), np.eye(4, dtype=np.float32))


def test_trs_order_is_translate_rotate_scale():
    s = math.sqrt(0.5)
    m = gltf.node_model({"translation": [1, 2, 3], "rotation": [0, s, 0, s], "scale": [2, 2, 2]})
    # scale first (x -> 2x), then the rotation (x -> -z), then the translation
    assert np.allclose(_apply(m, (1, 0, 0)), (1, 2, 1), atol=1e-5)
    mm = gltf.node_model({"matrix": list(range(16))})
    assert np.array_equal(mm.reshape(16), np.arange(16, dtype=np.float32))


def test_camera_view_is_the_inverse_of_its_placement():
    s = math.sqrt(0.5)
    node = {"translation": [0, 1, 5], "rotation": [0, 2 * s, 0, 2 * s]}   # unnormalised: normalised first
    v = gltf.camera_view(node)
    # the camera position maps to the view-space origin; the camera's local z
    # points along world +x (90 degrees about y), so world -x is view -z
    assert np.allclose(_apply(v, (0, 1, 5)), (0, 0, 0), atol=1e-5)
    assert np.allclose(_apply(v, (-1, 1, 5)), (0, 0, -1), atol=1e-5)


# ---------------------------------------------------------------- files
def _scene_doc(png_bytes, binary):
    """Two mesh nodes (one TRS, one matrix), one camera node, a child node that
    must be skipped, two materials with extensions and a texture."""
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], np.float32)
    N = np.tile(np.array([[0, 0, 1]], np.float32), (4, 1))
    U = np.array([[0, 0], [1, 0], [0, 1], [1, 1]], np.float32)
    I = np.array([0, 1, 2, 2, 1, 3], np.uint16)
    blob = P.tobytes() + N.tobytes() + U.tobytes() + I.tobytes()
    pad = (4 - len(blob) % 4) % 4
    blob += b"\0" * pad
    img_off = len(blob)
    blob += png_bytes + b"\0" * ((4 - len(png_bytes) % 4) % 4)
    views = [
        {"buffer": 0, "byteOffset": 0, "byteLength": 48},
        {"buffer": 0, "byteOffset": 48, "byteLength": 48},
        {"buffer": 0, "byteOffset": 96, "byteLength": 32},
        {"buffer": 0, "byteOffset": 128, "byteLength": 12},
        {"buffer": 0, "byteOffset": img_off, "byteLength": len(png_bytes)},
    ]
    acc = [
        {"bufferView": 0, "componentType": 5126, "count": 4, "type": "VEC3"},
        {"bufferView": 1, "componentType": 5126, "count": 4, "type": "VEC3"},
        {"bufferView": 2, "componentType": 5126, "count": 4, "type": "VEC2"},
        {"bufferView": 3, "componentType": 5123, "count": 6, "type": "SCALAR"},
        {"bufferView": 3, "byteOffset": 6, "componentType": 5123, "count": 3, "type": "SCALAR"},
    ]
    s = math.sqrt(0.5)
    doc = {
        "asset": {"version": "2.0"},
        "scene": 0,
        "scenes": [{"nodes": [0, 1, 2, 3]}],
        "nodes": [
            {"mesh": 0, "translation": [1, 2, 3], "rotation": [0, s, 0, s], "scale": [2, 2, 2], "children": [4]},
            {"mesh": 1, "matrix": [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 5, 0, 0, 1]},
            {"camera": 0, "translation": [0, 1, 5]},
            {"name": "empty"},
            {"mesh": 1, "translation": [9, 9, 9]},     # a child: the reference does not visit it
        ],
        "meshes": [
            {"primitives": [{"attributes": {"POSITION": 0, "NORMAL": 1, "TEXCOORD_0": 2}, "indices": 3,
                             "material": 1}]},
            {"primitives": [{"attributes": {"POSITION": 0, "NORMAL": 1, "TEXCOORD_0": 2}, "indices": 4,
                             "material": 0}]},
        ],
        "materials": [
            {"pbrMetallicRoughness": {"baseColorFactor": [0.2, 0.3, 0.4, 1.0], "metallicFactor": 0.0,
                                      "roughnessFactor": 0.7}},
            {"pbrMetallicRoughness": {"baseColorTexture": {"index": 0}, "metallicFactor": 1.0,
                                      "roughnessFactor": 0.1},
             "emissiveFactor": [1.0, 0.5, 0.25], "normalTexture": {"index": 0, "scale": 0.5},
             "extensions": {"KHR_materials_transmission": {"transmissionFactor": 0.25},
                            "KHR_materials_ior": {"ior": 1.5},
                            "KHR_materials_emissive_strength": {"emissiveStrength": 40.0}}},
        ],
        "textures": [{"source": 0}],
        "images": [{"bufferView": 4, "mimeType": "image/png"}],
        "cameras": [{"type": "perspective", "perspective": {"yfov": 0.8, "znear": 0.1}}],
        "accessors": acc,
        "bufferViews": views,
        "buffers": [{"byteLength": len(blob)}],
    }
    if not binary:
        doc["buffers"][0]["uri"] = "data:application/octet-stream;base64," + base64.b64encode(blob).decode()
    return doc, blob, (P, N, U, I)


def _write(tmp_path, binary):
    px = np.arange(4 * 3 * 3, dtype=np.uint8).reshape(4, 3, 3) * 7
    png = _encode_png(px, 2, [4, 1])
    doc, blob, arrays = _scene_doc(png, binary)
    if binary:
        js = json.dumps(doc).encode()
        js += b" " * ((4 - len(js) % 4) % 4)
        body = struct.pack("<II", len(js), 0x4E4F534A) + js + struct.pack("<II", len(blob), 0x004E4942) + blob
        path = tmp_path / "scene.glb"
        path.write_bytes(struct.pack("<4sII", b"glTF", 2, 12 + len(body)) + body)
    else:
        path = tmp_path / "scene.gltf"
        path.write_text(json.dumps(doc))
    return path, px, arrays


@pytest.mark.parametrize("binary", [False, True])
def test_import_whole_file(tmp_path, binary):
    path, px, (P, N, U, I) = _write(tmp_path, binary)
    sc, cams = gltf.load_gltf(str(path))
    # mesh 0: both triangles; mesh 1: the accessor at byte offset 6 (indices 2, 1, 3)
    assert list(sc.mesh_first) == [0, 2] and list(sc.mesh_count) == [2, 1]
    assert np.array_equal(sc.positions[:2], P[I.astype(np.int64).reshape(2, 3)])
    assert np.array_equal(sc.positions[2], P[[2, 1, 3]])
    assert np.array_equal(sc.uvs[:2], U[I.astype(np.int64).reshape(2, 3)])
    assert np.array_equal(sc.normals[2], N[[2, 1, 3]])
    assert list(sc.material_ids) == [1, 1, 0]
    assert np.allclose(_apply(sc.mesh_model[0].reshape(4, 4), (1, 0, 0)), (1, 2, 1), atol=1e-5)
    assert np.array_equal(sc.mesh_model[1].reshape(4, 4)[3], np.array([5, 0, 0, 1], np.float32))
    m0, m1 = sc.materials
    assert m0["albedo"] == pytest.approx((0.2, 0.3, 0.4)) and m0["roughness"] == pytest.approx(0.7)
    assert m0["albedo_texture_id"] == -1 and m0["ior"] == pytest.approx(1.45)
    assert m1["albedo_texture_id"] == 0 and m1["normal_texture_id"] == 0 and m1["normal_scale"] == 0.5
    assert m1["transmission"] == 0.25 and m1["ior"] == 1.5 and m1["emission_scale_nits"] == 40.0
    assert m1["emission"] == (1.0, 0.5, 0.25) and m1["metallic"] == 1.0
    assert sc.textures == [(3, 4, 3, 0)]
    assert np.array_equal(sc.pixels.reshape(4, 3, 3), px)
    assert len(cams) == 1 and cams[0].yfov == pytest.approx(0.8)
    assert np.allclose(_apply(cams[0].view.reshape(4, 4), (0, 1, 5)), (0, 0, 0), atol=1e-6)
