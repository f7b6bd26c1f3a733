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


# ---------------------------------------------------------------- C++ importer and C PNG decoder
import ctypes  # noqa: E402
import os  # noqa: E402
import subprocess  # noqa: E402

from kittlespt_amd import abi, build  # noqa: E402

_HERE = os.path.dirname(os.path.abspath(__file__))
_APP_SRC = os.path.join(_HERE, "cpp", "gltf_import.cpp")
_APP = os.path.join(_HERE, "cpp", "_build", "gltf_import")


def _c_decode(png):
    L = abi.lib()
    w, h = ctypes.c_int32(), ctypes.c_int32()
    buf = np.frombuffer(png, np.uint8)
    rc = L.kpt_png_decode_rgb8(buf.ctypes.data, len(png), ctypes.byref(w), ctypes.byref(h), None, 0)
    if rc != abi.KPT_OK:
        return rc, None
    out = np.zeros((h.value, w.value, 3), np.uint8)
    rc = L.kpt_png_decode_rgb8(buf.ctypes.data, len(png), ctypes.byref(w), ctypes.byref(h), out.ctypes.data, out.size)
    return rc, out


@pytest.mark.parametrize("ctype,chans", [(2, 3), (6, 4), (0, 1), (4, 2)])
def test_c_png_decoder_equals_python(ctype, chans):
    rng = np.random.default_rng(10 + ctype)
    px = rng.integers(0, 256, (9, 13, chans), dtype=np.uint8)
    png = _encode_png(px, ctype, [0, 1, 2, 3, 4, 4, 3, 1])
    rc, got = _c_decode(png)
    assert rc == abi.KPT_OK and np.array_equal(got, gltf.decode_png(png))


def test_c_png_decoder_palette_and_errors():
    pal = np.array([[255, 0, 0], [0, 255, 0], [0, 0, 255]], np.uint8)
    idx = np.array([[0, 1, 2], [2, 1, 0]], np.uint8)[..., None]
    png = _encode_png(idx, 3, [1, 2], palette=pal)
    rc, got = _c_decode(png)
    assert rc == abi.KPT_OK and np.array_equal(got, pal[idx[..., 0]])
    bad_index = _encode_png(np.full((2, 2, 1), 7, np.uint8), 3, [0], palette=pal)   # index past the palette
    assert _c_decode(bad_index)[0] == abi.KPT_ERR_INVALID_ARG
    assert _c_decode(b"\xff\xd8\xff\xe0JFIF")[0] == abi.KPT_ERR_INVALID_ARG
    assert _c_decode(png[:40])[0] == abi.KPT_ERR_INVALID_ARG                     # IDAT cut short
    sixteen = bytearray(png)
    sixteen[24] = 16                                                              # IHDR bit depth
    assert _c_decode(bytes(sixteen))[0] == abi.KPT_ERR_UNSUPPORTED


def _app():
    deps = [_APP_SRC, abi.LIB_PATH] + [os.path.join(build.INCLUDE, "kittlespt", f) for f in
                                       ("model_importer.hpp", "pod_types.hpp")] + [os.path.join(build.INCLUDE,
                                                                                                "kpt_abi.h")]
    if not os.path.exists(_APP) or os.path.getmtime(_APP) < max(os.path.getmtime(d) for d in deps):
        build.compile_cpp_app([_APP_SRC], _APP)
    return _APP


def _cpp_import(path, tmp_path):
    dump = tmp_path / "dump.bin"
    r = subprocess.run([_app(), str(path), str(dump)], capture_output=True, text=True, timeout=60)
    if r.returncode != 0:
        return None, r.stdout
    b = dump.read_bytes()
    pos = 0

    def take(dt, n):
        nonlocal pos
        a = np.frombuffer(b, dt, n, pos)
        pos += a.nbytes
        return a
    nm, nmat, ntex, ncam = take(np.int32, 4)
    meshes = []
    for _ in range(nm):
        nt = int(take(np.int32, 1)[0])
        model = take(np.float32, 16)
        tris = []
        for _ in range(nt):
            tris.append((take(np.float32, 24).copy(), int(take(np.int32, 1)[0])))
        meshes.append((model.copy(), tris))
    mats = [take(np.float32, 18).copy() for _ in range(nmat)]
    texs = []
    for _ in range(ntex):
        w, h, ch = (int(x) for x in take(np.uint32, 3))
        texs.append((w, h, ch, take(np.uint8, w * h * ch).copy()))
    cams = [(take(np.float32, 16).copy(), float(take(np.float32, 1)[0])) for _ in range(ncam)]
    assert pos == len(b)
    return (meshes, mats, texs, cams), r.stdout


def _assert_same_scene(cpp, sc, cams):
    meshes, mats, texs, ccams = cpp
    assert len(meshes) == len(sc.mesh_first)
    for k, (model, tris) in enumerate(meshes):
        assert np.array_equal(model.view(np.uint32), np.asarray(sc.mesh_model[k], np.float32).view(np.uint32))
        f, n = int(sc.mesh_first[k]), int(sc.mesh_count[k])
        assert len(tris) == n
        for j, (v, mat) in enumerate(tris):
            t = f + j
            want = np.concatenate([sc.positions[t].reshape(-1), sc.normals[t].reshape(-1),
                                   sc.uvs[t].reshape(-1)]).astype(np.float32)
            assert np.array_equal(v.view(np.uint32), want.view(np.uint32)), (k, j)
            assert mat == int(sc.material_ids[t])
    assert len(mats) == len(sc.materials)
    for got, m in zip(mats, sc.materials):
        want = np.array([m["albedo_texture_id"], *m["albedo"], m["orm_texture_id"], m["metallic"], m["roughness"],
                         m["transmission_texture_id"], m["transmission"], m["ior"], m["emission_texture_id"],
                         *m["emission"], m["emission_scale_nits"], m["normal_texture_id"], m["normal_scale"], 0.0],
                        np.float32)
        assert np.array_equal(got, want)
    assert [(w, h, ch) for w, h, ch, _ in texs] == [(w, h, ch) for w, h, ch, _ in sc.textures]
    for (w, h, ch, px), (_, _, _, off) in zip(texs, sc.textures):
        assert np.array_equal(px, sc.pixels[off:off + w * h * ch])
    assert len(ccams) == len(cams)
    for (view, yfov), c in zip(ccams, cams):
        assert yfov == np.float32(c.yfov)
        # both invert in double and round to float: equal, or one float ulp apart
        assert np.allclose(view, c.view, rtol=2e-7, atol=1e-7)


@pytest.mark.parametrize("binary", [False, True])
def test_cpp_importer_equals_python(tmp_path, binary):
    path, _, _ = _write(tmp_path, binary)
    cpp, log = _cpp_import(path, tmp_path)
    assert cpp is not None, log
    sc, cams = gltf.load_gltf(str(path))
    _assert_same_scene(cpp, sc, cams)


def _rich_files(tmp_path):
    """A .gltf with an external buffer and external images: interleaved
    position/normal (byte stride 24), 8- and 32-bit indices with accessor
    offsets, a primitive without indices, texture coordinates or material,
    a palette and a grey-alpha PNG, a camera with a non-normalised
    quaternion and a mesh node with a full TRS."""
    rng = np.random.default_rng(5)
    nv = 11
    P = rng.uniform(-3, 3, (nv, 3)).astype(np.float32)
    N = rng.normal(size=(nv, 3)).astype(np.float32)
    U = rng.uniform(0, 1, (nv, 2)).astype(np.float32)
    inter = np.concatenate([P, N], axis=1).astype(np.float32).tobytes()        # stride 24
    i8 = np.array([0, 1, 2, 3, 4, 5, 6, 7, 8], np.uint8).tobytes() + b"\0" * 3
    i32 = np.array([9, 9, 9, 10, 2, 4, 1, 3, 5, 8, 7, 6], np.uint32).tobytes()
    blob = inter + U.tobytes() + i8 + i32
    o_u = len(inter)
    o_i8 = o_u + U.nbytes
    o_i32 = o_i8 + len(i8)
    (tmp_path / "geo.bin").write_bytes(blob)
    pal = rng.integers(0, 256, (6, 3), dtype=np.uint8)
    (tmp_path / "pal.png").write_bytes(_encode_png(rng.integers(0, 6, (5, 4, 1), dtype=np.uint8), 3, [4, 1, 0],
                                                   palette=pal))
    ga = _encode_png(rng.integers(0, 256, (3, 6, 2), dtype=np.uint8), 4, [3, 2])
    doc = {
        "asset": {"version": "2.0"},
        "scenes": [{"nodes": [0, 1, 2]}],
        "nodes": [
            {"mesh": 0, "translation": [0.5, -2, 7], "rotation": [0.1, 0.7, -0.2, 0.68], "scale": [1, 2, 0.5]},
            {"camera": 0, "translation": [3, 1, -4], "rotation": [0.3, 0.2, 0.1, 2.0]},
            {"mesh": 1},
        ],
        "meshes": [
            {"name": "a", "primitives": [
                {"attributes": {"POSITION": 0, "NORMAL": 1, "TEXCOORD_0": 2}, "indices": 3, "material": 1},
                {"attributes": {"POSITION": 0, "NORMAL": 1}, "indices": 4}]},
            {"name": "b", "primitives": [{"attributes": {"POSITION": 0, "NORMAL": 1, "TEXCOORD_0": 2}}]},
        ],
        "materials": [
            {"name": "m0"},
            {"name": "m1", "pbrMetallicRoughness": {"baseColorTexture": {"index": 1},
                                                    "metallicRoughnessTexture": {"index": 0}},
             "emissiveTexture": {"index": 1},
             "extensions": {"KHR_materials_transmission": {"transmissionFactor": 0.75,
                                                           "transmissionTexture": {"index": 0}}}},
        ],
        "textures": [{"source": 0}, {"source": 1}],
        "images": [{"uri": "pal.png", "name": "pal"},
                   {"uri": "data:image/png;base64," + base64.b64encode(ga).decode(), "name": "ga"}],
        "cameras": [{"name": "cam", "type": "perspective", "perspective": {"yfov": 0.6}}],
        "accessors": [
            {"bufferView": 0, "componentType": 5126, "count": nv, "type": "VEC3"},
            {"bufferView": 0, "byteOffset": 12, "componentType": 5126, "count": nv, "type": "VEC3"},
            {"bufferView": 1, "componentType": 5126, "count": nv, "type": "VEC2"},
            {"bufferView": 2, "componentType": 5121, "count": 9, "type": "SCALAR"},
            {"bufferView": 3, "byteOffset": 12, "componentType": 5125, "count": 9, "type": "SCALAR"},
        ],
        "bufferViews": [
            {"buffer": 0, "byteOffset": 0, "byteLength": len(inter), "byteStride": 24},
            {"buffer": 0, "byteOffset": o_u, "byteLength": U.nbytes},
            {"buffer": 0, "byteOffset": o_i8, "byteLength": 9},
            {"buffer": 0, "byteOffset": o_i32, "byteLength": len(i32)},
        ],
        "buffers": [{"uri": "geo.bin", "byteLength": len(blob)}],
    }
    path = tmp_path / "rich.gltf"
    path.write_text(json.dumps(doc, indent=1))
    return path, doc


def test_cpp_importer_equals_python_on_layouts_and_defaults(tmp_path):
    path, _ = _rich_files(tmp_path)
    sc, cams = gltf.load_gltf(str(path))
    assert list(sc.mesh_count) == [6, 3] and len(sc.textures) == 2 and len(cams) == 1
    cpp, log = _cpp_import(path, tmp_path)
    assert cpp is not None, log
    _assert_same_scene(cpp, sc, cams)


def test_cpp_importer_reports_malformed_files(tmp_path):
    path, doc = _rich_files(tmp_path)
    cases = {
        "truncated.glb": None,
        "badjson.gltf": '{"asset": {"version": "2.0"}, "scenes": [',
        "nomesh.gltf": json.dumps({**doc, "scenes": [{"nodes": [1]}]}),
        "oob.gltf": json.dumps({**doc, "accessors": doc["accessors"][:3] + [
            {"bufferView": 2, "byteOffset": 8, "componentType": 5121, "count": 9, "type": "SCALAR"}] +
            doc["accessors"][4:]}),
    }
    glb_path, _, _ = _write(tmp_path, True)
    for name, text in cases.items():
        p = tmp_path / name
        if text is None:
            p.write_bytes(glb_path.read_bytes()[:200])
        else:
            p.write_text(text)
        cpp, log = _cpp_import(p, tmp_path)
        assert cpp is None and log.startswith("import failed: "), (name, log)
        with pytest.raises(Exception):
            gltf.load_gltf(str(p))


# ---------------------------------------------------------------- on-disk scenes
def test_scene_file_round_trip(tmp_path):
    """save_scene / load_scene: every array bit for bit, materials, texture
    descriptors, camera and sun; instanced and plain scenes; a file of
    another kind is refused."""
    from kittlespt_amd import scenes
    for sc in (scenes.scene_instanced_small(), scenes.scene_plain(subdiv=2), gltf.load_gltf(str(_write(tmp_path,
                                                                                                          True)[0]))[0]):
        p = tmp_path / "s.npz"
        scenes.save_scene(sc, str(p))
        back = scenes.load_scene(str(p))
        for k in ("positions", "normals", "uvs", "material_ids", "mesh_first", "mesh_count", "mesh_model", "pixels"):
            a, b = np.asarray(getattr(sc, k)), getattr(back, k)
            assert a.shape == b.shape and np.array_equal(a.view(np.uint8) if a.dtype.kind == "f" else a,
                                                         b.view(np.uint8) if b.dtype.kind == "f" else b), k
        assert (sc.instance_mesh is None) == (back.instance_mesh is None)
        if sc.instance_mesh is not None:
            assert np.array_equal(sc.instance_mesh, back.instance_mesh)
            assert np.array_equal(np.asarray(sc.instance_model, np.float32).view(np.uint32),
                                  back.instance_model.view(np.uint32))
        assert back.materials == [{k: (tuple(v) if isinstance(v, list) else v) for k, v in m.items()}
                                  for m in sc.materials]
        assert back.textures == [tuple(int(x) for x in t) for t in sc.textures]
        assert back.camera_eye == tuple(sc.camera_eye) and back.sun_theta == sc.sun_theta
    bad = tmp_path / "bad.npz"
    np.savez(bad, positions=np.zeros((1, 3, 3), np.float32))
    with pytest.raises(ValueError):
        scenes.load_scene(str(bad))
