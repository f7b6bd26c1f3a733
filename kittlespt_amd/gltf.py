"""glTF 2.0 scenes into a `scenes.Scene` (the SampleApp importer,
model_importer.hpp:51-456, feeding Renderer::loadScene).

What is read, as the reference reads it:

* the first scene's top-level nodes (model_importer.hpp:83-100; children are
  not visited, parseNode returns at once, :421-443); a node with a mesh
  becomes one mesh whose model matrix is the node's `matrix`, else
  T * R * S from `translation`, `rotation` (glm quaternion, w first) and
  `scale` (:155-198); a node with a camera becomes a camera whose view is
  the inverse of T * R(normalised quaternion) (or of `matrix`) with the
  camera's `yfov` (:380-419);
* every primitive's triangles expanded through its indices: POSITION,
  NORMAL, TEXCOORD_0 per corner, the primitive's material per triangle
  (:225-274);
* images decoded to 3 channels (RGB; alpha dropped, grey replicated, as
  stbi_load(..., 3) does, :276-318), from buffer views (.glb) or files next to
  the .gltf (the reference looks in a fixed "../models/" directory instead);
* materials: base colour factor (RGB) and texture, metallic-roughness factors
  and texture (ORM), normal texture and scale, emissive factor and texture,
  KHR_materials_transmission, KHR_materials_ior (default 1.45) and
  KHR_materials_emissive_strength (emission_scale_nits) (:320-377).

Defined where the reference is not: accessors are read as the glTF
specification lays them out (accessor and buffer-view offsets, byte stride,
8/16/32-bit indices; the reference assumes tightly packed attributes in the
index buffer's buffer and 16-bit indices, which is what such files hold), and
a primitive without indices is read in vertex order.  PNG images are decoded
here (all five row filters, 8-bit grey/RGB/palette/grey-alpha/RGBA); other
image formats raise ValueError.  The camera view is inverted in double
precision and rounded to float.
"""
from __future__ import annotations

import base64
import json
import math
import os
import struct
import ctypes
import zlib
from dataclasses import dataclass

import numpy as np

from . import abi, scenes

_CTYPES = {5120: np.int8, 5121: np.uint8, 5122: np.int16, 5123: np.uint16, 5125: np.uint32, 5126: np.float32}
_NCOMP = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT4": 16}


@dataclass
class Camera:
    name: str
    view: np.ndarray      # [16] float32, column-major (glm)
    yfov: float


# ---------------------------------------------------------------- PNG
def decode_png(data: bytes) -> np.ndarray:
    """8-bit PNG -> [H, W, 3] uint8 (stbi_load with 3 requested channels)."""
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError("not a PNG image")
    pos, idat, plte = 8, [], None
    w = h = depth = ctype = interlace = None
    while pos + 8 <= len(data):
        (n,) = struct.unpack(">I", data[pos:pos + 4])
        tag, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        pos += 12 + n
        if tag == b"IHDR":
            w, h, depth, ctype, _, _, interlace = struct.unpack(">IIBBBBB", body)
        elif tag == b"PLTE":
            plte = np.frombuffer(body, np.uint8).reshape(-1, 3)
        elif tag == b"IDAT":
            idat.append(body)
        elif tag == b"IEND":
            break
    if w is None or depth != 8 or interlace != 0:
        raise ValueError("only 8-bit non-interlaced PNG images are read")
    chans = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}.get(ctype)
    if chans is None:
        raise ValueError(f"PNG colour type {ctype} not read")
    stride = w * chans
    raw = np.frombuffer(zlib.decompress(b"".join(idat)), np.uint8)
    if raw.size != h * (stride + 1):
        raise ValueError("PNG data size does not match its header")
    out = np.empty((h, stride), np.uint8)
    L = abi.lib()
    f = L.kpt_png_unfilter
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p]
    if f(raw.ctypes.data_as(ctypes.c_void_p), h, stride, chans, out.ctypes.data_as(ctypes.c_void_p)) != abi.KPT_OK:
        raise ValueError("PNG row filter unknown")
    px = out.reshape(h, w, chans)
    if ctype == 0:
        rgb = np.repeat(px, 3, axis=2)
    elif ctype == 2:
        rgb = px
    elif ctype == 3:
        if plte is None:
            raise ValueError("palette PNG without PLTE")
        rgb = plte[px[..., 0]]
    elif ctype == 4:
        rgb = np.repeat(px[..., :1], 3, axis=2)
    else:
        rgb = px[..., :3]
    return np.ascontiguousarray(rgb)


# ---------------------------------------------------------------- matrices (glm, column-major)
def _f(x):
    return np.float32(x)


def quat_to_mat4(x, y, z, w):
    """glm::toMat4(glm::quat(w, x, y, z)) (mat3_cast) in float32, column-major."""
    x, y, z, w = _f(x), _f(y), _f(z), _f(w)
    qxx, qyy, qzz = x * x, y * y, z * z
    qxz, qxy, qyz = x * z, x * y, y * z
    qwx, qwy, qwz = w * x, w * y, w * z
    one, two = _f(1), _f(2)
    m = np.eye(4, dtype=np.float32)          # m[col][row]
    m[0, 0] = one - two * (qyy + qzz)
    m[0, 1] = two * (qxy + qwz)
    m[0, 2] = two * (qxz - qwy)
    m[1, 0] = two * (qxy - qwz)
    m[1, 1] = one - two * (qxx + qzz)
    m[1, 2] = two * (qyz + qwx)
    m[2, 0] = two * (qxz + qwy)
    m[2, 1] = two * (qyz - qwx)
    m[2, 2] = one - two * (qxx + qyy)
    return m


def mat4_mul(a, b):
    """glm mat4 * mat4 in float32: result column c = a[0]*b[c][0] + a[1]*b[c][1]
    + a[2]*b[c][2] + a[3]*b[c][3], summed left to right (arrays are [col][row])."""
    r = np.zeros((4, 4), np.float32)
    for c in range(4):
        acc = a[0] * b[c, 0]
        acc = acc + a[1] * b[c, 1]
        acc = acc + a[2] * b[c, 2]
        acc = acc + a[3] * b[c, 3]
        r[c] = acc
    return r


def node_model(node) -> np.ndarray:
    """A mesh node's model matrix (model_importer.hpp:155-198), [col][row]."""
    if node.get("matrix"):
        return np.array(node["matrix"], np.float32).reshape(4, 4)
    s = np.eye(4, dtype=np.float32)
    r = np.eye(4, dtype=np.float32)
    t = np.eye(4, dtype=np.float32)
    if node.get("scale"):
        sc = [_f(v) for v in node["scale"]]
        for k in range(3):
            s[k, k] = sc[k]
    if node.get("rotation"):
        q = node["rotation"]
        r = quat_to_mat4(q[0], q[1], q[2], q[3])
    if node.get("translation"):
        tr = node["translation"]
        t[3, :3] = [_f(tr[0]), _f(tr[1]), _f(tr[2])]
    return mat4_mul(mat4_mul(t, r), s)


def camera_view(node) -> np.ndarray:
    """inverse(T * R(normalised quaternion)) or inverse(matrix)
    (model_importer.hpp:380-410), [col][row] float32 (inverted in double)."""
    if node.get("matrix"):
        model = np.array(node["matrix"], np.float32).reshape(4, 4)
    else:
        r = np.eye(4, dtype=np.float32)
        t = np.eye(4, dtype=np.float32)
        if node.get("rotation"):
            x, y, z, w = (np.float32(v) for v in node["rotation"])
            ln = np.float32(math.sqrt(float(((x * x + y * y) + z * z) + w * w)))   # glm::normalize(quat)
            if ln > 0:
                inv = np.float32(1) / ln
                x, y, z, w = x * inv, y * inv, z * inv, w * inv
            r = quat_to_mat4(x, y, z, w)
        if node.get("translation"):
            tr = node["translation"]
            t[3, :3] = [_f(tr[0]), _f(tr[1]), _f(tr[2])]
        model = mat4_mul(t, r)
    # [col][row] storage is the transpose of the row-major matrix
    inv = np.linalg.inv(model.astype(np.float64).T).T
    return inv.astype(np.float32)


# ---------------------------------------------------------------- container
def _load(path):
    with open(path, "rb") as f:
        data = f.read()
    base = os.path.dirname(os.path.abspath(path))
    if data[:4] == b"glTF":
        magic, version, length = struct.unpack("<4sII", data[:12])
        if version != 2:
            raise ValueError(f"glTF binary version {version} not read")
        pos, doc, bin_chunk = 12, None, None
        while pos + 8 <= min(length, len(data)):
            clen, ctype = struct.unpack("<II", data[pos:pos + 8])
            body = data[pos + 8:pos + 8 + clen]
            if ctype == 0x4E4F534A:
                doc = json.loads(body.decode("utf-8"))
            elif ctype == 0x004E4942:
                bin_chunk = body
            pos += 8 + clen + ((4 - clen % 4) % 4)
        if doc is None:
            raise ValueError("glTF binary without a JSON chunk")
    else:
        doc, bin_chunk = json.loads(data.decode("utf-8")), None
    buffers = []
    for i, b in enumerate(doc.get("buffers", [])):
        uri = b.get("uri")
        if uri is None:
            if bin_chunk is None:
                raise ValueError(f"buffer {i} has no data")
            buffers.append(bin_chunk)
        elif uri.startswith("data:"):
            buffers.append(base64.b64decode(uri.split(",", 1)[1]))
        else:
            with open(os.path.join(base, uri), "rb") as f:
                buffers.append(f.read())
    return doc, buffers, base


def _view_bytes(doc, buffers, vi):
    v = doc["bufferViews"][vi]
    buf = buffers[v["buffer"]]
    off = v.get("byteOffset", 0)
    return buf[off:off + v["byteLength"]], v.get("byteStride")


def _accessor(doc, buffers, ai):
    a = doc["accessors"][ai]
    dt = np.dtype(_CTYPES[a["componentType"]])
    nc = _NCOMP[a["type"]]
    count = a["count"]
    if "bufferView" not in a:
        return np.zeros((count, nc), dt)
    raw, stride = _view_bytes(doc, buffers, a["bufferView"])
    off = a.get("byteOffset", 0)
    elem = dt.itemsize * nc
    stride = stride or elem
    if stride == elem:
        out = np.frombuffer(raw, dt, count * nc, off).reshape(count, nc)
    else:
        out = np.stack([np.frombuffer(raw, dt, nc, off + i * stride) for i in range(count)])
    return out.copy()


def load_gltf(path):
    """Returns (scene, cameras).  The scene holds one mesh per mesh node of the
    first scene (no instancing), its materials and RGB textures."""
    doc, buffers, base = _load(path)
    # images -> RGB textures
    tex_list, pix_parts, off = [], [], 0
    for i, im in enumerate(doc.get("images", [])):
        if "bufferView" in im:
            data, _ = _view_bytes(doc, buffers, im["bufferView"])
        elif im.get("uri", "").startswith("data:"):
            data = base64.b64decode(im["uri"].split(",", 1)[1])
        else:
            with open(os.path.join(base, im["uri"]), "rb") as f:
                data = f.read()
        if not bytes(data[:8]) == b"\x89PNG\r\n\x1a\n":
            raise ValueError(f"image {i}: only PNG images are decoded here")
        rgb = decode_png(bytes(data))
        h, w = rgb.shape[:2]
        tex_list.append((w, h, 3, off))
        pix_parts.append(rgb.reshape(-1))
        off += rgb.size
    texs = doc.get("textures", [])

    def tex_source(ref):
        if not ref or ref.get("index", -1) < 0:
            return -1
        return int(texs[ref["index"]].get("source", -1))

    mats = []
    for m in doc.get("materials", []):
        pbr = m.get("pbrMetallicRoughness", {})
        bc = pbr.get("baseColorFactor", [1.0, 1.0, 1.0, 1.0])
        ext = m.get("extensions", {})
        tr = ext.get("KHR_materials_transmission", {})
        mats.append(scenes.material(
            albedo=(float(bc[0]), float(bc[1]), float(bc[2])),
            albedo_texture_id=tex_source(pbr.get("baseColorTexture")),
            orm_texture_id=tex_source(pbr.get("metallicRoughnessTexture")),
            metallic=float(pbr.get("metallicFactor", 1.0)),
            roughness=float(pbr.get("roughnessFactor", 1.0)),
            transmission=float(tr.get("transmissionFactor", 0.0)),
            transmission_texture_id=tex_source(tr.get("transmissionTexture")),
            ior=float(ext.get("KHR_materials_ior", {}).get("ior", 1.45)),
            emission=tuple(float(x) for x in m.get("emissiveFactor", [0.0, 0.0, 0.0])),
            emission_texture_id=tex_source(m.get("emissiveTexture")),
            emission_scale_nits=float(ext.get("KHR_materials_emissive_strength", {}).get("emissiveStrength", 0.0)),
            normal_texture_id=tex_source(m.get("normalTexture")),
            normal_scale=float(m.get("normalTexture", {}).get("scale", 1.0))))
    pos_l, nrm_l, uv_l, mid_l, first, count, models, cams = [], [], [], [], [], [], [], []
    tri = 0
    scene_nodes = doc["scenes"][0].get("nodes", []) if doc.get("scenes") else []   # scene 0, as the reference
    for ni in scene_nodes:
        node = doc["nodes"][ni]
        if node.get("camera", -1) >= 0:
            c = doc["cameras"][node["camera"]]
            cams.append(Camera(c.get("name", ""), camera_view(node).reshape(16),
                               float(c.get("perspective", {}).get("yfov", math.radians(45.0)))))
        if node.get("mesh", -1) < 0:
            continue
        mesh = doc["meshes"][node["mesh"]]
        n0 = tri
        for prim in mesh["primitives"]:
            if prim.get("mode", 4) != 4:
                raise ValueError("only triangle primitives are read")
            at = prim["attributes"]
            P = _accessor(doc, buffers, at["POSITION"]).astype(np.float32)
            N = _accessor(doc, buffers, at["NORMAL"]).astype(np.float32)
            U = _accessor(doc, buffers, at["TEXCOORD_0"]).astype(np.float32) if "TEXCOORD_0" in at \
                else np.zeros((len(P), 2), np.float32)
            idx = _accessor(doc, buffers, prim["indices"]).reshape(-1).astype(np.int64) if "indices" in prim \
                else np.arange(len(P), dtype=np.int64)
            nt = len(idx) // 3
            idx = idx[:3 * nt].reshape(nt, 3)
            pos_l.append(P[idx])
            nrm_l.append(N[idx])
            uv_l.append(U[idx])
            mid_l.append(np.full(nt, int(prim.get("material", 0)), np.int32))
            tri += nt
        first.append(n0)
        count.append(tri - n0)
        models.append(node_model(node).reshape(16))
    if not first:
        raise ValueError("the first scene holds no mesh node")
    if not mats:
        mats = [scenes.material()]
    sc = scenes.Scene(
        positions=np.concatenate(pos_l).astype(np.float32),
        normals=np.concatenate(nrm_l).astype(np.float32),
        uvs=np.concatenate(uv_l).astype(np.float32),
        material_ids=np.concatenate(mid_l).astype(np.int32),
        mesh_first=np.array(first, np.int32),
        mesh_count=np.array(count, np.int32),
        mesh_model=np.array(models, np.float32),
        materials=mats,
        textures=tex_list,
        pixels=np.concatenate(pix_parts) if pix_parts else np.zeros(1, np.uint8))
    return sc, cams
