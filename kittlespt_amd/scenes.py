"""Seeded synthetic scenes and camera helpers for the benchmark configurations.

All geometry is generated with numpy's PCG64 generator from literal seeds, so a
scene is identical on every machine.  A scene is held as flat arrays (the
layout kpt_scene_desc takes) rather than per-triangle objects.

Configurations (sizes as stated in the benchmark definition):
  * ``scene_plain``     — 256x256 case: one icosphere BLAS (subdivision 5,
    20,480 triangles, radius 1 at (0,1,0)) and a 2-triangle ground quad;
    diffuse-dielectric ground (albedo 0.8), GGX conductor (roughness 0.3) and
    glass (IOR 1.5) sphere halves; Nishita sky with the sun 30 deg high.
  * ``scene_instanced`` — 1080p case: 8 seeded BLAS (4 displaced spheres,
    4 random triangle soups, 50k-200k triangles each) placed as 1,000
    instances (uniform rotation, translation in a 100^3 box, scale U[0.5,4]);
    16 materials (roughness U[0.02,1], metallic Bernoulli(0.4),
    IOR U[1.3,1.8]) with 1024^2 RGBA8 value-noise albedo textures.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import abi


def identity16():
    return np.eye(4, dtype=np.float32).reshape(16).copy()


def mat_colmajor(m4):
    """4x4 numpy (row, col) -> 16 floats column-major (glm layout)."""
    return np.ascontiguousarray(np.asarray(m4, dtype=np.float64).T.reshape(16), dtype=np.float32)


def perspective_fov_lh(fov, width, height, znear, zfar):
    """glm::perspectiveFovLH with the default [-1,1] depth range
    (vendored glm ext/matrix_clip_space.inl:412-429), float32 arithmetic."""
    f = np.float32
    rad = f(fov)
    h = f(np.cos(f(0.5) * rad)) / f(np.sin(f(0.5) * rad))
    w = f(h * f(height) / f(width))
    m = np.zeros((4, 4), dtype=np.float32)   # m[col][row] as glm stores it
    m[0][0] = w
    m[1][1] = h
    m[2][2] = (f(zfar) + f(znear)) / (f(zfar) - f(znear))
    m[2][3] = f(1)
    m[3][2] = -(f(2) * f(zfar) * f(znear)) / (f(zfar) - f(znear))
    return np.ascontiguousarray(m.reshape(16))


def look_at_view(eye, target, up=(0.0, 1.0, 0.0)):
    """View matrix as CameraController::getViewMatrix builds it
    (SampleApp camera_controller.cpp:85-97): inverse of the camera-to-world
    matrix with columns (right, up, forward, position)."""
    eye = np.asarray(eye, np.float64)
    fwd = np.asarray(target, np.float64) - eye
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, np.asarray(up, np.float64))
    right /= np.linalg.norm(right)
    upv = np.cross(right, fwd)
    upv /= np.linalg.norm(upv)
    inv_view = np.eye(4)
    inv_view[:3, 0] = right
    inv_view[:3, 1] = upv
    inv_view[:3, 2] = fwd
    inv_view[:3, 3] = eye
    return mat_colmajor(np.linalg.inv(inv_view))


def trs(translate=(0, 0, 0), rotate=None, scale=1.0):
    """Model matrix T * R * S as 16 column-major floats."""
    m = np.eye(4)
    r = np.eye(3) if rotate is None else np.asarray(rotate, np.float64)
    m[:3, :3] = r * scale
    m[:3, 3] = translate
    return mat_colmajor(m)


def quat_to_mat(q):
    x, y, z, w = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def random_rotation(rng):
    """Uniform rotation (Shoemake's subgroup algorithm)."""
    u1, u2, u3 = rng.random(3)
    a, b = math.sqrt(1 - u1), math.sqrt(u1)
    q = (a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2),
         b * math.sin(2 * math.pi * u3), b * math.cos(2 * math.pi * u3))
    return quat_to_mat(q)


# ----------------------------------------------------------------- geometry
def icosphere(subdiv):
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = np.array(v, dtype=np.float64)
    verts /= np.linalg.norm(verts, axis=1, keepdims=True)
    faces = np.array(f, dtype=np.int64)
    for _ in range(subdiv):
        e = np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]])
        e.sort(axis=1)
        uniq, inv = np.unique(e, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mid = verts[uniq[:, 0]] + verts[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = len(verts)
        verts = np.concatenate([verts, mid])
        n = len(faces)
        a, b, c = inv[:n] + base, inv[n:2 * n] + base, inv[2 * n:] + base
        faces = np.concatenate([
            np.stack([faces[:, 0], a, c], 1), np.stack([faces[:, 1], b, a], 1),
            np.stack([faces[:, 2], c, b], 1), np.stack([a, b, c], 1)])
    return verts, faces


def sphere_uv(d):
    u = 0.5 + np.arctan2(d[..., 2], d[..., 0]) / (2 * np.pi)
    v = 0.5 - np.arcsin(np.clip(d[..., 1], -1, 1)) / np.pi
    return np.stack([u, v], -1)


def uv_sphere(n_lat, n_lon):
    """Latitude-longitude sphere without degenerate pole triangles."""
    lat = np.linspace(0, np.pi, n_lat + 1)[1:-1]
    lon = np.linspace(0, 2 * np.pi, n_lon + 1)[:-1]
    th, ph = np.meshgrid(lat, lon, indexing="ij")
    ring = np.stack([np.sin(th) * np.cos(ph), np.cos(th), np.sin(th) * np.sin(ph)], -1).reshape(-1, 3)
    verts = np.concatenate([[[0, 1, 0]], ring, [[0, -1, 0]]])
    south = len(verts) - 1
    faces = []
    j = np.arange(n_lon)
    jn = (j + 1) % n_lon
    faces.append(np.stack([np.zeros(n_lon, np.int64), 1 + jn, 1 + j], 1))
    for i in range(n_lat - 2):
        a = 1 + i * n_lon + j
        b = 1 + i * n_lon + jn
        c = 1 + (i + 1) * n_lon + j
        d = 1 + (i + 1) * n_lon + jn
        faces.append(np.stack([a, b, c], 1))
        faces.append(np.stack([b, d, c], 1))
    last = 1 + (n_lat - 2) * n_lon
    faces.append(np.stack([last + j, last + jn, np.full(n_lon, south)], 1))
    return verts.astype(np.float64), np.concatenate(faces).astype(np.int64)


# ----------------------------------------------------------------- scene
@dataclass
class Scene:
    positions: np.ndarray            # [T,3,3] float32
    normals: np.ndarray              # [T,3,3] float32
    uvs: np.ndarray                  # [T,3,2] float32
    material_ids: np.ndarray         # [T] int32
    mesh_first: np.ndarray           # [M] int32
    mesh_count: np.ndarray           # [M] int32
    mesh_model: np.ndarray           # [M,16] float32
    materials: list                  # list of dict (MaterialSceneEntity fields)
    instance_mesh: np.ndarray | None = None   # [I] int32
    instance_model: np.ndarray | None = None  # [I,16] float32
    textures: list = field(default_factory=list)   # (width, height, channels, offset)
    pixels: np.ndarray = field(default_factory=lambda: np.zeros(1, np.uint8))
    camera_eye: tuple = (0.0, 0.0, -5.0)
    camera_target: tuple = (0.0, 0.0, 0.0)
    camera_fov: float = math.radians(45.0)
    sun_theta: float = math.radians(17.0)

    @property
    def num_instances(self):
        return len(self.instance_mesh) if self.instance_mesh is not None else len(self.mesh_first)

    def desc(self):
        """kpt_scene_desc over this scene's arrays; keep the returned tuple
        alive for as long as the descriptor is used."""
        keep = []

        def arr(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a

        pos = arr(self.positions, np.float32)
        nrm = arr(self.normals, np.float32)
        uvs = arr(self.uvs, np.float32)
        mids = arr(self.material_ids, np.int32)
        mf = arr(self.mesh_first, np.int32)
        mc = arr(self.mesh_count, np.int32)
        mm = arr(self.mesh_model, np.float32)
        mats = (abi.KptMaterial * max(1, len(self.materials)))()
        for i, m in enumerate(self.materials):
            mats[i] = material_struct(m)
        keep.append(mats)
        texs = (abi.KptTextureDesc * max(1, len(self.textures)))()
        for i, (w, h, ch, off) in enumerate(self.textures):
            texs[i] = abi.KptTextureDesc(w, h, ch, 0, off)
        keep.append(texs)
        pix = arr(self.pixels, np.uint8)
        d = abi.KptSceneDesc()
        d.num_triangles = len(pos)
        d.positions = abi.fptr(pos)
        d.normals = abi.fptr(nrm)
        d.uvs = abi.fptr(uvs)
        d.material_ids = abi.iptr(mids)
        d.num_meshes = len(mf)
        d.mesh_first = abi.iptr(mf)
        d.mesh_count = abi.iptr(mc)
        d.mesh_model = abi.fptr(mm)
        if self.instance_mesh is not None:
            im = arr(self.instance_mesh, np.int32)
            imod = arr(self.instance_model, np.float32)
            d.num_instances = len(im)
            d.instance_mesh = abi.iptr(im)
            d.instance_model = abi.fptr(imod)
        else:
            d.num_instances = 0
        d.num_materials = len(self.materials)
        d.materials = C.cast(mats, C.POINTER(abi.KptMaterial))
        d.num_textures = len(self.textures)
        d.textures = C.cast(texs, C.POINTER(abi.KptTextureDesc))
        d.pixels = pix.ctypes.data_as(C.POINTER(C.c_uint8))
        d.num_pixel_bytes = int(pix.size)
        return d, keep


# ---------------------------------------------------------------- on-disk scenes
SCENE_FORMAT = "kittlespt-scene/1"
_ARRAYS = ("positions", "normals", "uvs", "material_ids", "mesh_first", "mesh_count", "mesh_model", "pixels")
_DTYPES = {"positions": np.float32, "normals": np.float32, "uvs": np.float32, "material_ids": np.int32,
           "mesh_first": np.int32, "mesh_count": np.int32, "mesh_model": np.float32, "pixels": np.uint8,
           "instance_mesh": np.int32, "instance_model": np.float32}


def save_scene(scene: Scene, path: str) -> None:
    """Writes a scene as one uncompressed .npz: the flat arrays kpt_scene_desc
    takes (float32 / int32 / uint8, the layout loadScene receives) and a JSON
    header (format tag, materials, texture descriptors, camera, sun).  Loading
    needs no pickling (numpy.load with allow_pickle off)."""
    import json
    header = {"format": SCENE_FORMAT, "materials": [{k: (list(v) if isinstance(v, tuple) else v)
                                                     for k, v in m.items()} for m in scene.materials],
              "textures": [list(map(int, t)) for t in scene.textures], "camera_eye": list(scene.camera_eye),
              "camera_target": list(scene.camera_target), "camera_fov": float(scene.camera_fov),
              "sun_theta": float(scene.sun_theta), "instanced": scene.instance_mesh is not None}
    arrays = {k: np.ascontiguousarray(getattr(scene, k), _DTYPES[k]) for k in _ARRAYS}
    if scene.instance_mesh is not None:
        arrays["instance_mesh"] = np.ascontiguousarray(scene.instance_mesh, np.int32)
        arrays["instance_model"] = np.ascontiguousarray(scene.instance_model, np.float32)
    text = json.dumps(header, default=lambda o: o.item() if isinstance(o, np.generic) else o)
    arrays["header"] = np.frombuffer(text.encode("utf-8"), np.uint8)
    with open(path, "wb") as f:
        np.savez(f, **arrays)


def load_scene(path: str) -> Scene:
    """Reads a scene written by save_scene (ValueError on another format)."""
    import json
    with np.load(path, allow_pickle=False) as z:
        if "header" not in z.files:
            raise ValueError(f"{path}: not a scene file (no header)")
        header = json.loads(bytes(z["header"]).decode("utf-8"))
        if header.get("format") != SCENE_FORMAT:
            raise ValueError(f"{path}: scene format {header.get('format')!r} is not {SCENE_FORMAT!r}")
        arrays = {}
        for k in _ARRAYS:
            if k not in z.files:
                raise ValueError(f"{path}: array {k} missing")
            arrays[k] = np.asarray(z[k], _DTYPES[k])
        inst = {}
        if header.get("instanced"):
            inst = {k: np.asarray(z[k], _DTYPES[k]) for k in ("instance_mesh", "instance_model")}
    mats = [{k: (tuple(v) if isinstance(v, list) else v) for k, v in m.items()} for m in header["materials"]]
    return Scene(materials=mats, textures=[tuple(t) for t in header["textures"]],
                 camera_eye=tuple(header["camera_eye"]), camera_target=tuple(header["camera_target"]),
                 camera_fov=header["camera_fov"], sun_theta=header["sun_theta"], **arrays, **inst)


def material(albedo=(0.8, 0.8, 0.8), metallic=0.0, roughness=0.5, transmission=0.0, ior=1.45,
             albedo_texture_id=-1, emission=(1.0, 1.0, 1.0), emission_scale_nits=0.0, orm_texture_id=-1,
             transmission_texture_id=-1, emission_texture_id=-1, normal_texture_id=-1, normal_scale=1.0):
    """MaterialSceneEntity (pod_types.hpp:60-123) with its defaults."""
    return dict(albedo=albedo, metallic=metallic, roughness=roughness, transmission=transmission, ior=ior,
                albedo_texture_id=albedo_texture_id, emission=emission, emission_scale_nits=emission_scale_nits,
                orm_texture_id=orm_texture_id, transmission_texture_id=transmission_texture_id,
                emission_texture_id=emission_texture_id, normal_texture_id=normal_texture_id,
                normal_scale=normal_scale)


def material_struct(m):
    s = abi.KptMaterial()
    s.albedo_texture_id = m["albedo_texture_id"]
    s.albedo[:] = [float(x) for x in m["albedo"]]
    s.orm_texture_id = m["orm_texture_id"]
    s.metallic = m["metallic"]
    s.roughness = m["roughness"]
    s.transmission_texture_id = m["transmission_texture_id"]
    s.transmission = m["transmission"]
    s.ior = m["ior"]
    s.emission_texture_id = m["emission_texture_id"]
    s.emission_factor[:] = [float(x) for x in m["emission"]]
    s.emission_scale_nits = m["emission_scale_nits"]
    s.normal_texture_id = m["normal_texture_id"]
    s.normal_scale = m["normal_scale"]
    return s


def _pack(meshes, materials, mesh_models=None, **kw):
    """meshes: list of (positions[T,3,3], normals[T,3,3], uvs[T,3,2], mat_ids[T])."""
    counts = np.array([len(m[0]) for m in meshes], np.int32)
    first = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int32)
    if mesh_models is None:
        mesh_models = np.stack([identity16() for _ in meshes])
    return Scene(
        positions=np.concatenate([m[0] for m in meshes]).astype(np.float32),
        normals=np.concatenate([m[1] for m in meshes]).astype(np.float32),
        uvs=np.concatenate([m[2] for m in meshes]).astype(np.float32),
        material_ids=np.concatenate([m[3] for m in meshes]).astype(np.int32),
        mesh_first=first, mesh_count=counts, mesh_model=np.asarray(mesh_models, np.float32),
        materials=materials, **kw)


def ground_quad(half=20.0, mat=0):
    p = np.array([[[-half, 0, -half], [half, 0, half], [half, 0, -half]],
                  [[-half, 0, -half], [-half, 0, half], [half, 0, half]]], np.float64)
    n = np.tile(np.array([0, 1, 0], np.float64), (2, 3, 1))
    uv = (p[..., [0, 2]] + half) / (2 * half)
    return p, n, uv, np.full(2, mat, np.int32)


def scene_plain(subdiv=5):
    """The 256x256 plain-path scene (icosphere subdivision 5 + ground quad)."""
    verts, faces = icosphere(subdiv)
    tri_p = verts[faces]                    # unit sphere, object space
    tri_n = tri_p.copy()
    tri_uv = sphere_uv(tri_p)
    cen_x = tri_p.mean(axis=1)[:, 0]
    mats_sphere = np.where(cen_x < 0.0, 1, 2).astype(np.int32)
    materials = [
        material(albedo=(0.8, 0.8, 0.8), metallic=0.0, roughness=1.0, transmission=0.0, ior=1.45),
        material(albedo=(0.9, 0.9, 0.9), metallic=1.0, roughness=0.3),
        material(albedo=(1.0, 1.0, 1.0), metallic=0.0, roughness=0.02, transmission=1.0, ior=1.5),
    ]
    meshes = [ground_quad(), (tri_p, tri_n, tri_uv, mats_sphere)]
    models = np.stack([identity16(), trs(translate=(0.0, 1.0, 0.0))])
    return _pack(meshes, materials, models, camera_eye=(0.0, 1.2, -3.5), camera_target=(0.0, 1.0, 0.0),
                 camera_fov=math.radians(45.0), sun_theta=math.radians(30.0))


def value_noise_texture(rng, size=1024, cells=16, octaves=3):
    """RGBA8 value noise: bilinear-smoothstep interpolation of random lattices."""
    img = np.zeros((size, size, 3), np.float64)
    amp, total = 1.0, 0.0
    for o in range(octaves):
        c = cells * (2 ** o)
        lat = rng.random((c + 1, c + 1, 3))
        lat[-1, :] = lat[0, :]
        lat[:, -1] = lat[:, 0]
        x = np.linspace(0, c, size, endpoint=False)
        i0 = np.floor(x).astype(np.int64)
        fx = x - i0
        sx = fx * fx * (3 - 2 * fx)
        a = lat[i0][:, i0] * (1 - sx)[None, :, None] + lat[i0][:, i0 + 1] * sx[None, :, None]
        b = lat[i0 + 1][:, i0] * (1 - sx)[None, :, None] + lat[i0 + 1][:, i0 + 1] * sx[None, :, None]
        img += amp * (a * (1 - sx)[:, None, None] + b * sx[:, None, None])
        total += amp
        amp *= 0.5
    img /= total
    rgba = np.empty((size, size, 4), np.uint8)
    rgba[..., :3] = np.clip(img * 255.0 + 0.5, 0, 255).astype(np.uint8)
    rgba[..., 3] = 255
    return rgba


def displaced_sphere(rng, n_tris):
    n_lat = max(4, int(round(math.sqrt(n_tris / 4.0))))
    n_lon = max(8, int(round(n_tris / (2.0 * (n_lat - 1)))))
    verts, faces = uv_sphere(n_lat, n_lon)
    k = rng.normal(size=(6, 3)) * 3.0
    ph = rng.random(6) * 2 * np.pi
    a = rng.random(6) * 0.05
    r = 1.0 + (a[None, :] * np.sin(verts @ k.T + ph[None, :])).sum(1)
    dv = verts * r[:, None]
    tri_p = dv[faces]
    tri_n = verts[faces]
    tri_uv = sphere_uv(verts)[faces]
    return tri_p, tri_n, tri_uv


def triangle_soup(rng, n_tris, extent=0.08):
    c = rng.uniform(-1.0, 1.0, size=(n_tris, 1, 3))
    p = c + rng.uniform(-extent, extent, size=(n_tris, 3, 3))
    gn = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    gn /= np.maximum(np.linalg.norm(gn, axis=1, keepdims=True), 1e-12)
    n = np.repeat(gn[:, None, :], 3, axis=1)
    uv = rng.random((n_tris, 3, 2))
    return p, n, uv


def scene_instanced(num_instances=1000, tex_size=1024, tris_range=(50_000, 200_000), num_blas=8,
                    num_materials=16, seed=0):
    """The 1080p instanced scene (8 BLAS, 1,000 instances, 16 textured materials)."""
    meshes = []
    for b in range(num_blas):
        rng = np.random.default_rng(seed + b + 1)         # BLAS seeds 1..8
        n = int(rng.integers(tris_range[0], tris_range[1] + 1))
        if b < num_blas // 2:
            p, nr, uv = displaced_sphere(rng, n)
        else:
            p, nr, uv = triangle_soup(rng, n)
        t = np.arange(len(p))
        mids = ((2 * b + (t % 2)) % num_materials).astype(np.int32)
        meshes.append((p, nr, uv, mids))
    rng = np.random.default_rng(seed + 100)
    materials, pixels, textures = [], [], []
    off = 0
    for m in range(num_materials):
        tex = value_noise_texture(np.random.default_rng(seed + 200 + m), size=tex_size)
        textures.append((tex_size, tex_size, 4, off))
        pixels.append(tex.reshape(-1))
        off += tex.size
        materials.append(material(
            albedo=(1.0, 1.0, 1.0), roughness=float(rng.uniform(0.02, 1.0)),
            metallic=float(rng.random() < 0.4), ior=float(rng.uniform(1.3, 1.8)), transmission=0.0,
            albedo_texture_id=m))
    rng = np.random.default_rng(seed + 300)
    inst_mesh = rng.integers(0, num_blas, size=num_instances).astype(np.int32)
    models = []
    for i in range(num_instances):
        rot = random_rotation(rng)
        tr = rng.uniform(0.0, 100.0, size=3)
        sc = float(rng.uniform(0.5, 4.0))
        models.append(trs(translate=tr, rotate=rot, scale=sc))
    sc = _pack(meshes, materials, None, instance_mesh=inst_mesh, instance_model=np.stack(models),
               textures=textures, pixels=np.concatenate(pixels) if pixels else np.zeros(1, np.uint8),
               camera_eye=(50.0, 60.0, -60.0), camera_target=(50.0, 45.0, 50.0), camera_fov=math.radians(50.0),
               sun_theta=math.radians(30.0))
    return sc


def scene_instanced_small(seed=0):
    """A reduced instanced scene for tests: same generators, 40 instances of
    4 small BLAS and 64^2 textures."""
    return scene_instanced(num_instances=40, tex_size=64, tris_range=(600, 1500), num_blas=4, num_materials=8,
                           seed=seed)


def multi_view_cameras(scene, n_views=64, seed=64, lo=128, hi=1024):
    """The multi-view workload's cameras: n_views views around the instance
    box (azimuth uniform, elevation U[0.1, 0.9] rad, 110 units from its
    centre), square resolutions with side round(exp(U[ln lo, ln hi]))."""
    rng = np.random.default_rng(seed)
    views = []
    for _ in range(n_views):
        side = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        az, el = rng.uniform(0, 2 * math.pi), rng.uniform(0.1, 0.9)
        eye = (50 + 110 * math.cos(el) * math.cos(az), 45 + 110 * math.sin(el), 50 + 110 * math.cos(el) * math.sin(az))
        proj = perspective_fov_lh(scene.camera_fov, side, side, 1.0, 100.0)
        views.append((proj, look_at_view(eye, (50.0, 45.0, 50.0)), side, side))
    return views

