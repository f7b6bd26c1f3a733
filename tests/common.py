"""Shared set-up for the tests: settings, cameras and ray batches."""
import math

import numpy as np

from kittlespt_amd import abi, scenes
from kittlespt_amd.context import Context


def settings(max_depth=4, seed_mode=abi.SEED_XOR, integrator=abi.INTEGRATOR_MEGAKERNEL,
             sun_theta=math.radians(30.0)):
    s = abi.KptSettings()
    s.max_depth = max_depth
    s.use_temporal_accumulation = 0
    s.bloom_final_blend = 0.175
    s.bloom_internal_blend = 0.65
    s.sun_angular_diameter_rad = 0.0087
    s.sun_phi_rad = 4.18879
    s.sun_theta_rad = sun_theta
    s.sun_emission_factor = 1.0
    s.seed_mode = seed_mode
    s.integrator = integrator
    return s


def view_proj(scene, width, height):
    proj = scenes.perspective_fov_lh(scene.camera_fov, width, height, 1.0, 100.0)
    view = scenes.look_at_view(scene.camera_eye, scene.camera_target)
    return proj, view


def host_camera(scene, width, height):
    """kpt_camera as Renderer::setView computes it, from a host-only context."""
    c = Context(-2)
    p, v = view_proj(scene, width, height)
    c.set_view(p, v)
    cam = c.camera()
    c.close()
    return cam


def camera_rays(cam, width, height):
    """Primary rays of every pixel through the same inverse matrices
    (float64 here; only used as test inputs)."""
    ip = np.array(cam.inv_proj, np.float64).reshape(4, 4).T
    iv = np.array(cam.inv_view, np.float64).reshape(4, 4).T
    ys, xs = np.mgrid[0:height, 0:width]
    u = (xs + 0.5) / width * 2 - 1
    v = (ys + 0.5) / height * 2 - 1
    p = np.stack([u, v, np.ones_like(u), np.ones_like(u)], -1).reshape(-1, 4) @ ip.T
    d = p[:, :3] / p[:, 3:4]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    dw = np.concatenate([d, np.zeros((len(d), 1))], 1) @ iv.T
    dw = dw[:, :3] / np.linalg.norm(dw[:, :3], axis=1, keepdims=True)
    rays = np.zeros(len(dw), abi.RAY_DTYPE)
    rays["org"] = iv[:3, 3]
    rays["dir"] = dw
    rays["tmin"] = 0.0
    rays["tmax"] = np.inf
    return rays


def random_rays(n, lo, hi, target_center, target_spread, seed=0, tmax=np.inf):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3))
    tgt = rng.normal(size=(n, 3)) * target_spread + np.asarray(target_center)
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros(n, abi.RAY_DTYPE)
    rays["org"] = o
    rays["dir"] = d
    rays["tmin"] = 0.0
    rays["tmax"] = tmax
    return rays


def furnace_cube(half=5.0, albedo=(0.4, 0.25, 0.6), emission=(1.0, 0.5, 2.0), roughness=1.0, ior=1.0,
                 scale_nits=1.0):
    """A closed cube of 12 equal-area triangles facing inward, every one an
    emitter of uniform radiance `emission` * `scale_nits` and an opaque
    dielectric of albedo `albedo`.  With IOR 1 the dielectric's Fresnel term
    is 0 and its lobe is Lambertian (f = albedo / pi), so the radiance inside
    has a closed form (the furnace test).  The camera sits at the centre."""
    c = np.array([[x, y, z] for x in (-half, half) for y in (-half, half) for z in (-half, half)], np.float64)
    faces = []
    for axis in range(3):
        for sgn in (-1.0, 1.0):
            idx = [i for i in range(8) if c[i, axis] * sgn > 0]
            # quad corners in cyclic order around the face
            q = sorted(idx, key=lambda i: math.atan2(c[i, (axis + 2) % 3], c[i, (axis + 1) % 3]))
            inward = np.zeros(3)
            inward[axis] = -sgn
            for tri in ((q[0], q[1], q[2]), (q[0], q[2], q[3])):
                p = c[list(tri)]
                if np.dot(np.cross(p[1] - p[0], p[2] - p[0]), inward) < 0:
                    p = p[[0, 2, 1]]
                faces.append((p, np.tile(inward, (3, 1))))
    P = np.stack([f[0] for f in faces])
    N = np.stack([f[1] for f in faces])
    UV = np.zeros((len(P), 3, 2))
    mats = [scenes.material(albedo=albedo, metallic=0.0, roughness=roughness, transmission=0.0, ior=ior,
                            emission=emission, emission_scale_nits=scale_nits)]
    sc = scenes._pack([(P, N, UV, np.zeros(len(P), np.int32))], mats, None,
                      camera_eye=(0.0, 0.0, 0.0), camera_target=(0.0, 0.0, 1.0),
                      camera_fov=math.radians(30.0), sun_theta=math.radians(30.0))
    return sc


def instance_chain(n, ratio=1.5):
    """A TLAS that the agglomerative builder makes a chain n - 1 levels deep:
    instance k's box spans x in [r^k, r^(k+1)] and y, z in [-r^k, r^k]
    (r = ratio), so the cluster of instances 0..k-1 is always the best
    match of instance k.  Each instance holds two quads in the planes
    z = -1 and z = +1 of its unit box, so a ray along the x axis (y = z = 0)
    enters every box and hits nothing: closest-hit traversal from the origin
    pushes the farther child (the next instance) at every level, n - 1
    entries deep, before it reaches instance 0."""
    quad = lambda z: np.array([[[0, -1, z], [1, -1, z], [1, 1, z]], [[0, -1, z], [1, 1, z], [0, 1, z]]], np.float64)  # noqa: E731
    P = np.concatenate([quad(-1.0), quad(1.0)])
    N = np.zeros_like(P)
    N[:2, :, 2] = 1.0
    N[2:, :, 2] = -1.0
    UV = np.zeros((4, 3, 2))
    mats = [scenes.material(albedo=(0.7, 0.7, 0.7), roughness=0.6)]
    models = []
    for k in range(n):
        x0, x1, s = ratio ** k, ratio ** (k + 1), ratio ** k
        m = np.diag([x1 - x0, s, s, 1.0])
        m[0, 3] = x0
        models.append(scenes.mat_colmajor(m))
    sc = scenes._pack([(P, N, UV, np.zeros(4, np.int32))], mats, None,
                      instance_mesh=np.zeros(n, np.int32), instance_model=np.stack(models),
                      camera_eye=(-2.0, 0.0, 0.0), camera_target=(1.0, 0.0, 0.0), camera_fov=math.radians(20.0))
    return sc


def coincident_quads(n=24, seed=0, plane_z=7.3):
    """n instances of one axis-aligned quad (a flat BLAS box, z = 0 in object
    space) laid over each other in the world plane z = plane_z: rotations by
    multiples of 90 degrees about z (exact), or by a few 1e-6 rad about x and
    y (boxes no longer exactly flat), translations within the plane and along
    z by a few ulps of plane_z.  Rays from above hit several instances at
    distances equal to within an ulp or two: the case in which entering a
    BLAS with the closest distance so far as tmax could disagree with the
    reference's per-BLAS traversal from the original tmax."""
    rng = np.random.default_rng(seed)
    P = np.array([[[-1, -1, 0], [1, -1, 0], [1, 1, 0]], [[-1, -1, 0], [1, 1, 0], [-1, 1, 0]]], np.float64)
    N = np.tile(np.array([0.0, 0.0, 1.0]), (2, 3, 1))
    UV = np.zeros((2, 3, 2))
    ulp = np.spacing(np.float32(plane_z))
    models = []
    for k in range(n):
        if k % 3 == 2:
            ax, ay = rng.uniform(-3e-6, 3e-6, 2)
            r = _rot_x(ax) @ _rot_y(ay)
        else:
            c, s = [(1, 0), (0, 1), (-1, 0), (0, -1)][k % 4]
            r = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]], np.float64)
        t = (rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3), plane_z + float(ulp) * int(rng.integers(-2, 3)))
        models.append(scenes.trs(translate=t, rotate=r, scale=float(rng.choice([1.0, 1.5, 2.0]))))
    mats = [scenes.material(albedo=(0.6, 0.6, 0.6), roughness=0.5)]
    return scenes._pack([(P, N, UV, np.zeros(2, np.int32))], mats, None, instance_mesh=np.zeros(n, np.int32),
                        instance_model=np.stack(models), camera_eye=(0.0, 0.0, 0.0),
                        camera_target=(0.0, 0.0, 1.0), camera_fov=math.radians(40.0))


def _rot_x(a):
    return np.array([[1, 0, 0], [0, math.cos(a), -math.sin(a)], [0, math.sin(a), math.cos(a)]])


def _rot_y(a):
    return np.array([[math.cos(a), 0, math.sin(a)], [0, 1, 0], [-math.sin(a), 0, math.cos(a)]])
