"""BLASes built on the device (kpt_blas.hip, tuning knob blas_device = 1):
node for node and slot for slot the host builder's trees, which restate
BLASBuilder (as_builder.cuh:110-261) and equal the oracle's."""
import time

import numpy as np
import pytest

from kittlespt_amd import abi, scenes
from kittlespt_amd.context import Context

from common import settings, view_proj

pytestmark = pytest.mark.gpu


def _build_both(sc):
    out = []
    for dev in (0, 1):
        ctx = Context(0)
        ctx.set_tuning(blas_device=dev)
        t0 = time.perf_counter()
        ctx.upload(sc)
        dt = time.perf_counter() - t0
        trees = [ctx.blas(m) for m in range(len(sc.mesh_count))]
        out.append((ctx, trees, ctx.info(), dt))
    return out


def _same_trees(sc, host, dev):
    for m in range(len(sc.mesh_count)):
        (hn, ht), (dn, dt) = host[1][m], dev[1][m]
        assert len(hn) == len(dn), f"mesh {m}: {len(hn)} vs {len(dn)} nodes"
        assert hn.tobytes() == dn.tobytes(), f"mesh {m}: nodes differ"
        assert np.array_equal(ht, dt), f"mesh {m}: triangle order differs"
    assert host[2] == dev[2]


def _mesh(p):
    p = np.asarray(p, np.float64).reshape(-1, 3, 3)
    n = np.tile(np.array([0.0, 1.0, 0.0]), (len(p), 3, 1))
    return p, n, np.zeros((len(p), 3, 2)), np.zeros(len(p), np.int32)


def test_device_blas_equals_host_instanced():
    sc = scenes.scene_instanced(num_instances=16, tex_size=8, tris_range=(20_000, 60_000), num_blas=8, num_materials=4)
    print(f"[blas] {int(sc.mesh_count.sum())} triangles in 8 meshes", flush=True)
    host, dev = _build_both(sc)
    print(f"[blas] upload with host build {host[3] * 1e3:.1f} ms, device build {dev[3] * 1e3:.1f} ms", flush=True)
    _same_trees(sc, host, dev)
    host[0].close()
    dev[0].close()


def test_device_blas_edge_cases():
    """Few triangles, an empty mesh, coincident centroids (the split aborts
    after the partition), two clusters whose bins leave every plane an empty
    side (no axis wins: the partition runs on z against 0, here a rotation
    when every centroid has z > 0), and signed zeros (the later of equal
    bounds wins)."""
    rng = np.random.default_rng(3)
    one = rng.uniform(-1, 1, (1, 9))
    eight = rng.uniform(-1, 1, (8, 9))
    nine = rng.uniform(-1, 1, (9, 9))
    seventeen = rng.uniform(-1, 1, (17, 9))
    same = np.tile(rng.uniform(-1, 1, (1, 9)), (64, 1))
    ca = rng.uniform(0.0, 0.1, (20, 9)) + np.array([0, 0, 1.0] * 3)
    cb = rng.uniform(0.0, 0.1, (20, 9)) + np.array([5, 5, 6.0] * 3)
    two_pos = np.concatenate([ca, cb])[rng.permutation(40)]
    cneg = ca - np.array([0, 0, 2.0] * 3)
    two_mixed = np.concatenate([cneg, cb])[rng.permutation(40)]
    zeros = rng.choice(np.array([-0.0, 0.0, 1.0, -1.0]), (200, 9))
    soup = rng.uniform(-3, 3, (3000, 9))
    meshes = [_mesh(x) for x in (one, eight, nine, seventeen, same, two_pos, two_mixed, zeros, soup)]
    empty = (np.zeros((0, 3, 3)), np.zeros((0, 3, 3)), np.zeros((0, 3, 2)), np.zeros(0, np.int32))
    meshes.insert(3, empty)
    sc = scenes._pack(meshes, [scenes.material()])
    print(f"[blas] edge cases, {len(meshes)} meshes", flush=True)
    host, dev = _build_both(sc)
    _same_trees(sc, host, dev)
    assert len(host[1][5][0]) == 1                       # coincident centroids: the root stays a leaf
    host[0].close()
    dev[0].close()


def test_device_build_is_the_default_on_a_device():
    ctx = Context(0)
    assert ctx.tuning()["blas_device"] == 1
    ctx.upload(scenes.scene_instanced_small())
    assert ctx.build_timing()[1] is True
    ctx.close()


def test_scene_the_device_builder_cannot_take_is_built_on_the_host():
    """A vertex position that is not finite (the device builder's bin keys
    order finite floats only): the upload still succeeds, on host threads,
    with the host builder's trees."""
    rng = np.random.default_rng(71)
    p = rng.uniform(-1, 1, (40, 3, 3))
    p[7, 1, 2] = np.nan
    p[19, 0, 0] = np.inf
    sc = scenes._pack([_mesh(p), _mesh(rng.uniform(-1, 1, (30, 3, 3)))],
                      [scenes.material()], None)
    ctx = Context(0)
    host = Context(-2)
    try:
        assert ctx.tuning()["blas_device"] == 1
        ctx.upload(sc)
        assert ctx.build_timing()[1] is False
        host.upload(sc)
        for m in range(2):
            assert ctx.blas(m)[0].tobytes() == host.blas(m)[0].tobytes()
            assert np.array_equal(ctx.blas(m)[1], host.blas(m)[1])
        ctx.upload(scenes.scene_instanced_small())          # finite again: on the device
        assert ctx.build_timing()[1] is True
    finally:
        ctx.close()
        host.close()


def test_device_blas_renders_the_same():
    sc = scenes.scene_instanced_small()
    W, H = 64, 48
    imgs = []
    for dev in (0, 1):
        ctx = Context(0)
        ctx.set_tuning(blas_device=dev)
        ctx.upload(sc)
        ctx.set_settings(settings(max_depth=4, integrator=abi.INTEGRATOR_WAVEFRONT, sun_theta=sc.sun_theta))
        ctx.resize(W, H)
        p, v = view_proj(sc, W, H)
        ctx.set_view(p, v)
        print(f"[blas] render with device build {dev}", flush=True)
        ctx.render(1)
        imgs.append(ctx.readback(1))
        ctx.close()
    assert np.array_equal(imgs[0].view(np.uint32), imgs[1].view(np.uint32))
