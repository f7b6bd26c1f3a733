"""GPU parity: the gfx950 kernels against the CPU oracle on identical inputs.

Bar: traversal bit-exact (instance id, primitive id, t, barycentrics, occlusion);
radiance of every pixel equal to the oracle's bit for bit (no pixel may
differ; the mean relative error bound of 1e-4 is asserted beside it);
G-buffer words bit-exact.
"""
import numpy as np
import pytest

import oracle
from kittlespt_amd import abi, scenes
from kittlespt_amd.context import Context

from common import camera_rays, host_camera, random_rays, settings, view_proj

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def plain():
    sc = scenes.scene_plain()
    ctx = Context(0)
    ctx.upload(sc)
    orc = oracle.OracleScene(sc)
    yield sc, ctx, orc
    ctx.close()
    orc.close()


@pytest.fixture(scope="module")
def inst_small():
    sc = scenes.scene_instanced_small()
    ctx = Context(0)
    ctx.upload(sc)
    orc = oracle.OracleScene(sc)
    yield sc, ctx, orc
    ctx.close()
    orc.close()


def _assert_hits_equal(g, o):
    assert np.array_equal(g["prim_id"], o["prim_id"]), f"prim mismatches: {(g['prim_id'] != o['prim_id']).sum()}"
    assert np.array_equal(g["inst_id"], o["inst_id"])
    hit = o["prim_id"] >= 0
    assert np.array_equal(g["t"][hit].view(np.uint32), o["t"][hit].view(np.uint32))
    assert np.array_equal(g["bary"][hit].view(np.uint32), o["bary"][hit].view(np.uint32))
    assert np.all(np.isinf(g["t"][~hit]))


def test_trace_closest_plain_camera(plain):
    sc, ctx, orc = plain
    cam = host_camera(sc, 128, 128)
    rays = camera_rays(cam, 128, 128)
    print("case: plain scene, 128x128 camera rays", flush=True)
    g = ctx.trace_closest(rays)
    o, _, _ = orc.trace(rays, oracle.MODE_CLOSEST)
    _assert_hits_equal(g, o)
    assert (o["prim_id"] >= 0).mean() > 0.3


def test_trace_closest_plain_random(plain):
    sc, ctx, orc = plain
    rays = random_rays(50_000, (-3, 0.05, -3), (3, 3, 3), (0, 1, 0), 0.8, seed=7)
    print("case: plain scene, 50k random rays", flush=True)
    g = ctx.trace_closest(rays)
    o, _, _ = orc.trace(rays, oracle.MODE_CLOSEST)
    _assert_hits_equal(g, o)


def test_trace_any_plain(plain):
    sc, ctx, orc = plain
    rays = random_rays(50_000, (-3, 0.05, -3), (3, 3, 3), (0, 1, 0), 0.8, seed=8)
    rays["tmax"] = np.random.default_rng(9).uniform(0.1, 6.0, len(rays)).astype(np.float32)
    print("case: plain scene, 50k any-hit rays with finite tmax", flush=True)
    g = ctx.trace_any(rays)
    o, _, _ = orc.trace(rays, oracle.MODE_ANY)
    assert np.array_equal(g, o)
    assert 0.05 < o.mean() < 0.95


def test_trace_instanced(inst_small):
    sc, ctx, orc = inst_small
    rays = random_rays(40_000, (-10, -10, -10), (110, 110, 110), (50, 50, 50), 30.0, seed=11)
    print("case: small instanced scene, 40k random rays", flush=True)
    g = ctx.trace_closest(rays)
    o, _, _ = orc.trace(rays, oracle.MODE_CLOSEST)
    _assert_hits_equal(g, o)
    ga = ctx.trace_any(rays)
    oa, _, _ = orc.trace(rays, oracle.MODE_ANY)
    assert np.array_equal(ga, oa)


def test_trace_counts_bounded_by_reference(plain):
    """The GPU traversal does no more work than the reference's (it enters
    each BLAS with the closest hit so far as tmax)."""
    sc, ctx, orc = plain
    cam = host_camera(sc, 64, 64)
    rays = camera_rays(cam, 64, 64)
    g, gc = ctx.trace_closest(rays, counted=True)
    o, oc, _ = orc.trace(rays, oracle.MODE_CLOSEST)
    _assert_hits_equal(g, o)
    for k in ("blas_interior_pops", "blas_leaf_pops", "tri_tests", "blas_entries", "tlas_pops"):
        assert 0 < gc[k] <= oc[k], (k, gc[k], oc[k])


def _render_both(sc, ctx, orc, W, H, spp, integrator, max_depth=4, seed_mode=abi.SEED_XOR):
    st = settings(max_depth=max_depth, seed_mode=seed_mode, integrator=integrator,
                  sun_theta=sc.sun_theta)
    ctx.set_settings(st)
    ctx.resize(W, H)
    p, v = view_proj(sc, W, H)
    ctx.set_view(p, v)
    cam = ctx.camera()
    ctx.render(spp)
    img = ctx.estimate().reshape(-1, 3)
    gbuf = ctx.readback(2).reshape(-1, 4)
    vbuf = ctx.readback(3).reshape(-1, 4)
    if spp > 1:
        cam.prev_inv_view[:] = cam.inv_view[:]
        cam.prev_inv_proj[:] = cam.inv_proj[:]
    pix = np.arange(W * H, dtype=np.int32)
    orgb, og, ov = orc.render_pixels(st, cam, W, H, pix, 0, spp, want_buffers=True)
    return img, gbuf, vbuf, orgb, og, ov


def _rel_err(a, b):
    return np.abs(a.astype(np.float64) - b) / np.maximum(np.abs(b.astype(np.float64)), 1e-6)


@pytest.mark.parametrize("integrator", [abi.INTEGRATOR_MEGAKERNEL, abi.INTEGRATOR_WAVEFRONT])
def test_render_plain_parity(plain, integrator):
    sc, ctx, orc = plain
    print(f"case: plain scene render 64x64 x 2 spp, integrator {integrator}", flush=True)
    img, gbuf, vbuf, orgb, og, ov = _render_both(sc, ctx, orc, 64, 64, 2, integrator)
    err = _rel_err(img, orgb)
    diff = int((img.view(np.uint32) != orgb.view(np.uint32)).any(axis=1).sum())
    print(f"mean rel err {err.mean():.3e}, max {err.max():.3e}, bit-different pixels {diff}/{len(img)}")
    assert err.mean() <= 1e-4
    assert diff == 0
    assert np.array_equal(gbuf[:, 2:4], og[:, 2:4])                       # albedo bytes + depth bits
    finite = np.isfinite(gbuf[:, :2].view(np.float32)).all(axis=1)
    assert np.array_equal(gbuf[finite, :2], og[finite, :2])               # octahedral normal bits
    assert np.array_equal(vbuf[:, 2:], ov[:, 2:])                         # primitive / instance ids
    assert np.allclose(vbuf[:, :2], ov[:, :2], atol=0, rtol=0, equal_nan=True)


def test_render_instanced_parity(inst_small):
    sc, ctx, orc = inst_small
    print("case: instanced scene render 48x32 x 1 spp wavefront, depth 8", flush=True)
    img, gbuf, vbuf, orgb, og, ov = _render_both(sc, ctx, orc, 48, 32, 1, abi.INTEGRATOR_WAVEFRONT, max_depth=8)
    err = _rel_err(img, orgb)
    diff = int((img.view(np.uint32) != orgb.view(np.uint32)).any(axis=1).sum())
    print(f"mean rel err {err.mean():.3e}, bit-different pixels {diff}")
    assert err.mean() <= 1e-4
    assert diff == 0


def test_wavefront_equals_megakernel(plain):
    sc, ctx, orc = plain
    W, H = 96, 64
    out = []
    for integ in (abi.INTEGRATOR_MEGAKERNEL, abi.INTEGRATOR_WAVEFRONT):
        ctx.set_settings(settings(max_depth=4, integrator=integ, sun_theta=sc.sun_theta))
        ctx.resize(W, H)
        p, v = view_proj(sc, W, H)
        ctx.set_view(p, v)
        ctx.render(3)
        out.append((ctx.readback(0), ctx.readback(2)))
    assert np.array_equal(out[0][0].view(np.uint32), out[1][0].view(np.uint32))
    assert np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("knobs", [{"ray_sort": 1}, {"ray_sort": 1, "refill": 1}, {"sort": 0, "blocks_per_cu": 3},
                                   {"tail": 1}, {"tail": 2}, {"tail": 4, "ray_sort": 1}, {"sort": 2, "tail": 2},
                                   {"stack": 8, "block": 64}, {"stack": 32, "block": 128, "tail": 1},
                                   {"stack": 24, "block": 64, "refill": 1, "tail": 2},
                                   {"tail": 1, "suspend": 0}, {"tail": 1, "suspend": 64, "blocks_per_cu": 1},
                                   {"tail": 1, "suspend": 64, "blocks_per_cu": 1, "stack": 8, "block": 64},
                                   {"overlap": 0}, {"overlap": 0, "tail": 1}])
def test_launch_shapes_do_not_change_the_image(inst_small, knobs):
    """Every tuning knob only reorders work: the image and G-buffer are the
    default launch shape's, bit for bit (80x64: the camera rays run in 8x8
    tile order)."""
    sc, ctx, orc = inst_small
    W, H = 80, 64
    out = []
    for kn in ({}, knobs):
        ctx.set_tuning(refill=0, chunk=128, refill_min=16, blocks_per_cu=8, sort=1, graph=0, ray_sort=0, tail=0,
                       stack=16, block=256, suspend=16, overlap=1)
        ctx.set_tuning(**kn)
        ctx.set_settings(settings(max_depth=6, integrator=abi.INTEGRATOR_WAVEFRONT, sun_theta=sc.sun_theta))
        ctx.resize(W, H)
        p, v = view_proj(sc, W, H)
        ctx.set_view(p, v)
        print(f"[shapes] {kn}", flush=True)
        ctx.render(3)
        out.append((ctx.readback(1), ctx.readback(2)))
    ctx.set_tuning(refill=0, chunk=128, refill_min=16, blocks_per_cu=8, sort=1, graph=0, ray_sort=0, tail=0,
                   stack=16, block=256, suspend=16, overlap=1)
    assert np.array_equal(out[0][0].view(np.uint32), out[1][0].view(np.uint32))
    assert np.array_equal(out[0][1], out[1][1])


def test_reference_seed_mode(plain):
    sc, ctx, orc = plain
    img, gbuf, vbuf, orgb, og, ov = _render_both(sc, ctx, orc, 32, 32, 3, abi.INTEGRATOR_WAVEFRONT,
                                                 seed_mode=abi.SEED_REFERENCE)
    assert _rel_err(img, orgb).mean() <= 1e-4
    assert not (img.view(np.uint32) != orgb.view(np.uint32)).any()


def test_stage_stats_of_the_tail(inst_small):
    """Per-stage measurements: with the tail from bounce 1, each frame has one
    closest-hit trace launch and one tail launch; counting mode fills the
    traversal events and the tail's path and vertex counts, and does not
    change the image."""
    sc, ctx, orc = inst_small
    W, H, spp = 80, 64, 2
    ctx.set_tuning(refill=0, chunk=128, refill_min=16, blocks_per_cu=8, sort=1, graph=0, ray_sort=0, tail=1)
    ctx.set_settings(settings(max_depth=6, integrator=abi.INTEGRATOR_WAVEFRONT, sun_theta=sc.sun_theta))
    ctx.resize(W, H)
    p, v = view_proj(sc, W, H)
    ctx.set_view(p, v)
    imgs, stats = [], []
    for counting in (False, True):
        ctx.reset()
        ctx.set_trace_counting(counting)
        print(f"[stages] counting={counting}", flush=True)
        ctx.render(spp)
        imgs.append(ctx.readback(1))
        stats.append({s: ctx.stage_stats(s) for s in ("trace", "tail")})
    ctx.set_trace_counting(False)
    ctx.set_tuning(tail=0)
    assert np.array_equal(imgs[0].view(np.uint32), imgs[1].view(np.uint32))
    for st in stats:
        assert st["trace"]["launches"] == spp and st["tail"]["launches"] == spp
        assert st["trace"]["ms"] > 0 and st["tail"]["ms"] > 0
    t, c = stats[1]["tail"], stats[1]["trace"]
    assert 0 < t["paths"] <= W * H * spp
    assert t["vertices"] >= t["paths"]
    for k in ("blas_interior_pops", "blas_leaf_pops", "tri_tests", "tlas_pops"):
        assert t["events"][k] > 0 and c["events"][k] > 0, k
    assert stats[0]["tail"]["paths"] == 0 and stats[0]["tail"]["events"]["tri_tests"] == 0


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_coincident_instances_match_oracle(seed):
    """Overlapping coplanar instances a few ulps apart (flat BLAS boxes,
    exact and ulp-level ties between instances on most rays): the kernels'
    closest hits (with their BLAS-entry bound) equal the reference
    traversal's bit for bit, and any-hit agrees."""
    from common import coincident_quads
    z = 7.3 + seed
    sc = coincident_quads(seed=seed, plane_z=z)
    ctx = Context(0)
    orc = oracle.OracleScene(sc)
    try:
        ctx.upload(sc)
        rng = np.random.default_rng(seed)
        n = 200_000
        tgt = np.column_stack([rng.uniform(-1.5, 1.5, n), rng.uniform(-1.5, 1.5, n), np.full(n, z)])
        o = np.column_stack([rng.uniform(-3, 3, n), rng.uniform(-3, 3, n), rng.uniform(-1, 2, n)])
        d = tgt - o
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        rays = np.zeros(n, abi.RAY_DTYPE)
        rays["org"], rays["dir"], rays["tmax"] = o, d, np.inf
        print(f"case: coincident quads seed {seed}, 200k rays", flush=True)
        g = ctx.trace_closest(rays)
        r, _, _ = orc.trace(rays, oracle.MODE_CLOSEST)
        _assert_hits_equal(g, r)
        rays["tmax"] = rng.uniform(1.0, 12.0, n).astype(np.float32)
        assert np.array_equal(ctx.trace_any(rays), orc.trace(rays, oracle.MODE_ANY)[0])
    finally:
        ctx.close()
        orc.close()
