"""GPU parity at the benchmark configurations' own workloads, against the
oracle on pixel subsets:

* C1 — the 1,000-instance, 1.24 M-triangle scene at 1920x1080, depth 8:
  4,096 random pixels of a 2-spp frame, wavefront (two tunings) and
  megakernel, bit-exact radiance and G/V-buffer words;
* C2 — the 64 cameras of uneven square resolution (log-uniform in
  [128, 1024]) rendered as one batch at 4 spp with material sort and a
  hipGraph per sample: 64 random pixels of every view;
* C3 — one cached 1080p frame of the C1 scene: every cache query equal to
  the oracle's bit for bit, the other paths bit-exact, cached pixels within
  the network's tolerance;
* C4 — 3840x2160, the first 100 instances moved between two frames with
  temporal accumulation, auto exposure, 6-level bloom and AgX: 4,096 random
  pixels of each frame (reprojection, motion vectors) bit-exact, and the
  whole post-processed output bit-exact;
* a TLAS 39 levels deep (the traversal stack spills past its 16 LDS
  entries) bit-exact, and one 139 deep (past the 128-entry limit) reported
  as KPT_ERR_OVERFLOW once, after which the context renders normally.
"""
import math

import numpy as np
import pytest

import oracle
from kittlespt_amd import abi, scenes
from kittlespt_amd.context import Context

from common import host_camera, instance_chain, random_rays, settings, view_proj

pytestmark = pytest.mark.gpu

W1, H1 = 1920, 1080


@pytest.fixture(scope="module")
def c1():
    print("[configs] generating the 1,000-instance scene", flush=True)
    sc = scenes.scene_instanced()
    ctx = Context(0)
    ctx.upload(sc)
    print("[configs] building the oracle's copy", flush=True)
    orc = oracle.OracleScene(sc)
    yield sc, ctx, orc
    ctx.close()
    orc.close()


def _bits_differ(a, b):
    return (np.ascontiguousarray(a, np.float32).view(np.uint32) !=
            np.ascontiguousarray(b, np.float32).view(np.uint32)).any(axis=-1)


def _gbuf_equal(g, o):
    if not np.array_equal(g[:, 2:], o[:, 2:]):
        return False
    fin = np.isfinite(o[:, :2].view(np.float32)).all(1)
    return np.array_equal(g[fin, :2], o[fin, :2])


TUNINGS = {"tail1": dict(tail=1, sort=0, blocks_per_cu=8, refill=0, chunk=128, graph=0, ray_sort=0),
           "stages_sorted": dict(tail=0, sort=1, blocks_per_cu=8, refill=0, chunk=128, graph=0, ray_sort=0)}


@pytest.mark.parametrize("mode", ["tail1", "stages_sorted", "megakernel"])
def test_c1_frame_matches_oracle(c1, mode):
    sc, ctx, orc = c1
    integ = abi.INTEGRATOR_MEGAKERNEL if mode == "megakernel" else abi.INTEGRATOR_WAVEFRONT
    if mode in TUNINGS:
        ctx.set_tuning(**TUNINGS[mode])
    st = settings(max_depth=8, integrator=integ, sun_theta=sc.sun_theta)
    ctx.set_settings(st)
    ctx.resize(W1, H1)
    p, v = view_proj(sc, W1, H1)
    ctx.set_view(p, v)
    cam = ctx.camera()
    print(f"[c1] 1920x1080 x 2 spp, depth 8, {mode}", flush=True)
    ctx.render(2)
    img = ctx.estimate().reshape(-1, 3)
    gbuf, vbuf = ctx.readback(2).reshape(-1, 4), ctx.readback(3).reshape(-1, 4)
    pix = np.sort(np.random.default_rng(41).choice(W1 * H1, 4096, replace=False)).astype(np.int32)
    cam.prev_inv_view[:] = cam.inv_view[:]
    cam.prev_inv_proj[:] = cam.inv_proj[:]
    o_rgb, o_g, o_v = orc.render_pixels(st, cam, W1, H1, pix, 0, 2, want_buffers=True)
    diff = _bits_differ(img[pix], o_rgb)
    print(f"[c1] {mode}: {int(diff.sum())} of {len(pix)} pixels differ; "
          f"hit pixels {(o_v[:, 3] >= 0).mean():.2f}")
    assert not diff.any()
    assert _gbuf_equal(gbuf[pix], o_g)
    assert np.array_equal(vbuf[pix, 2:], o_v[:, 2:])
    assert (o_v[:, 3] >= 0).mean() > 0.3                   # the frame sees the instances, not only sky


@pytest.mark.parametrize("mode", ["graph_sorted", "tail1"])
def test_c2_views_match_oracle(c1, mode):
    sc, ctx, orc = c1
    views = scenes.multi_view_cameras(sc)
    if mode == "graph_sorted":
        ctx.set_tuning(graph=1, sort=1, tail=0, blocks_per_cu=8, refill=0, chunk=128, ray_sort=0)
    else:
        ctx.set_tuning(**TUNINGS["tail1"])
    st = settings(max_depth=8, integrator=abi.INTEGRATOR_WAVEFRONT, sun_theta=sc.sun_theta)
    ctx.set_settings(st)
    print(f"[c2] 64 views ({sum(w * h for _, _, w, h in views)} pixels) x 4 spp, {mode}", flush=True)
    out = ctx.render_views(views, spp=4)
    ctx.set_tuning(graph=0)
    rng = np.random.default_rng(43)
    total = bad = 0
    for k, (proj, view, w, h) in enumerate(views):
        hc = Context(-2)
        hc.set_view(proj, view)
        cam = hc.camera()
        hc.close()
        pix = np.sort(rng.choice(w * h, 64, replace=False)).astype(np.int32)
        o = orc.render_pixels(st, cam, w, h, pix, 0, 4)
        g = out[k].reshape(-1, 4)[pix, :3]
        d = int(_bits_differ(g, o).sum())
        total += len(pix)
        bad += d
        if d:
            print(f"[c2] view {k} ({w}x{h}): {d} pixels differ")
    print(f"[c2] {mode}: {bad} of {total} pixels differ over {len(views)} views")
    assert bad == 0


@pytest.mark.parametrize("mode", ["tail1", "stages_sorted"])
def test_c3_cached_frame_matches_oracle(c1, mode):
    """A 1080p frame of the C1 scene in which every path ends in the cache at
    its bounce-1 vertex after that vertex's next-event estimate (the bench's
    mode), the network at random weights (the query, not the training, is
    what the oracle pins); bounce 1 in the path-per-lane tail or as sorted
    stage kernels."""
    from test_cache_oracle import he_weights
    sc, ctx, orc = c1
    ctx.set_tuning(**TUNINGS[mode])
    st = settings(max_depth=8, integrator=abi.INTEGRATOR_WAVEFRONT, sun_theta=sc.sun_theta)
    ctx.set_settings(st)
    ctx.resize(W1, H1)
    p, v = view_proj(sc, W1, H1)
    ctx.set_view(p, v)
    cam = ctx.camera()
    box = ((-10.0, -10.0, -10.0), (110.0, 110.0, 110.0))
    w = he_weights(51) * np.float32(0.5)
    ctx.cache_configure(enabled=True, train=False, seed=7, bounds=box, radiance_scale=1e4, nee_at_vertex=1)
    ctx.cache_set_weights(w)
    print(f"[c3] one cached 1080p frame, {mode}", flush=True)
    try:
        ctx.render(1)
        gp, gd, gb = ctx.cache_queries()
        est = ctx.estimate().reshape(-1, 3)
    finally:
        ctx.cache_configure(enabled=False, train=False, seed=7)
    print("[c3] oracle cache queries of the whole frame", flush=True)
    Lb, q = orc.render_cache_queries(st, cam, W1, H1, 0, nee=True)
    hit = np.nonzero(q[:, 0] > 0)[0]
    paths = gd[:, 3].view(np.uint32).astype(np.int64)
    print(f"[c3] {len(paths)} GPU queries, {len(hit)} oracle queries")
    assert len(hit) > 100_000 and np.array_equal(np.sort(paths), hit)
    order = np.argsort(paths)
    gp, gd, gb = gp[order], gd[order], gb[order]
    u = np.uint32
    assert np.array_equal(gp[:, :4].view(u), q[hit, 1:5].view(u))
    assert np.array_equal(gd[:, :3].view(u), q[hit, 5:8].view(u))
    assert np.array_equal(gb[:, :3].view(u), q[hit, 8:11].view(u))
    o = oracle.cache_forward(w, gp, gd, box[0], box[1], mode=0).astype(np.float32)
    want = Lb.copy()
    want[hit] = Lb[hit] + gb[:, :3] * (np.float32(1e4) * np.maximum(o, np.float32(0)))
    want = np.where(np.isfinite(want).all(1, keepdims=True), np.clip(want, 0, 1e8), 0).astype(np.float32)
    miss = np.setdiff1d(np.arange(W1 * H1), hit)
    assert not _bits_differ(est[miss], want[miss]).any()
    # the network's forward pass equals the oracle's within 2e-3 of its
    # output range on almost every query (the matrix cores sum in FP32 in
    # their own order, 10x the error the forward test observes); a sum that
    # lands on the other side of a BF16 rounding point of a hidden
    # activation moves the output further, rarely (a few in 10^6 queries)
    err = np.abs(est[hit] - want[hit])
    bound = gb[:, :3] * np.float32(1e4) * np.float32(2e-3 * np.abs(o).max()) + 1e-6 * np.abs(want[hit])
    ratio = (err / bound).max(axis=1)
    rel = err / (np.abs(want[hit]) + 1e-3 * np.abs(want).mean())
    print(f"[c3] cached pixels: max rel err {rel.max():.2e}, mean {rel.mean():.2e}, "
          f"beyond the bound {int((ratio > 1).sum())}, max err / bound {ratio.max():.2e}")
    assert (ratio > 1).sum() <= len(hit) // 100_000 and ratio.max() < 10 and rel.mean() < 2e-6


def _axis_angle(axis, ang):
    x, y, z = axis / np.linalg.norm(axis)
    c, s = math.cos(ang), math.sin(ang)
    return np.array([[c + x * x * (1 - c), x * y * (1 - c) - z * s, x * z * (1 - c) + y * s, 0],
                     [y * x * (1 - c) + z * s, c + y * y * (1 - c), y * z * (1 - c) - x * s, 0],
                     [z * x * (1 - c) - y * s, z * y * (1 - c) + x * s, c + z * z * (1 - c), 0],
                     [0, 0, 0, 1]], np.float64)


@pytest.mark.parametrize("tlas", ["rebuild", "refit"])
def test_c4_animated_4k_frames_match_oracle(c1, tlas):
    """Two 3840x2160 frames; between them the first 100 instances move
    (seeded linear + angular velocities, 1/60 s).  The TLAS is rebuilt on the
    host (the reference's per-frame rebuild) or refitted on the device.  Each
    frame's accumulation, G- and V-buffer on 4,096 random pixels equal the
    oracle's temporal frame fed the GPU's previous buffers; the post chain's
    output (auto exposure, bloom, AgX) equals the oracle's over the whole
    frame.  A refitted tree visits nodes in another order, so an exact tie
    between two hits could resolve differently: the refit run reports its
    differing pixels and allows 0.1 % of them."""
    sc, ctx, orc = c1
    W, H = 3840, 2160
    ctx.set_tuning(**TUNINGS["tail1"], tlas_device=0 if tlas == "rebuild" else 2)
    st = settings(max_depth=8, integrator=abi.INTEGRATOR_WAVEFRONT, sun_theta=sc.sun_theta)
    st.use_temporal_accumulation, st.enable_auto_exposure, st.bloom_generate_bloom = 1, 1, 1
    ctx.set_settings(st)
    ctx.resize(W, H)
    p, v = view_proj(sc, W, H)
    ctx.set_view(p, v)
    ctx.set_exposure(abi.KptExposure.camera(aperture=2.8, iso=100, shutter=1 / 60))
    dt = 1000.0 / 60.0
    ctx.set_frame_delta(dt)
    rng = np.random.default_rng(100)
    base = np.array(sc.instance_model[:100], np.float64).reshape(100, 4, 4).transpose(0, 2, 1)
    vel = rng.uniform(-3.0, 3.0, (100, 3))
    axes = rng.normal(size=(100, 3))
    omega = rng.uniform(0.2, 1.5, 100)
    pix = np.sort(np.random.default_rng(45).choice(W * H, 4096, replace=False)).astype(np.int32)
    prev_acc = np.zeros((H, W, 4), np.float32)
    prev_gb = np.zeros((H, W, 4), np.uint32)
    allowed = 0 if tlas == "rebuild" else len(pix) // 1000
    try:
        for f in range(2):
            if f == 1:
                models = np.empty((100, 16), np.float32)
                for i in range(100):
                    m = base[i].copy()
                    m[:3, 3] += vel[i] / 60.0
                    m = m @ _axis_angle(axes[i], omega[i] / 60.0)
                    models[i] = m.T.reshape(-1)
                ctx.set_instance_transforms(0, models)
                for i in range(100):
                    orc.set_instance_transform(i, models[i])
            e, film = ctx.exposure(), ctx.film()
            cam = ctx.camera()
            print(f"[c4] {tlas}: 4K frame {f}", flush=True)
            ctx.render(1)
            acc, gb, vb, out = ctx.readback(1), ctx.readback(2), ctx.readback(3), ctx.readback(4)
            r_o, g_o, v_o = orc.render_temporal_pixels(st, cam, W, H, f, prev_acc, prev_gb, pix)
            d = _bits_differ(acc.reshape(-1, 4)[pix], r_o)
            print(f"[c4] {tlas} frame {f}: {int(d.sum())} of {len(pix)} pixels differ; "
                  f"history > 0 on {(r_o[:, 3] > 0).mean():.2f}")
            assert int(d.sum()) <= allowed
            same = ~d
            assert _gbuf_equal(gb.reshape(-1, 4)[pix][same], g_o[same])
            assert not _bits_differ(vb.reshape(-1, 4)[pix][same], v_o[same]).any()
            if f == 1:
                assert (r_o[:, 3] > 0).mean() > 0.3           # reprojection accepts most of the frame
                print("[c4] oracle post chain over the whole 4K frame", flush=True)
                est = acc.copy()
                est[..., 3] = 0.0                              # the post chain reads the colour only
                _, out_o, _, _ = oracle.post_frame(est, gb, st, e, film, dt)
                od = _bits_differ(out, out_o)
                print(f"[c4] {tlas} post output: {int(od.sum())} of {W * H} pixels differ")
                assert int(od.sum()) <= (0 if tlas == "rebuild" else W * H // 1000)
            prev_acc, prev_gb = acc, gb
    finally:
        ctx.set_instance_transforms(0, sc.instance_model[:100])
        for i in range(100):
            orc.set_instance_transform(i, sc.instance_model[i])
        ctx.set_tuning(tlas_device=0)


# ----------------------------------------------------------------- traversal stack depth
def _chain_rays(n, seed):
    """Rays from the chain's origin: a bundle along +x (through every box,
    hitting nothing) and a cone that hits the quads of various instances."""
    rng = np.random.default_rng(seed)
    d = np.column_stack([np.ones(n), rng.uniform(-1.2, 1.2, n), rng.uniform(-1.2, 1.2, n)])
    d[: n // 8, 1:] = rng.uniform(-1e-4, 1e-4, (n // 8, 2))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros(n, abi.RAY_DTYPE)
    rays["org"] = rng.uniform(-0.05, 0.05, (n, 3))
    rays["dir"] = d
    rays["tmax"] = np.inf
    return rays


def test_deep_tlas_spills_the_stack_and_matches_oracle():
    sc = instance_chain(40)
    ctx = Context(0)
    orc = oracle.OracleScene(sc)
    try:
        ctx.upload(sc)
        rays = _chain_rays(8192, 47)
        o, _, depth = orc.trace(rays, oracle.MODE_CLOSEST)
        print(f"[deep] 40-instance chain: oracle stack depth {depth}", flush=True)
        assert depth >= 32                                   # past the 16 LDS entries, into the spill region
        g = ctx.trace_closest(rays)
        assert np.array_equal(g["prim_id"], o["prim_id"]) and np.array_equal(g["inst_id"], o["inst_id"])
        hit = o["prim_id"] >= 0
        assert 0.1 < hit.mean() < 0.95
        assert np.array_equal(g["t"][hit].view(np.uint32), o["t"][hit].view(np.uint32))
        assert np.array_equal(g["bary"][hit].view(np.uint32), o["bary"][hit].view(np.uint32))
        rays["tmax"] = np.random.default_rng(48).uniform(1.0, 1e4, len(rays)).astype(np.float32)
        assert np.array_equal(ctx.trace_any(rays), orc.trace(rays, oracle.MODE_ANY)[0])
        # the wavefront kernels' LDS share of the stack is a tuning knob: the
        # shallowest and deepest LDS parts, at every workgroup size, give
        # the same bits (the 128-entry capacity does not change)
        for integ, tail, stack, block in ((abi.INTEGRATOR_MEGAKERNEL, 0, 16, 256), (abi.INTEGRATOR_WAVEFRONT, 1, 16, 256),
                                          (abi.INTEGRATOR_WAVEFRONT, 0, 16, 256), (abi.INTEGRATOR_WAVEFRONT, 1, 8, 64),
                                          (abi.INTEGRATOR_WAVEFRONT, 0, 32, 128), (abi.INTEGRATOR_WAVEFRONT, 1, 24, 256)):
            ctx.set_tuning(tail=tail, stack=stack, block=block)
            st = settings(max_depth=4, integrator=integ)
            ctx.set_settings(st)
            ctx.resize(64, 64)
            p, v = view_proj(sc, 64, 64)
            ctx.set_view(p, v)
            cam = ctx.camera()
            print(f"[deep] render 64x64 x 2 spp, integrator {integ}, tail {tail}, LDS stack {stack}, "
                  f"workgroup {block}", flush=True)
            ctx.render(2)
            img = ctx.estimate().reshape(-1, 3)
            cam.prev_inv_view[:] = cam.inv_view[:]
            cam.prev_inv_proj[:] = cam.inv_proj[:]
            ref = orc.render_pixels(st, cam, 64, 64, np.arange(64 * 64, dtype=np.int32), 0, 2)
            assert not _bits_differ(img, ref).any()
    finally:
        ctx.close()
        orc.close()


def test_stack_overflow_is_reported_once():
    """A TLAS 139 levels deep: the on-axis rays need 139 stack entries, past
    the 128 the kernels hold; the lanes' queries end early, every kernel
    exits normally, the render and the synchronising call after a trace
    return KPT_ERR_OVERFLOW once, and the next render of a shallow scene on
    the same context succeeds."""
    sc = instance_chain(140, ratio=1.3)
    orc = oracle.OracleScene(sc)
    _, _, depth = orc.trace(_chain_rays(256, 49), oracle.MODE_CLOSEST)
    orc.close()
    assert depth > 130
    ctx = Context(0)
    try:
        ctx.upload(sc)
        rays = _chain_rays(4096, 49)
        print("[overflow] trace of 4096 rays through a 139-level TLAS", flush=True)
        with pytest.raises(abi.KptError) as ei:
            ctx.trace_closest(rays)
        assert ei.value.code == abi.KPT_ERR_OVERFLOW
        ctx.set_settings(settings(max_depth=2, integrator=abi.INTEGRATOR_WAVEFRONT))
        ctx.resize(32, 32)
        p, v = view_proj(sc, 32, 32)
        ctx.set_view(p, v)
        print("[overflow] render of the 139-level scene", flush=True)
        with pytest.raises(abi.KptError) as ei:
            ctx.render(1)
        assert ei.value.code == abi.KPT_ERR_OVERFLOW
        # the deepest LDS part keeps the 128-entry capacity
        ctx.set_tuning(stack=32, block=64)
        print("[overflow] render of the 139-level scene, LDS stack 32, workgroup 64", flush=True)
        with pytest.raises(abi.KptError) as ei:
            ctx.render(1)
        assert ei.value.code == abi.KPT_ERR_OVERFLOW
        shallow = instance_chain(10)
        ctx.upload(shallow)
        p, v = view_proj(shallow, 32, 32)
        ctx.set_view(p, v)
        print("[overflow] render of a 9-level scene on the same context", flush=True)
        ctx.render(1)
        ctx.trace_closest(_chain_rays(1024, 50))
    finally:
        ctx.close()
