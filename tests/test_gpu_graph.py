"""Per-frame hipGraph replay (tuning knob `graph`): every frame captured as
one graph and replayed (the executable graph updated in place from the next
frame's capture) must leave exactly the buffers the eager launches leave,
replay after replay, with the program's own eager work (readbacks, setting
changes) and synchronisations between the replays.
"""
import numpy as np
import pytest

from kittlespt_amd import abi, scenes
from common import settings, view_proj

pytestmark = pytest.mark.gpu


def _frames(ctx, sc, graph, frames, st, W, H, cache_w=None):
    ctx.set_tuning(graph=graph)
    ctx.set_settings(st)
    ctx.resize(W, H)
    p, v = view_proj(sc, W, H)
    ctx.set_view(p, v)
    ctx.set_exposure(abi.KptExposure.camera(aperture=16.0, iso=100, shutter=1 / 1000))
    ctx.set_frame_delta(16.0)
    if cache_w is not None:
        ctx.cache_configure(enabled=True, train=False, seed=9)
        ctx.cache_set_weights(cache_w)
    else:
        ctx.cache_configure(enabled=False, train=False, seed=9)
    out = []
    for f in range(frames):
        ctx.render(1)
        out.append(tuple(ctx.readback(k).copy() for k in (0, 1, 2, 4)) + (ctx.film().as_dict(),))
    return out


@pytest.mark.parametrize("mode", ["post", "temporal", "cache"])
def test_graph_replay_equals_eager(mode):
    from kittlespt_amd.context import Context
    sc = scenes.scene_plain(subdiv=3)
    W, H = 72, 40
    st = settings(max_depth=4, integrator=abi.INTEGRATOR_WAVEFRONT, sun_theta=sc.sun_theta)
    st.enable_auto_exposure, st.bloom_generate_bloom = 1, 1
    if mode == "temporal":
        st.use_temporal_accumulation = 1
    w = None
    if mode == "cache":
        rng = np.random.default_rng(3)
        w = (rng.uniform(-1, 1, abi.CACHE_PARAMS) * 0.2).astype(np.float32)
    runs = []
    for graph in (0, 1):                   # a fresh context each: the adapted scene luminance persists in one
        ctx = Context(0)
        try:
            ctx.upload(sc)
            print(f"[graph] {mode}: {'graph' if graph else 'eager'} frames", flush=True)
            runs.append(_frames(ctx, sc, graph, 5, st, W, H, w))
        finally:
            ctx.close()
    eager, graph = runs
    for f, (a, b) in enumerate(zip(eager, graph)):
        for k, name in enumerate(("render", "accum", "gbuffer", "output")):
            assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), f"frame {f}: {name} differs"
        assert a[4] == b[4], f"frame {f}: exposure state differs"
