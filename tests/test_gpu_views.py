"""Batched multi-view rendering (kpt_render_views): views of uneven
resolution traced as one set of paths must each equal rendering that view
alone (kpt_render from frame 0), bit for bit — the single-view render is
itself checked against the oracle in test_gpu_parity.py.
"""
import numpy as np
import pytest

from kittlespt_amd import abi, scenes
from common import settings

pytestmark = pytest.mark.gpu


def _view(sc, eye, target, w, h):
    proj = scenes.perspective_fov_lh(sc.camera_fov, w, h, 1.0, 100.0)
    view = scenes.look_at_view(eye, target)
    return proj, view, w, h


@pytest.mark.parametrize("which", ["plain", "instanced"])
def test_views_equal_single_renders(which):
    from kittlespt_amd.context import Context
    sc = scenes.scene_plain(subdiv=3) if which == "plain" else scenes.scene_instanced_small()
    eye, tgt = np.array(sc.camera_eye, np.float64), np.array(sc.camera_target, np.float64)
    views = [_view(sc, eye, tgt, 64, 48),
             _view(sc, eye + (0.3, 0.1, -0.2), tgt, 37, 29),
             _view(sc, eye * 1.2, tgt + (0.1, 0, 0), 96, 17),
             _view(sc, eye, tgt, 1, 1)]
    st = settings(max_depth=4, integrator=abi.INTEGRATOR_WAVEFRONT, sun_theta=sc.sun_theta)
    ctx = Context(0)
    try:
        ctx.upload(sc)
        ctx.set_settings(st)
        print(f"[views] {which}: {len(views)} views in one batch, 3 spp", flush=True)
        batch = ctx.render_views(views, spp=3)
        for k, (proj, view, w, h) in enumerate(views):
            ctx.resize(w, h)
            ctx.set_view(proj, view)
            ctx.render(3)
            single = ctx.estimate()
            got = batch[k][..., :3]
            diff = (got.view(np.uint32) != single.view(np.uint32)).any(-1)
            assert not diff.any(), f"view {k} ({w}x{h}): {int(diff.sum())} pixels differ"
            assert (batch[k][..., 3] == 1).all()
    finally:
        ctx.close()
