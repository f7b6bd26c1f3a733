"""Edge cases of the boundary on the GPU: an empty scene (no meshes) and a
scene whose only mesh has no triangles render as all-miss frames equal to
the oracle's (the reference prints and returns a default TLAS for an empty
scene, as_builder.cuh:316-319; here the frame is valid), traces through them
miss, and a context goes on rendering a real scene after them."""
import numpy as np
import pytest

import oracle
from kittlespt_amd import abi, scenes
from kittlespt_amd.context import Context
from common import random_rays, settings, view_proj

pytestmark = pytest.mark.gpu


def _empty(with_mesh):
    base = scenes.scene_plain(subdiv=1)
    m = 1 if with_mesh else 0
    return scenes.Scene(positions=np.zeros((0, 3, 3), np.float32), normals=np.zeros((0, 3, 3), np.float32),
                        uvs=np.zeros((0, 3, 2), np.float32), material_ids=np.zeros(0, np.int32),
                        mesh_first=np.zeros(m, np.int32), mesh_count=np.zeros(m, np.int32),
                        mesh_model=np.eye(4, dtype=np.float32).reshape(1, 16)[:m], materials=base.materials,
                        textures=[], pixels=np.zeros(1, np.uint8), camera_eye=base.camera_eye,
                        camera_target=base.camera_target, camera_fov=base.camera_fov, sun_theta=base.sun_theta)


@pytest.mark.parametrize("with_mesh", [False, True])
def test_empty_scene_renders_the_sky(with_mesh):
    sc = _empty(with_mesh)
    W, H = 32, 24
    ctx = Context(0)
    orc = oracle.OracleScene(sc)
    try:
        ctx.upload(sc)
        rays = random_rays(512, -5.0, 5.0, (0.0, 0.0, 0.0), 3.0, seed=5)
        hits = ctx.trace_closest(rays)
        assert (hits["prim_id"] == -1).all() and (hits["inst_id"] == -1).all()
        assert not ctx.trace_any(rays).any()
        for integ, tail in ((abi.INTEGRATOR_MEGAKERNEL, 0), (abi.INTEGRATOR_WAVEFRONT, 0),
                            (abi.INTEGRATOR_WAVEFRONT, 1)):
            ctx.set_tuning(tail=tail)
            st = settings(max_depth=3, integrator=integ, sun_theta=sc.sun_theta)
            ctx.set_settings(st)
            ctx.resize(W, H)
            p, v = view_proj(sc, W, H)
            ctx.set_view(p, v)
            cam = ctx.camera()
            print(f"[edge] empty scene (mesh={with_mesh}), integrator {integ}, tail {tail}", flush=True)
            ctx.render(2)
            img = ctx.estimate().reshape(-1, 3)
            ref = orc.render_pixels(st, cam, W, H, np.arange(W * H, dtype=np.int32), 0, 2)
            assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))
            # the sky above the horizon; below it the sky model's planet is black
            assert np.isfinite(img).all() and (img > 0).any() and (img == 0).any()
        # the same context then renders a real scene
        real = scenes.scene_plain(subdiv=1)
        ctx.upload(real)
        p, v = view_proj(real, W, H)
        ctx.set_view(p, v)
        ctx.render(1)
        assert np.isfinite(ctx.estimate()).all()
    finally:
        ctx.close()
        orc.close()
