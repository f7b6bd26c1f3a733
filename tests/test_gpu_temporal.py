"""GPU parity of temporal accumulation with a moving camera and a moving
instance (reprojectAccumulate, kernels.cu:299-338; setMeshTransform keeping the
previous model, renderer.cu:250-269; the per-frame TLAS rebuild,
renderer.cu:115) against the oracle, bit-exact, for both integrators.

Each frame the oracle takes the GPU's previous accumulation and G-buffer
(checked bit-exact the frame before) and renders the whole frame itself.
"""
import numpy as np
import pytest

import oracle
from kittlespt_amd import abi, scenes
from common import settings, view_proj

pytestmark = pytest.mark.gpu


def _gbuf_equal(a, b):
    a, b = a.reshape(-1, 4), b.reshape(-1, 4)
    if not np.array_equal(a[:, 2:], b[:, 2:]):
        return False
    fin = np.isfinite(b[:, :2].view(np.float32)).all(1)
    return np.array_equal(a[fin, :2], b[fin, :2])


@pytest.mark.parametrize("integrator", [abi.INTEGRATOR_WAVEFRONT, abi.INTEGRATOR_MEGAKERNEL])
@pytest.mark.parametrize("which", ["plain", "instanced"])
def test_temporal_parity_with_motion(integrator, which):
    from kittlespt_amd.context import Context
    sc = scenes.scene_plain(subdiv=3) if which == "plain" else scenes.scene_instanced_small()
    W, H = 80, 56
    ctx = Context(0)
    orc = oracle.OracleScene(sc)
    try:
        ctx.upload(sc)
        st = settings(max_depth=3, integrator=integrator, sun_theta=sc.sun_theta)
        st.use_temporal_accumulation = 1
        ctx.set_settings(st)
        ctx.resize(W, H)
        eye = np.array(sc.camera_eye, np.float64)
        moving = 1 if which == "plain" else 3
        base = np.array(sc.mesh_model[1] if sc.instance_model is None else sc.instance_model[moving],
                        np.float32).reshape(4, 4)
        prev_acc = np.zeros((H, W, 4), np.float32)
        prev_gb = np.zeros((H, W, 4), np.uint32)
        for f in range(4):
            # camera dolly and an instance translation every frame after the first
            sc.camera_eye = tuple(eye + np.array([0.04, 0.0, 0.05]) * f * (1 if which == "plain" else 10))
            p, v = view_proj(sc, W, H)
            ctx.set_view(p, v)
            if f > 0:
                m = base.copy()
                m[3, :3] += np.float32(0.05 * f) * (1 if which == "plain" else 8)
                ctx.set_instance_transform(moving, m.reshape(-1))
                orc.set_instance_transform(moving, m.reshape(-1))
            cam = ctx.camera()
            print(f"[temporal] {which} integrator {integrator} frame {f}", flush=True)
            ctx.render(1)
            acc, gb, vb = ctx.readback(1), ctx.readback(2), ctx.readback(3)
            r_o, g_o, v_o = orc.render_temporal(st, cam, W, H, f, prev_acc, prev_gb)
            assert _gbuf_equal(gb, g_o), f"frame {f}: gbuffer differs"
            assert np.array_equal(vb.view(np.uint32), v_o.view(np.uint32)), f"frame {f}: vbuffer differs"
            diff = (acc.view(np.uint32) != r_o.view(np.uint32)).any(-1)
            assert not diff.any(), f"frame {f}: {int(diff.sum())} pixels differ"
            if f > 0:
                assert (acc[..., 3] > 0).any() and (acc[..., 3] == 0).any()
            prev_acc, prev_gb = acc, gb
    finally:
        ctx.close()
        orc.close()
