"""The TLAS rebuilt on the device after instance edits (kpt_tlas.hip): node for
node the oracle's restatement of TLASBuilder::buildTLAS (as_builder.cuh:312-392)
over the same instance boxes, and the same images as the host-built tree.
"""
import numpy as np
import pytest

import oracle
from kittlespt_amd import abi, scenes
from kittlespt_amd.context import Context

from common import settings, view_proj

pytestmark = pytest.mark.gpu


def _moved(sc, first, count, rng, scale=3.0, rotate=False):
    models = np.array(sc.instance_model[first:first + count], np.float32).reshape(count, 16).copy()
    models[:, 12:15] += rng.uniform(-scale, scale, (count, 3)).astype(np.float32)
    if rotate:
        c, s = np.cos(0.3), np.sin(0.3)
        r = np.array([[c, 0, -s, 0], [0, 1, 0, 0], [s, 0, c, 0], [0, 0, 0, 1]], np.float32)
        for i in range(count):
            m = models[i].reshape(4, 4)   # column-major: rows of this view are columns
            models[i] = (r.T @ m.T).T.reshape(16) if i % 2 else models[i]
    return models


def test_device_tlas_equals_oracle_over_animation_rounds():
    sc = scenes.scene_instanced(num_instances=1000, tex_size=8, tris_range=(60, 120), num_blas=8, num_materials=4)
    ctx = Context(0)
    ctx.upload(sc)
    assert ctx.tuning()["tlas_device"] == 0          # the host builds by default
    ctx.set_tuning(tlas_device=1)
    orc = oracle.OracleScene(sc)
    rng = np.random.default_rng(5)
    for rnd in range(3):
        models = _moved(sc, 0, 100, rng, rotate=(rnd == 2))
        ctx.set_instance_transforms(0, models)
        for i in range(100):
            orc.set_instance_transform(i, models[i])
        print(f"[tlas] 1000 instances, round {rnd}", flush=True)
        pt, ot = ctx.tlas(), orc.tlas()
        assert len(pt) == len(ot) == 2000
        assert np.array_equal(pt.view(np.uint32), ot.view(np.uint32)), f"round {rnd}"
    ctx.close()
    orc.close()


@pytest.mark.parametrize("n", [1, 2, 3, 17])
def test_device_tlas_small_counts_and_ties(n):
    """Few instances, and (for n >= 3) several instances moved onto the same
    place: equal merged areas, where the first minimum in list order decides."""
    sc = scenes.scene_instanced(num_instances=n, tex_size=8, tris_range=(40, 60), num_blas=min(n, 3), num_materials=2)
    ctx = Context(0)
    ctx.upload(sc)
    ctx.set_tuning(tlas_device=1)
    orc = oracle.OracleScene(sc)
    models = np.array(sc.instance_model[:n], np.float32).reshape(n, 16).copy()
    models[:, 12:15] += np.float32(0.5)
    if n >= 3:
        models[1:, :] = models[0]                   # all but the first on one spot
        models[n // 2, 12] += np.float32(0.25)      # one slightly off
    ctx.set_instance_transforms(0, models)
    for i in range(n):
        orc.set_instance_transform(i, models[i])
    print(f"[tlas] {n} instances with ties", flush=True)
    pt, ot = ctx.tlas(), orc.tlas()
    assert len(pt) == len(ot) == 2 * n
    assert np.array_equal(pt.view(np.uint32), ot.view(np.uint32))
    ctx.close()
    orc.close()


def test_device_and_host_rebuilds_render_the_same():
    sc = scenes.scene_instanced_small()
    W, H = 80, 64
    imgs = []
    for dev in (1, 0):
        ctx = Context(0)
        ctx.upload(sc)
        ctx.set_tuning(tlas_device=dev)
        ctx.set_settings(settings(max_depth=6, integrator=abi.INTEGRATOR_WAVEFRONT, sun_theta=sc.sun_theta))
        ctx.resize(W, H)
        p, v = view_proj(sc, W, H)
        ctx.set_view(p, v)
        rng = np.random.default_rng(9)
        n = len(sc.instance_model)
        k = max(1, n // 3)
        out = []
        for f in range(3):
            ctx.set_instance_transforms(0, _moved(sc, 0, k, rng, scale=1.0))
            print(f"[tlas] render after edit {f}, device build {dev}", flush=True)
            ctx.render(1)
            out.append((ctx.readback(1), ctx.readback(2)))
        imgs.append(out)
        ctx.close()
    for f in range(3):
        assert np.array_equal(imgs[0][f][0].view(np.uint32), imgs[1][f][0].view(np.uint32)), f"frame {f}"
        assert np.array_equal(imgs[0][f][1], imgs[1][f][1])
