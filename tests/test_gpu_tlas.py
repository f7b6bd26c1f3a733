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


def _union_ok(nodes):
    """Every merged node's box is fminf/fmaxf of its children's (host C
    library semantics: the second operand on ties)."""
    lo, hi, lr = nodes["pmin"], nodes["pmax"], nodes["lr"]
    for i in range(1, len(nodes)):
        if lr[i] == 0:
            continue
        a, b = lr[i] & 0xFFFF, lr[i] >> 16
        want_lo = np.where(lo[a] < lo[b], lo[a], lo[b])
        want_hi = np.where(hi[a] > hi[b], hi[a], hi[b])
        if not (np.array_equal(lo[i].view(np.uint32), want_lo.view(np.uint32)) and
                np.array_equal(hi[i].view(np.uint32), want_hi.view(np.uint32))):
            return False
    return True


def test_device_refit_keeps_topology_and_refits_boxes():
    """tlas_device = 2: the topology of the last build, boxes recomputed from
    the moved instances on the device; closest hits agree with the oracle's
    rebuilt tree (they may differ only where a different tree order decides an
    exact tie)."""
    from common import random_rays
    sc = scenes.scene_instanced(num_instances=1000, tex_size=8, tris_range=(60, 120), num_blas=8, num_materials=4)
    ctx = Context(0)
    ctx.upload(sc)
    ctx.set_tuning(tlas_device=2)
    orc = oracle.OracleScene(sc)
    before = ctx.tlas()
    rng = np.random.default_rng(21)
    for rnd in range(2):
        models = _moved(sc, 0, 100, rng)
        ctx.set_instance_transforms(0, models)
        for i in range(100):
            orc.set_instance_transform(i, models[i])
        print(f"[refit] round {rnd}", flush=True)
        pt, ot = ctx.tlas(), orc.tlas()
        assert np.array_equal(pt["lr"], before["lr"]) and np.array_equal(pt["blas"][1:], before["blas"][1:])
        # leaves: the instances' current world boxes, as the rebuilt tree holds them
        n = len(sc.instance_model)
        assert np.array_equal(pt[1:n + 1].view(np.uint32), ot[1:n + 1].view(np.uint32))
        assert _union_ok(pt)
        assert np.array_equal(pt[0]["pmin"].view(np.uint32), pt[2 * n - 1]["pmin"].view(np.uint32))
        rays = random_rays(100_000, (-10, -10, -10), (110, 110, 110), (50, 50, 50), 30.0, seed=31 + rnd)
        g = ctx.trace_closest(rays)
        o, _, _ = orc.trace(rays, oracle.MODE_CLOSEST)
        same = (g["prim_id"] == o["prim_id"]) & (g["inst_id"] == o["inst_id"])
        print(f"[refit] closest-hit agreement {same.mean():.6f}", flush=True)
        assert same.mean() >= 0.9999
        assert (o["prim_id"] >= 0).mean() > 0.05
    ctx.close()
    orc.close()
