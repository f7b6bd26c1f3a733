"""DebugData heat map (blas.cuh:39-42, tlas.cu:87, kernels.cu:440-443): the
per-pixel BLAS and TLAS node pops of each path's closest-hit traces, from the
counted wavefront kernels, against the oracle's traversal of the same
sample (BLASes entered as the kernels enter them), and bounded by the
reference's own traversal."""
import numpy as np
import pytest

import oracle
from kittlespt_amd import abi, scenes
from kittlespt_amd.context import Context

from common import settings, view_proj

pytestmark = pytest.mark.gpu

W, H = 48, 32


@pytest.fixture(scope="module")
def inst():
    sc = scenes.scene_instanced_small()
    ctx = Context(0)
    ctx.upload(sc)
    orc = oracle.OracleScene(sc)
    yield sc, ctx, orc
    ctx.close()
    orc.close()


def _counted_render(sc, ctx, tail):
    st = settings(max_depth=4, integrator=abi.INTEGRATOR_WAVEFRONT, sun_theta=sc.sun_theta)
    ctx.set_settings(st)
    ctx.resize(W, H)
    p, v = view_proj(sc, W, H)
    ctx.set_view(p, v)
    ctx.set_tuning(refill=0, chunk=64, blocks_per_cu=8, sort=1, graph=0, ray_sort=0, tail=tail)
    ctx.reset()
    ctx.set_trace_counting(True)
    try:
        ctx.render(1)
        heat = ctx.heat_map().reshape(-1, 2)
        rgba = ctx.heat_map(rgba=True).reshape(-1, 4)
        counts = ctx.render_counters()
    finally:
        ctx.set_trace_counting(False)
    return st, ctx.camera(), heat, rgba, counts


@pytest.mark.parametrize("tail", [0, 1])
def test_heat_map_matches_oracle(inst, tail):
    sc, ctx, orc = inst
    print(f"case: heat map, small instanced scene {W}x{H}, tail {tail}", flush=True)
    st, cam, heat, rgba, counts = _counted_render(sc, ctx, tail)
    px = np.arange(W * H, dtype=np.int32)
    want = oracle.heat_map(orc, st, cam, W, H, px, frame=0, entry_bound=True)
    ref = oracle.heat_map(orc, st, cam, W, H, px, frame=0, entry_bound=False)
    diff = np.argwhere(heat != want)
    assert len(diff) == 0, (len(diff), diff[:5], heat[diff[:5, 0]], want[diff[:5, 0]])
    assert np.all(heat <= ref)
    assert heat[:, 0].sum() > 0 and heat[:, 1].sum() > 0
    if tail == 0:
        # every closest-hit trace runs in the counted stage kernels (shadow
        # rays are not counted there): the frame totals are the map's sums
        assert heat[:, 0].sum() == counts["blas_interior_pops"] + counts["blas_leaf_pops"]
        assert heat[:, 1].sum() == counts["tlas_pops"]
    f = heat.astype(np.float32)
    want_rgba = np.stack([np.zeros(len(f), np.float32), f[:, 0] * np.float32(0.02), f[:, 1] * np.float32(0.05),
                          np.ones(len(f), np.float32)], 1)
    assert np.array_equal(rgba, want_rgba)


def test_heat_map_needs_a_counted_render(inst):
    sc, ctx, orc = inst
    st = settings(max_depth=2, integrator=abi.INTEGRATOR_WAVEFRONT, sun_theta=sc.sun_theta)
    ctx.set_settings(st)
    ctx.resize(W, H)
    p, v = view_proj(sc, W, H)
    ctx.set_view(p, v)
    ctx.set_trace_counting(False)
    ctx.render(1)
    with pytest.raises(abi.KptError) as e:
        ctx.heat_map()
    assert e.value.code == abi.KPT_ERR_STATE
