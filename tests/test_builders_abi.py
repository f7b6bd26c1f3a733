"""Host builders and the C-ABI, without a GPU.

* libkpt_hip.so loads and exports every entry point include/kpt_abi.h
  declares.
* The product's BLAS/TLAS builders (kittlespt_amd/csrc/host/kpt_build.cpp)
  produce the oracle's reference-order trees node for node.
* Tree invariants and small hand-checkable cases pin the oracle's builders.
"""
import os
import re

import numpy as np
import pytest

import oracle
from kittlespt_amd import abi, scenes
from kittlespt_amd.context import Context

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_abi_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "kpt_abi.h")).read()
    declared = set(re.findall(r"^\s*(?:int|const char\*)\s+(kpt_\w+)\s*\(", hdr, re.M))
    assert len(declared) >= 25
    L = abi.lib()
    for name in sorted(declared):
        assert hasattr(L, name), name
    listed = {s[0] for s in abi.SIGNATURES}
    assert declared == listed
    assert L.kpt_get_abi_version() == 1


def test_host_only_context_errors():
    c = Context(-2)
    with pytest.raises(abi.KptError) as e:
        c.render(1)
    assert e.value.code in (abi.KPT_ERR_NO_DEVICE, abi.KPT_ERR_STATE)
    bad = scenes.scene_plain(subdiv=1)
    bad.material_ids = bad.material_ids.copy()
    bad.material_ids[0] = 99
    with pytest.raises(abi.KptError) as e:
        c.upload(bad)
    assert e.value.code == abi.KPT_ERR_INVALID_ARG
    out = np.zeros(8, np.int32)
    assert abi.lib().kpt_get_heat_map(c.h, out.ctypes.data, None, 4) == abi.KPT_ERR_NO_DEVICE
    assert abi.lib().kpt_get_heat_map(c.h, None, None, 4) == abi.KPT_ERR_INVALID_ARG
    c.close()


def test_blas_device_knob():
    c = Context(-2)
    assert c.tuning()["blas_device"] == 0
    with pytest.raises(abi.KptError) as e:
        c.set_tuning(blas_device=2)
    assert e.value.code == abi.KPT_ERR_INVALID_ARG
    c.set_tuning(blas_device=1)                 # no device: built on host threads
    c.upload(scenes.scene_plain(subdiv=1))
    assert c.build_timing()[1] is False
    c.set_tuning(blas_device=0)
    c.upload(scenes.scene_plain(subdiv=1))
    c.close()


def test_stack_and_workgroup_knobs():
    """LDS stack entries per lane (8..32) and workgroup size (64, 128, 256)
    of the wavefront's trace kernels, and the tail's parking threshold
    (0..64): defaults, accepted values, rejections."""
    c = Context(-2)
    t = c.tuning()
    assert (t["stack"], t["block"]) == (16, 256)
    for stack, block in ((8, 64), (32, 128), (24, 256)):
        c.set_tuning(stack=stack, block=block)
        assert (c.tuning()["stack"], c.tuning()["block"]) == (stack, block)
    for bad in ({"stack": 7}, {"stack": 33}, {"block": 96}, {"block": 512}, {"block": 0}):
        with pytest.raises(abi.KptError) as e:
            c.set_tuning(**bad)
        assert e.value.code == abi.KPT_ERR_INVALID_ARG
    assert (c.tuning()["stack"], c.tuning()["block"]) == (24, 256)
    assert c.tuning()["suspend"] == 16
    for v in (0, 8, 64):
        c.set_tuning(suspend=v)
        assert c.tuning()["suspend"] == v
    for bad in (-1, 65):
        with pytest.raises(abi.KptError) as e:
            c.set_tuning(suspend=bad)
        assert e.value.code == abi.KPT_ERR_INVALID_ARG
    assert c.tuning()["suspend"] == 64
    assert c.tuning()["overlap"] == 1
    c.set_tuning(overlap=0)
    assert c.tuning()["overlap"] == 0
    for bad in (-1, 2):
        with pytest.raises(abi.KptError) as e:
            c.set_tuning(overlap=bad)
        assert e.value.code == abi.KPT_ERR_INVALID_ARG
    c.close()


def test_settings_validation_effects_and_upscale():
    from common import settings
    c = Context(-2)
    st = settings()
    st.postprocess_enable_effects = 1
    c.set_settings(st)
    assert c.settings().postprocess_enable_effects == 1
    st.postprocess_enable_effects = 2
    with pytest.raises(abi.KptError) as e:
        c.set_settings(st)
    assert e.value.code == abi.KPT_ERR_INVALID_ARG
    st.postprocess_enable_effects = 0
    st.upscale_enable = 1
    c.set_settings(st)
    assert c.settings().upscale_enable == 1
    st.upscale_enable = 3
    with pytest.raises(abi.KptError) as e:
        c.set_settings(st)
    assert e.value.code == abi.KPT_ERR_INVALID_ARG
    c.close()


def test_host_only_set_exposure_matches_oracle():
    """Renderer::setExposure's film scalar (host side of the C-ABI) equals the
    oracle's getStandardOutputBasedExposure, bit for bit; defaults match the
    reference's Film before any setExposure."""
    c = Context(-2)
    f0 = c.film()
    assert (f0.luminance_exposure_scalar, f0.white_point_ev, f0.black_point_ev) == (1.0, 6.5, -10.0)
    for e in (abi.KptExposure.camera(), abi.KptExposure.camera(4.0, 800, 1 / 250, 1.0, 5.0, -8.0)):
        c.set_exposure(e)
        got = c.film()
        want = oracle.set_exposure(e)
        assert got.luminance_exposure_scalar == want.luminance_exposure_scalar
        assert (got.white_point_ev, got.black_point_ev) == (want.white_point_ev, want.black_point_ev)
        assert c.exposure().as_dict() == e.as_dict()
    with pytest.raises(abi.KptError):
        c.set_frame_delta(-1.0)
    c.close()


def _compare_trees(sc):
    ctx = Context(-2)
    ctx.upload(sc)
    orc = oracle.OracleScene(sc)
    for m in range(len(sc.mesh_first)):
        pn, pt = ctx.blas(m)
        on, ot = orc.blas(m)
        used = len(pn)
        assert np.array_equal(pn["first"], on["first"][:used])
        assert np.array_equal(pn["count"], on["count"][:used])
        assert np.array_equal(pn["pmin"], on["pmin"][:used]) and np.array_equal(pn["pmax"], on["pmax"][:used])
        assert np.array_equal(pt, ot)
        # nodes the reference reserves (2N-1) but never allocates stay default
        assert np.all(on["count"][used:] == 0) and np.all(on["first"][used:] == -1)
    pt, ot = ctx.tlas(), orc.tlas()
    assert np.array_equal(pt["lr"], ot["lr"]) and np.array_equal(pt["blas"][pt["lr"] == 0], ot["blas"][ot["lr"] == 0])
    assert np.array_equal(pt["pmin"], ot["pmin"]) and np.array_equal(pt["pmax"], ot["pmax"])
    info_p, info_o = ctx.info(), orc.info()
    assert info_p[4] == info_o[4] and info_p[5] == info_o[5]
    ctx.close()
    return orc


def test_product_trees_equal_oracle_plain():
    _compare_trees(scenes.scene_plain())


def test_product_trees_equal_oracle_instanced():
    _compare_trees(scenes.scene_instanced_small())


def test_tlas_of_1000_moving_instances_equals_oracle():
    """The TLAS of the 1,000-instance layout, rebuilt after moving 100
    instances (the per-frame update of the animated workload), node for node."""
    sc = scenes.scene_instanced(num_instances=1000, tex_size=8, tris_range=(60, 120), num_blas=8, num_materials=4)
    ctx = Context(-2)
    ctx.upload(sc)
    orc = oracle.OracleScene(sc)
    rng = np.random.default_rng(5)
    models = np.array(sc.instance_model[:100], np.float32).reshape(100, 16).copy()
    models[:, 12:15] += rng.uniform(-3, 3, (100, 3)).astype(np.float32)
    ctx.set_instance_transforms(0, models)
    for i in range(100):
        orc.set_instance_transform(i, models[i])
    pt, ot = ctx.tlas(), orc.tlas()
    assert len(pt) == len(ot) == 2000
    assert np.array_equal(pt.view(np.uint32), ot.view(np.uint32))
    ctx.close()
    orc.close()


def _leaves(nodes, root=0):
    out = []
    st = [root]
    while st:
        n = st.pop()
        if nodes["count"][n] > 0:
            out.append(n)
        else:
            st += [nodes["first"][n], nodes["first"][n] + 1]
    return out


def test_blas_invariants():
    sc = scenes.scene_instanced_small()
    orc = oracle.OracleScene(sc)
    for m in range(len(sc.mesh_first)):
        nodes, tri = orc.blas(m)
        first, count = int(sc.mesh_first[m]), int(sc.mesh_count[m])
        assert sorted(tri.tolist()) == list(range(first, first + count))   # a permutation
        leaves = _leaves(nodes)
        covered = np.zeros(count, np.int32)
        P = sc.positions
        for lf in leaves:
            s, c = nodes["first"][lf], nodes["count"][lf]
            covered[s:s + c] += 1
            pts = P[tri[s:s + c]].reshape(-1, 3)
            assert np.array_equal(pts.min(0), nodes["pmin"][lf]) and np.array_equal(pts.max(0), nodes["pmax"][lf])
            if c > 8:   # only an aborted split leaves more than 8 triangles
                cen = P[tri[s:s + c]].sum(1) * np.float32(0.3333)
                assert np.ptp(cen, axis=0).min() == 0 or c > 8
        assert np.all(covered == 1)
        # every interior box encloses its children
        for n in range(len(nodes)):
            if nodes["count"][n] == 0 and nodes["first"][n] > 0:
                for ch in (nodes["first"][n], nodes["first"][n] + 1):
                    assert np.all(nodes["pmin"][ch] >= nodes["pmin"][n]) and np.all(nodes["pmax"][ch] <= nodes["pmax"][n])


def _tiny_scene(pos):
    pos = np.asarray(pos, np.float32)
    n = len(pos)
    nrm = np.zeros_like(pos)
    nrm[..., 2] = 1
    return scenes.Scene(positions=pos, normals=nrm, uvs=np.zeros((n, 3, 2), np.float32),
                        material_ids=np.zeros(n, np.int32), mesh_first=np.array([0], np.int32),
                        mesh_count=np.array([n], np.int32), mesh_model=np.stack([scenes.identity16()]),
                        materials=[scenes.material()])


def _tri_at(x, y=0.0, s=0.05):
    return [[x, y, 0], [x + s, y, 0], [x, y + s, 0]]


def test_blas_small_cases():
    # 1 and 8 triangles: the root is a leaf
    for n in (1, 8):
        orc = oracle.OracleScene(_tiny_scene([_tri_at(i) for i in range(n)]))
        nodes, _ = orc.blas(0)
        assert nodes["count"][0] == n
    # 17 coincident triangles: every axis has zero centroid extent, the split aborts
    orc = oracle.OracleScene(_tiny_scene([_tri_at(0.0)] * 17))
    nodes, _ = orc.blas(0)
    assert nodes["count"][0] == 17
    # 9 triangles along x: one split, children allocated as the pair (1, 2),
    # left side holds the centroids below the plane, binned SAH is minimal
    xs = np.arange(9, dtype=np.float64)
    orc = oracle.OracleScene(_tiny_scene([_tri_at(x) for x in xs]))
    nodes, tri = orc.blas(0)
    assert nodes["count"][0] == 0 and nodes["first"][0] == 1
    nl, nr = nodes["count"][1], nodes["count"][2]
    assert nl + nr == 9 and nl > 0 and nr > 0
    cx = np.float32(0.3333) * (np.float32(0) + np.float32(0.05) + np.float32(0))
    lx = xs[tri[:nl]] if True else None
    assert lx.max() < xs[tri[nl:]].min()
    # exhaustive check of the 7 bin planes along x with numpy
    cen = np.array([np.float32((np.float32(x) + np.float32(x + 0.05) + np.float32(x)) * np.float32(0.3333))
                    for x in xs])
    lo, hi = cen.min(), cen.max()
    costs = []
    for i in range(7):
        plane = lo + (hi - lo) / 8 * (i + 1)
        L = cen < plane
        def sa(mask):
            pts = np.array([_tri_at(x) for x in xs[mask]], np.float32).reshape(-1, 3)
            d = pts.max(0) - pts.min(0)
            return 2 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0])
        costs.append(L.sum() * sa(L) + (~L).sum() * sa(~L))
    assert nl == int((cen < lo + (hi - lo) / 8 * (int(np.argmin(costs)) + 1)).sum())
    del cx


def test_tlas_small_cases():
    sc = _tiny_scene([_tri_at(0.0)])
    sc.instance_mesh = np.zeros(3, np.int32)
    sc.instance_model = np.stack([scenes.trs(translate=(10.0 * i, 0, 0)) for i in (0, 1, 5)])
    orc = oracle.OracleScene(sc)
    t = orc.tlas()
    assert len(t) == 2 * 3          # node 0 (root copy) + 3 leaves + 2 interior
    # the two close instances merge first, then the far one joins
    assert (t["lr"][4] & 0xFFFF, t["lr"][4] >> 16) in ((1, 2), (2, 1))
    assert t["lr"][0] == t["lr"][5]
    leaves = t["blas"][t["lr"] == 0]
    assert sorted(leaves[1:].tolist()) == [0, 1, 2] if len(leaves) == 4 else sorted(leaves.tolist()) == [0, 1, 2]


def test_host_only_context_builds_on_host_threads_whatever_the_knob():
    """blas_device = 1 on a context without a device: the upload builds on
    host threads (kpt_last_build_timing reports it) instead of failing."""
    sc = scenes.scene_plain(subdiv=2)
    a, b = Context(-2), Context(-2)
    try:
        a.set_tuning(blas_device=1)
        a.upload(sc)
        assert a.build_timing()[1] is False
        b.upload(sc)
        for m in range(len(sc.mesh_count)):
            assert a.blas(m)[0].tobytes() == b.blas(m)[0].tobytes()
    finally:
        a.close()
        b.close()



def _tree_covers(nodes, tri, first, count):
    assert sorted(tri.tolist()) == list(range(first, first + count))   # a permutation
    covered = np.zeros(count, np.int32)
    for lf in _leaves(nodes):
        s, c = nodes["first"][lf], nodes["count"][lf]
        covered[s:s + c] += 1
    assert np.all(covered == 1)


def test_host_builders_take_non_finite_positions():
    """NaN and infinite vertex positions, and instance boxes whose merged
    areas overflow: the reference's binning converts a NaN bin position to
    an integer and its best-match search returns -1 when no merged area is
    below infinity (as_builder.cuh:375-391); the host builders end with
    trees that cover every triangle and instance instead."""
    rng = np.random.default_rng(71)
    p = rng.uniform(-1, 1, (40, 3, 3)).astype(np.float32)
    p[7, 1, 2] = np.nan
    p[19, 0, 0] = np.inf
    p[23, 2, 1] = -np.inf
    sc = _tiny_scene(p)
    sc.instance_mesh = np.zeros(5, np.int32)
    sc.instance_model = np.stack([scenes.trs(translate=(3.0 * i, 0, 0)) for i in range(5)])
    ctx = Context(-2)
    try:
        ctx.upload(sc)
        nodes, tri = ctx.blas(0)
        _tree_covers(nodes, tri, 0, 40)
        t = ctx.tlas()
        assert len(t) == 2 * 5
        assert sorted(t["blas"][1:][t["lr"][1:] == 0].tolist()) == list(range(5))
        # finite meshes far apart: every merged area overflows to infinity
        q = np.array([_tri_at(0.0, s=1e19), _tri_at(0.0, s=1e19)], np.float32)
        sc2 = _tiny_scene(q)
        sc2.instance_mesh = np.zeros(4, np.int32)
        sc2.instance_model = np.stack([scenes.trs(translate=(1e19 * i, 0, 0)) for i in range(4)])
        ctx.upload(sc2)
        t = ctx.tlas()
        assert len(t) == 2 * 4
        assert sorted(t["blas"][1:][t["lr"][1:] == 0].tolist()) == list(range(4))
    finally:
        ctx.close()
