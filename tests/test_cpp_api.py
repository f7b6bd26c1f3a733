"""The C++ Renderer class (include/kittlespt/renderer.hpp), which mirrors
KittlesPT/src/renderer.hpp:13-58 over the C-ABI.

CPU: the header compiles and links against libkpt_hip.so.  GPU: an application
written against the header (tests/cpp/renderer_api.cpp) renders a scene and
its image and V-buffer are bit-identical to the same scene rendered through
the Python binding (which makes the same C-ABI calls, with one kpt_render of
all samples where the application calls executeRendering once per sample).
"""
import os
import subprocess

import numpy as np
import pytest

from kittlespt_amd import abi, build, scenes
from common import settings, view_proj

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpp", "renderer_api.cpp")
EXE = os.path.join(HERE, "cpp", "_build", "renderer_api")


def _exe():
    deps = [SRC, abi.LIB_PATH] + [os.path.join(build.INCLUDE, "kittlespt", f) for f in
                                  ("renderer.hpp", "pod_types.hpp")] + [os.path.join(build.INCLUDE, "kpt_abi.h")]
    if not os.path.exists(EXE) or os.path.getmtime(EXE) < max(os.path.getmtime(d) for d in deps):
        build.compile_cpp_app([SRC], EXE)
    return EXE


def write_dump(path, sc, W, H, spp, depth, seed_mode, integrator, sun):
    """Flat scene dump read by tests/cpp/renderer_api.cpp (format in its header)."""
    I = 0 if sc.instance_mesh is None else len(sc.instance_mesh)
    T, M = len(sc.positions), len(sc.mesh_first)
    hdr = np.array([T, M, I, len(sc.materials), len(sc.textures), W, H, spp, depth, seed_mode, integrator], np.int32)
    mats = []
    for m in sc.materials:
        mats += [m["albedo_texture_id"], *m["albedo"], m["orm_texture_id"], m["metallic"], m["roughness"],
                 m["transmission_texture_id"], m["transmission"], m["ior"], m["emission_texture_id"],
                 *m["emission"], m["emission_scale_nits"], m["normal_texture_id"], m["normal_scale"], 0.0]
    tex = []
    pix = []
    for (w, h, ch, off) in sc.textures:
        tex += [w, h, ch]
        pix.append(np.asarray(sc.pixels[off:off + w * h * ch], np.uint8))
    pix = np.concatenate(pix) if pix else np.zeros(0, np.uint8)
    proj, view = view_proj(sc, W, H)
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(np.asarray(sun, np.float32).tobytes())
        for a, dt in ((sc.positions, np.float32), (sc.normals, np.float32), (sc.uvs, np.float32),
                      (sc.material_ids, np.int32), (sc.mesh_first, np.int32), (sc.mesh_count, np.int32),
                      (sc.mesh_model, np.float32)):
            f.write(np.ascontiguousarray(a, dt).tobytes())
        if I:
            f.write(np.ascontiguousarray(sc.instance_mesh, np.int32).tobytes())
            f.write(np.ascontiguousarray(sc.instance_model, np.float32).tobytes())
        f.write(np.asarray(mats, np.float32).tobytes())
        f.write(np.asarray(tex, np.uint32).tobytes())
        f.write(np.array([pix.size], np.uint64).tobytes())
        f.write(pix.tobytes())
        f.write(np.asarray(proj, np.float32).tobytes())
        f.write(np.asarray(view, np.float32).tobytes())


def test_cpp_header_compiles_and_links():
    exe = _exe()
    assert os.access(exe, os.X_OK)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["plain", "instanced"])
def test_cpp_renderer_matches_binding(tmp_path, which):
    from kittlespt_amd.context import Context
    sc = scenes.scene_plain(subdiv=3) if which == "plain" else scenes.scene_instanced_small()
    W, H, spp, depth = 64, 48, 3, 4
    sun = (0.0087, 4.18879, float(sc.sun_theta), 1.0)
    dump, out = str(tmp_path / "scene.bin"), str(tmp_path / "out.bin")
    write_dump(dump, sc, W, H, spp, depth, abi.SEED_XOR, abi.INTEGRATOR_WAVEFRONT, sun)
    print(f"[cpp api] {which} {W}x{H} spp {spp}", flush=True)
    r = subprocess.run([_exe(), dump, out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(out, np.float32)
    got = raw[:2 * H * W * 4].reshape(2, H, W, 4)
    batch = raw[2 * H * W * 4:]

    ctx = Context(0)
    try:
        ctx.upload(sc)
        ctx.set_settings(settings(max_depth=depth, integrator=abi.INTEGRATOR_WAVEFRONT, sun_theta=sun[2]))
        ctx.resize(W, H)
        p, v = view_proj(sc, W, H)
        ctx.set_view(p, v)
        ctx.set_frame_delta(16.0)
        ctx.render(spp)
        img, vb = ctx.readback(4), ctx.readback(3)
        views = ctx.render_views([(p, v, W, H), (p, v, W // 2, H // 2)], spp=spp)
    finally:
        ctx.close()
    assert np.array_equal(got[0].view(np.uint32), img.view(np.uint32))
    dbg = vb.copy()
    dbg[..., 2], dbg[..., 3] = 0.0, 1.0          # the debug texture: (velocity.xy, 0, 1)
    assert np.array_equal(got[1].view(np.uint32), dbg.view(np.uint32))
    want = np.concatenate([x.reshape(-1) for x in views])
    assert batch.shape == want.shape and np.array_equal(batch.view(np.uint32), want.view(np.uint32))
