"""GPU parity of the post chain (kpt_post.hip) against the oracle's
restatement (oracle/kpt_oracle_post.cpp): modulation, histogram + auto
exposure, bloom and the AgX film transform, frame after frame, bit-exact.

The oracle is fed the GPU's own demodulated estimate and G-buffer of each
frame (both already checked against the oracle's path tracer in
test_gpu_parity.py), so these tests isolate the post chain.
"""
import numpy as np
import pytest

import oracle
from kittlespt_amd import abi, scenes
from common import settings, view_proj

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def plain_ctx():
    from kittlespt_amd.context import Context
    sc = scenes.scene_plain(subdiv=3)
    ctx = Context(0)
    ctx.upload(sc)
    yield sc, ctx
    ctx.close()


def _run(sc, ctx, W, H, frames, ae, bloom, karis, dt=16.0, exposure=None):
    st = settings(max_depth=3, integrator=abi.INTEGRATOR_WAVEFRONT, sun_theta=sc.sun_theta)
    st.enable_auto_exposure, st.bloom_generate_bloom, st.bloom_use_karis_average = ae, bloom, karis
    ctx.set_settings(st)
    ctx.resize(W, H)
    p, v = view_proj(sc, W, H)
    ctx.set_view(p, v)
    # daylight exposure (f/16, 1/1000 s, ISO 100) so the sky is not clipped
    ctx.set_exposure(exposure or abi.KptExposure.camera(aperture=16.0, iso=100, shutter=1 / 1000))
    ctx.set_frame_delta(dt)
    e, film = ctx.exposure(), ctx.film()
    for f in range(frames):
        print(f"[post] {W}x{H} ae={ae} bloom={bloom} karis={karis} frame {f}", flush=True)
        ctx.render(1)
        est = np.zeros((H, W, 4), np.float32)
        est[..., :3] = ctx.estimate()
        gb = ctx.readback(2)
        r_o, out_o, _, _ = oracle.post_frame(est, gb, st, e, film, dt)
        r_g, out_g = ctx.readback(0), ctx.readback(4)
        assert np.array_equal(r_g.view(np.uint32), r_o.view(np.uint32)), "modulated render differs"
        diff = (out_g.view(np.uint32) != out_o.view(np.uint32)).any(-1)
        assert not diff.any(), f"{int(diff.sum())} output pixels differ, max abs {np.abs(out_g - out_o).max()}"
        fg, eg = ctx.film(), ctx.exposure()
        assert fg.as_dict() == film.as_dict()
        assert eg.as_dict() == e.as_dict()
    return out_g


@pytest.mark.parametrize("ae,bloom,karis", [(0, 0, 0), (0, 1, 0), (0, 1, 1), (1, 0, 0), (1, 1, 1)])
def test_post_chain_parity(plain_ctx, ae, bloom, karis):
    sc, ctx = plain_ctx
    out = _run(sc, ctx, 160, 96, 3, ae, bloom, karis)
    # AgX output is display-referred: [0, 1] up to rounding of the outset matrix
    assert np.isfinite(out).all() and (out[..., :3] >= 0).all() and (out[..., :3] <= 1.001).all()
    assert out[..., :3].std() > 0.01                    # an image, not a clipped frame


def test_post_chain_odd_size_bloom_and_adaptation(plain_ctx):
    sc, ctx = plain_ctx
    # 97x61: five mip levels with odd extents; slow adaptation (dt 1 ms) over 4 frames
    _run(sc, ctx, 97, 61, 4, 1, 1, 0, dt=1.0,
         exposure=abi.KptExposure.camera(aperture=4.0, iso=400, shutter=1 / 125, ev_comp=0.5))


@pytest.mark.parametrize("W,H", [(160, 96), (97, 61), (64, 32)])
def test_effects_pass_parity(plain_ctx, W, H):
    """postprocess_enable_effects: the output is computeEffects of the back
    buffer (oracle.post_effects), bit for bit; the back buffer is the
    effects-off output of the same frame (64x32 has the centre pixel whose
    fisheye coordinates are NaN)."""
    sc, ctx = plain_ctx
    outs = []
    for eff in (0, 1):
        st = settings(max_depth=3, integrator=abi.INTEGRATOR_WAVEFRONT, sun_theta=sc.sun_theta)
        st.bloom_generate_bloom = 1
        st.postprocess_enable_effects = eff
        ctx.set_settings(st)
        ctx.resize(W, H)
        p, v = view_proj(sc, W, H)
        ctx.set_view(p, v)
        ctx.set_exposure(abi.KptExposure.camera(aperture=16.0, iso=100, shutter=1 / 1000))
        ctx.set_frame_delta(16.0)
        print(f"[fx] {W}x{H} effects={eff}", flush=True)
        ctx.render(1)
        outs.append(ctx.readback(4))
    want = oracle.post_effects(outs[0])
    diff = (outs[1].view(np.uint32) != want.view(np.uint32)).any(-1)
    assert not diff.any(), f"{int(diff.sum())} pixels differ"
    assert not np.array_equal(outs[1], outs[0])          # the pass changed the image


@pytest.mark.parametrize("ae,bloom,fsr", [(0, 0, 0), (0, 1, 0), (1, 1, 0), (0, 0, 1), (1, 1, 1)])
def test_upscale_frames_parity(plain_ctx, ae, bloom, fsr):
    """upscale_enable: the path trace, modulation and histogram at
    int(W/2.5) x int(H/2.5), the bilinear upscale (or, with upscale_fsr, FSR
    1.0's EASU + RCAS) to W x H, then bloom and the post pass at the output
    size — against the oracle's post chain at the two sizes, frame after
    frame."""
    sc, ctx = plain_ctx
    W, H = 160, 96
    st = settings(max_depth=3, integrator=abi.INTEGRATOR_WAVEFRONT, sun_theta=sc.sun_theta)
    st.enable_auto_exposure, st.bloom_generate_bloom = ae, bloom
    st.upscale_enable = 1
    st.upscale_fsr = fsr
    ctx.set_settings(st)
    ctx.resize(W, H)
    rw, rh, ow, oh = ctx.resolution()
    assert (rw, rh, ow, oh) == (int(np.float32(W) / np.float32(2.5)), int(np.float32(H) / np.float32(2.5)), W, H)
    p, v = view_proj(sc, rw, rh)
    ctx.set_view(p, v)
    ctx.set_exposure(abi.KptExposure.camera(aperture=16.0, iso=100, shutter=1 / 1000))
    ctx.set_frame_delta(16.0)
    e, film = ctx.exposure(), ctx.film()
    for f in range(3):
        print(f"[upscale] ae={ae} bloom={bloom} fsr={fsr} frame {f}", flush=True)
        ctx.render(1)
        est = np.zeros((rh, rw, 4), np.float32)
        est[..., :3] = ctx.estimate()
        gb = ctx.readback(2)
        r_o, out_o, _, _ = oracle.post_frame(est, gb, st, e, film, 16.0, out_size=(W, H))
        r_g, out_g = ctx.readback(0), ctx.readback(4)
        assert r_g.shape == (rh, rw, 4) and out_g.shape == (H, W, 4)
        assert np.array_equal(r_g.view(np.uint32), r_o.view(np.uint32)), "modulated render differs"
        diff = (out_g.view(np.uint32) != out_o.view(np.uint32)).any(-1)
        assert not diff.any(), f"{int(diff.sum())} output pixels differ"
    st.upscale_enable = 0
    st.upscale_fsr = 0
    ctx.set_settings(st)


def test_resize_recomputes_the_projection(plain_ctx):
    """Renderer::resizeResolution (renderer.cu:78-84): the projection becomes
    perspectiveFovLH(fov, render width, render height, 0.1, 100) with the
    field of view of the one set before."""
    sc, ctx = plain_ctx
    st = settings(max_depth=3, integrator=abi.INTEGRATOR_WAVEFRONT, sun_theta=sc.sun_theta)
    ctx.set_settings(st)
    ctx.resize(64, 48)
    fov = 0.9
    p = scenes.perspective_fov_lh(fov, 64, 48, 1.0, 100.0)
    _, v = view_proj(sc, 64, 48)
    ctx.set_view(p, v)
    ctx.resize(100, 40)
    want = np.linalg.inv(scenes.perspective_fov_lh(fov, 100, 40, 0.1, 100.0).reshape(4, 4).astype(np.float64))
    got = np.array(ctx.camera().inv_proj, np.float64).reshape(4, 4)
    assert np.allclose(got, want, rtol=1e-5, atol=1e-6)
