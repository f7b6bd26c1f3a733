"""Pins the oracle's temporal accumulation (reprojectAccumulate /
testReprojection, kernels.cu:196-338) with properties the reference's
definition fixes:

* static camera and scene: zero motion vectors, every surface pixel passes the
  depth test, and the history blend alpha = 1 / min(n + 1, 16) makes the
  result the running mean of the samples since the first accepted frame (the
  very first frame's sample has history 0 and alpha 1, so it is replaced);
* the history length counts accepted frames; sky pixels (no instance) and
  pixels reprojected off screen restart at 0 with the current sample;
* after 16 accepted frames the blend is an exponential average with 1/16.
"""
import numpy as np
import pytest

import oracle
from kittlespt_amd import abi, scenes
from common import host_camera, settings


@pytest.fixture(scope="module")
def plain():
    sc = scenes.scene_plain(subdiv=2)
    orc = oracle.OracleScene(sc)
    yield sc, orc
    orc.close()


def _static_cam(sc, W, H):
    cam = host_camera(sc, W, H)
    cam.prev_inv_view[:] = cam.inv_view[:]
    cam.prev_inv_proj[:] = cam.inv_proj[:]
    return cam


def _run(orc, st, cam, W, H, frames):
    acc = np.zeros((H, W, 4), np.float32)
    gb = np.zeros((H, W, 4), np.uint32)
    outs = []
    for f in range(frames):
        acc, gb, vb = orc.render_temporal(st, cam, W, H, f, acc, gb)
        outs.append((acc.copy(), gb.copy(), vb.copy()))
    return outs


def test_static_scene_is_running_mean(plain):
    sc, orc = plain
    W, H = 24, 16
    st = settings(max_depth=3, sun_theta=sc.sun_theta)
    st.use_temporal_accumulation = 1
    cam = _static_cam(sc, W, H)
    outs = _run(orc, st, cam, W, H, 6)
    r, _, v = outs[-1]
    inst = v[..., 3].reshape(-1)
    # motion vectors vanish up to one ulp of uv (the previous-frame path goes
    # through the model inverse and back)
    assert np.abs(v[..., :2]).max() <= 2.4e-7
    still = (v[..., :2] == 0).all(-1).reshape(-1)
    surf = inst >= 0
    assert surf.sum() > 50 and (~surf).sum() > 20 and (surf & still).sum() > 150
    hist = r[..., 3].reshape(-1)
    assert (hist[surf & still] == 5).all() and (hist[~surf] == 0).all()
    # an ulp of motion moves the truncated depth-test tap to a neighbour: such
    # pixels are kept (history 5) or restart (history 0), nothing in between
    assert np.isin(hist[surf & ~still], [0, 5]).all()
    pix = np.arange(W * H, dtype=np.int32)
    samples = np.stack([orc.sample_pixels(st, cam, W, H, pix, f) for f in range(6)]).astype(np.float64)
    mean15 = samples[1:].mean(0)
    got = r[..., :3].reshape(-1, 3).astype(np.float64)
    keep = surf & still
    assert np.allclose(got[keep], mean15[keep], rtol=2e-5, atol=1e-6)
    restart = hist == 0
    assert np.array_equal(got[restart], samples[5][restart])  # sky / rejected: the current sample


def test_history_saturates_at_16(plain):
    sc, orc = plain
    W, H = 8, 6
    st = settings(max_depth=2, sun_theta=sc.sun_theta)
    st.use_temporal_accumulation = 1
    cam = _static_cam(sc, W, H)
    outs = _run(orc, st, cam, W, H, 20)
    hist = outs[-1][0][..., 3]
    v = outs[-1][2]
    assert hist.max() == 19 and (hist[v[..., 3] >= 0] == 19).all()
    # exponential average with alpha 1/16 once the history is >= 15
    pix = np.arange(W * H, dtype=np.int32)
    s19 = orc.sample_pixels(st, cam, W, H, pix, 19).astype(np.float64)
    prev = outs[-2][0][..., :3].reshape(-1, 3).astype(np.float64)
    want = prev + (s19 - prev) / 16.0
    surf = v[..., 3].reshape(-1) >= 0
    assert np.allclose(outs[-1][0][..., :3].reshape(-1, 3)[surf], want[surf], rtol=1e-5, atol=1e-7)


def test_camera_motion_rejects_offscreen(plain):
    sc, orc = plain
    W, H = 32, 24
    st = settings(max_depth=2, sun_theta=sc.sun_theta)
    st.use_temporal_accumulation = 1
    cam = _static_cam(sc, W, H)
    acc, gb, _ = orc.render_temporal(st, cam, W, H, 0, np.zeros((H, W, 4), np.float32), np.zeros((H, W, 4), np.uint32))
    acc, gb, _ = orc.render_temporal(st, cam, W, H, 1, acc, gb)
    # pan: the new camera looks to the side; the previous camera is the old one
    sc2 = scenes.scene_plain(subdiv=2)
    sc2.camera_target = (sc.camera_target[0] + 0.6, sc.camera_target[1], sc.camera_target[2])
    cam2 = host_camera(sc2, W, H)
    cam2.prev_inv_view[:] = cam.inv_view[:]
    cam2.prev_inv_proj[:] = cam.inv_proj[:]
    r, _, v = orc.render_temporal(st, cam2, W, H, 2, acc, gb)
    vel = v[..., :2].reshape(-1, 2)
    inst = v[..., 3].reshape(-1)
    ys, xs = np.divmod(np.arange(W * H), W)
    ppx = xs.astype(np.float32) - vel[:, 0] * np.float32(W)
    ppy = ys.astype(np.float32) - vel[:, 1] * np.float32(H)
    off = (ppx < 0) | (ppx >= W) | (ppy < 0) | (ppy >= H)
    surf = inst >= 0
    assert (surf & off).sum() > 10 and (surf & ~off).sum() > 10
    hist = r[..., 3].reshape(-1)
    assert (hist[surf & off] == 0).all()
    assert (hist[surf & ~off] <= 2).all() and (hist[surf & ~off] == 2).mean() > 0.5
