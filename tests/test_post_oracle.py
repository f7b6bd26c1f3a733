"""Pins the oracle's post chain (oracle/kpt_oracle_post.cpp) against closed
forms, textbook values and invariants, each chosen so a plausible slip (a
transposed matrix, a dropped term, a wrong tap weight or offset, a wrong bin
mapping) fails one of them.

Reference routines: histogram.cu:18-153, autoexposure_program.hpp:18-94,
renderer.cu:312-329, bloom.cu:6-145, mipchain.cuh:44-50,
device_texture_buffer.cuh:37-71, color.cuh:33-82, 409-424, film.cuh:35-188,
kernels.cu:453-504.
"""
import math

import numpy as np
import pytest

import oracle
from kittlespt_amd import abi

WP, BP = 6.5, -10.0


# ---------------------------------------------------------------- histogram
def test_lum_to_bin_closed_form():
    rng = np.random.default_rng(0)
    lum = np.exp(rng.uniform(math.log(1e-4), math.log(1e5), 20000)).astype(np.float32)
    got = oracle.lum_to_bin(lum, BP, WP).astype(np.int64)
    l2 = np.log2(lum.astype(np.float64))
    t = np.clip((l2 - BP) / (WP - BP), 0, 1) * 254 + 1
    want = np.where(lum < np.float32(0.005), 0, np.floor(t)).astype(np.int64)
    frac = t - np.floor(t)
    away = (frac > 1e-4) & (frac < 1 - 1e-4)          # rounding decides only at bin edges
    assert np.array_equal(got[away], want[away])
    assert (np.abs(got - want) <= 1).all()
    # specials: below epsilon -> 0; at or above the white point -> 255; black point -> 1
    assert list(oracle.lum_to_bin([0.0, 0.004999, 0.005], BP, WP)) == [0, 0, 37]
    assert list(oracle.lum_to_bin([2.0 ** WP, 1e30, np.inf], BP, WP)) == [255, 255, 255]
    assert oracle.lum_to_bin([2.0 ** BP * 8], BP, WP)[0] == 1 + int(3 / 16.5 * 254)


def test_meter_weight_closed_form():
    W, H = 640, 360
    rng = np.random.default_rng(1)
    xy = np.stack([rng.integers(0, W, 4000), rng.integers(0, H, 4000)], 1).astype(np.int32)
    got = oracle.meter_weight(W, H, xy).astype(np.float64)
    d = np.hypot(W // 2 - xy[:, 0], H // 2 - xy[:, 1]) / 0.4 / (min(W, H) / 2)
    y = np.clip(d, 0, 1)
    want = 1 - y * y * (3 - 2 * y)
    assert np.allclose(got, want, rtol=0, atol=2e-6)
    assert oracle.meter_weight(W, H, [[W // 2, H // 2]])[0] == 1.0
    assert oracle.meter_weight(W, H, [[0, 0]])[0] == 0.0            # beyond 0.4 * half height


def _const_image(H, W, rgb):
    img = np.zeros((H, W, 4), np.float32)
    img[..., :3] = rgb
    img[..., 3] = 1
    return img


def test_histogram_constant_and_black_images():
    H, W = 48, 80
    img = _const_image(H, W, (0.18, 0.18, 0.18))
    bins = oracle.histogram(img, WP, BP)
    b = int(oracle.lum_to_bin([np.float32(0.2126) * 0.18 + np.float32(0.7152) * 0.18 + np.float32(0.0722) * 0.18],
                              BP, WP)[0])
    assert np.count_nonzero(bins) == 1 and bins[b] > 0
    xy = np.stack(np.meshgrid(np.arange(W), np.arange(H)), -1).reshape(-1, 2)
    w = oracle.meter_weight(W, H, xy).astype(np.float64)
    assert abs(int(bins[b]) - np.round(w * 2 ** 23).sum()) <= len(w)
    black = oracle.histogram(_const_image(H, W, (0, 0, 0)), WP, BP)
    assert black[0] == W * H * 2 ** 23 and np.count_nonzero(black) == 1   # black pixels count 1 each


def test_histogram_average_closed_forms():
    for b in (1, 37, 116, 200, 255):
        bins = np.zeros(256, np.uint64)
        bins[b] = 12345 * 2 ** 23
        got = oracle.histogram_average(bins, WP, BP, last=0.0, dt_ms=100.0)   # 1 - e^-50 == 1 in float
        want = 2.0 ** ((b - 1) / 254 * (WP - BP) + BP)
        assert got == pytest.approx(want, rel=2e-6)
    # weighted mean of two bins
    bins = np.zeros(256, np.uint64)
    bins[50], bins[150] = 3 * 2 ** 23, 1 * 2 ** 23
    got = oracle.histogram_average(bins, WP, BP, 0.0, 100.0)
    assert got == pytest.approx(2.0 ** ((75 - 1) / 254 * 16.5 - 10), rel=2e-6)
    # black pixels do not move the mean; all-black frames read as the black point
    bins[0] = 10 ** 6 * 2 ** 23
    assert oracle.histogram_average(bins, WP, BP, 0.0, 100.0) == pytest.approx(got, rel=1e-6)
    only_black = np.zeros(256, np.uint64)
    only_black[0] = 5 * 2 ** 23
    assert oracle.histogram_average(only_black, WP, BP, 0.0, 100.0) == 2.0 ** BP
    # adaptation: lerp(last, avg, 1 - exp(-dt * 0.5))
    bins = np.zeros(256, np.uint64)
    bins[116] = 2 ** 23
    avg = 2.0 ** ((116 - 1) / 254 * 16.5 - 10)
    got = oracle.histogram_average(bins, WP, BP, last=0.5, dt_ms=2.0)
    assert got == pytest.approx(0.5 + (1 - math.exp(-1.0)) * (avg - 0.5), rel=2e-6)


# ---------------------------------------------------------------- exposure
def _program_closed_form(N, iso0, t0, L, comp=0.0, min_iso=100, max_iso=6400, min_t=0.00025, max_t=0.0333):
    ev = math.log2(L * 100 / 12.5) - comp + 6
    t = 0.01
    iso = min(max(N * N * 100 / (t * 2 ** ev), min_iso), max_iso)
    evd = ev - math.log2(N * N * 100 / (t * iso))
    t = min(max(t * 2 ** -evd, min_t), max_t)
    scalar = 0.18 / ((10 / 0.65) * (N * N / (iso * t)))
    return iso, t, scalar


@pytest.mark.parametrize("L", [1e-3, 0.02, 0.3, 5.0, 300.0])
def test_exposure_program_closed_form(L):
    e = abi.KptExposure.camera(aperture=2.8, iso=100, shutter=1 / 60)
    oracle.exposure_program(e, L)
    f = oracle.set_exposure(e)
    iso, t, scalar = _program_closed_form(2.8, 100, 1 / 60, L)
    assert e.iso == pytest.approx(iso, rel=1e-5)
    assert e.shutter_speed_secs == pytest.approx(t, rel=1e-5)
    assert f.luminance_exposure_scalar == pytest.approx(scalar, rel=1e-5)
    assert (f.white_point_ev, f.black_point_ev) == (WP, BP)


def test_standard_output_exposure_known_answer():
    # ISO 100, f/2.8, 1/60 s: 0.18 / ((10/0.65) * 2.8^2 / (100/60))
    f = oracle.set_exposure(abi.KptExposure.camera(2.8, 100, 1 / 60))
    assert f.luminance_exposure_scalar == pytest.approx(0.18 / ((10 / 0.65) * 7.84 / (100 / 60)), rel=1e-6)


# ---------------------------------------------------------------- colour and film
def test_yxy_round_trip_and_exposure_linearity():
    rng = np.random.default_rng(2)
    rgb = rng.uniform(0.01, 10.0, (500, 3)).astype(np.float32)
    yxy, back = oracle.yxy_roundtrip(rgb, 1.0)
    scale = rgb.max(1, keepdims=True).astype(np.float64)   # float cancellation is relative to the row
    assert (np.abs(back - rgb) <= 4e-6 * scale).all()
    y = rgb.astype(np.float64) @ np.array([0.2126729, 0.7151522, 0.0721750])
    assert np.allclose(yxy[:, 0], y, rtol=2e-6)
    xyz = rgb.astype(np.float64) @ np.array([[0.4124564, 0.2126729, 0.0193339], [0.3575761, 0.7151522, 0.1191920],
                                              [0.1804375, 0.0721750, 0.9503041]])
    assert np.allclose(yxy[:, 1], xyz[:, 0] / xyz.sum(1), rtol=2e-6)
    _, scaled = oracle.yxy_roundtrip(rgb, 0.37)
    assert (np.abs(scaled - 0.37 * rgb.astype(np.float64)) <= 4e-6 * scale).all()


def test_agx_curve_and_srgb_known_answers():
    c = oracle.agx_curve(np.array([0.0, 1.0], np.float32))
    assert c[0] == np.float32(-0.00232)
    assert c[1] == pytest.approx(15.5 - 40.14 + 31.96 - 6.868 + 0.4298 + 0.1191 - 0.00232, rel=1e-6)
    x = np.linspace(0, 1, 4001, dtype=np.float32)
    assert (np.diff(oracle.agx_curve(x).astype(np.float64)) >= 0).all()
    s = oracle.srgb_encode(np.array([0.0, 0.0031308, 0.5, 1.0], np.float32))
    assert s[0] == 0.0
    assert s[1] == pytest.approx(12.92 * 0.0031308, rel=1e-6)
    assert s[2] == pytest.approx(0.7353569830524495, rel=2e-6)        # IEC 61966-2-1
    assert s[3] == pytest.approx(1.0, abs=2e-7)


def test_film_invariants():
    g = np.linspace(0.001, 30.0, 300, dtype=np.float32)
    out = oracle.film(np.stack([g, g, g], 1))
    # neutral stays neutral (catches a transposed 709->2020, inset or outset matrix)
    assert np.abs(out - out.mean(1, keepdims=True)).max() < 3e-4
    assert (np.diff(out[:, 1].astype(np.float64)) >= 0).all()
    assert (out >= 0).all() and (out <= 1.0).all()
    # black maps to black (the curve's -0.00232 turns NaN in the look and is clamped)
    assert (oracle.film([[0.0, 0.0, 0.0]]) == 0).all()
    # everything above 2^white * 0.18 clips to the same code value
    hi = oracle.film([[20.0] * 3, [1000.0] * 3])
    assert np.array_equal(hi[0], hi[1])
    # the exposure scalar is the same as scaling the input
    rng = np.random.default_rng(3)
    rgb = rng.uniform(0.01, 2, (200, 3)).astype(np.float32)
    a = oracle.film(rgb, scalar=0.5)
    b = oracle.film(rgb * np.float32(0.5), scalar=1.0)
    assert np.allclose(a, b, atol=2e-5)


# ---------------------------------------------------------------- texture reads and bloom
def test_bilinear_semantics():
    img = np.zeros((4, 5, 4), np.float32)
    img[..., 0] = np.arange(20, dtype=np.float32).reshape(4, 5)
    img[..., 3] = np.arange(20, dtype=np.float32).reshape(4, 5) * 10
    q = oracle.bilinear(img, [[2, 1], [2.5, 1], [2, 1.5], [10, 10], [-0.5, 0]])
    assert q[0, 0] == 7 and q[1, 0] == 7.5 and q[2, 0] == 9.5
    assert q[3, 0] == 19                              # clamped to the last texel
    assert q[4, 0] == 0 + -0.5 * (1 - 0)              # truncation toward zero extrapolates
    pa = oracle.bilinear(img, [[2.5, 1.5]], filter_alpha=False)
    assert pa[0, 3] == img[1, 2, 3]                   # alpha point-sampled from the first tap


def test_bloom_down_impulse_response():
    H, W = 32, 32
    for (sy, sx), total in (((10, 10), 0.5), ((11, 11), 0.5), ((10, 11), 0.0), ((11, 10), 0.0)):
        img = np.zeros((H, W, 4), np.float32)
        img[sy, sx, :3] = 1.0
        d = oracle.bloom_down(img)
        assert d[..., 0].sum() == pytest.approx(total, abs=1e-7)
    img = np.zeros((H, W, 4), np.float32)
    img[10, 10, :3] = 1.0
    d = oracle.bloom_down(img)[..., 0]
    assert d[5, 5] == 0.125                       # centre tap e in four 4x4 boxes
    assert d[5, 4] == d[5, 6] == d[4, 5] == d[6, 5] == 0.0625
    assert d[4, 4] == d[4, 6] == d[6, 4] == d[6, 6] == 0.03125
    img = np.zeros((H, W, 4), np.float32)
    img[11, 11, :3] = 1.0
    d = oracle.bloom_down(img)[..., 0]
    assert d[5, 5] == d[5, 6] == d[6, 5] == d[6, 6] == 0.125   # diagonal taps j k l m


def test_bloom_constant_images_and_mass():
    c = _const_image(32, 48, (0.7, 0.2, 3.0))
    assert np.array_equal(oracle.bloom_down(c)[..., :3], c[:16, :24, :3])
    k = oracle.bloom_down(c, karis=True)[..., :3]
    assert np.allclose(k, c[:16, :24, :3], rtol=1e-6)
    up = oracle.bloom_up(c[:16, :24], _const_image(32, 48, (0.0, 0.0, 0.0)), 0.65)
    assert np.allclose(up[..., :3], np.float32(0.65) * c[..., :3], rtol=1e-6)
    src = np.zeros((16, 16, 4), np.float32)
    src[8, 7, :3] = 1.0
    up = oracle.bloom_up(src, np.zeros((32, 32, 4), np.float32), 0.65)
    assert up[..., 0].sum() == pytest.approx(0.65 * 4.0, rel=1e-6)
    assert oracle.bloom_levels(1920, 1080) == 6 and oracle.bloom_levels(100, 37) == 5


# ---------------------------------------------------------------- whole frame
def _frame_inputs(H=40, W=64, seed=4):
    rng = np.random.default_rng(seed)
    est = np.zeros((H, W, 4), np.float32)
    est[..., :3] = rng.exponential(0.5, (H, W, 3))
    est[..., 3] = 1
    gb = np.zeros((H, W, 4), np.uint32)
    alb = rng.integers(0, 256, (H, W, 3)).astype(np.uint32)
    gb[..., 2] = alb[..., 0] | (alb[..., 1] << 8) | (alb[..., 2] << 16)
    return est, gb, alb


def _settings(ae=0, bloom=0, karis=0):
    s = abi.KptSettings()
    s.enable_auto_exposure, s.bloom_generate_bloom, s.bloom_use_karis_average = ae, bloom, karis
    s.bloom_final_blend, s.bloom_internal_blend = 0.175, 0.65
    return s


def test_post_frame_composition():
    est, gb, alb = _frame_inputs()
    e = abi.KptExposure.camera()
    film = oracle.set_exposure(e)
    render, out, _, _ = oracle.post_frame(est, gb, _settings(), e, film)
    want_render = est[..., :3] * (alb.astype(np.float32) / np.float32(255.0))
    assert np.array_equal(render[..., :3], want_render) and (render[..., 3] == 1).all()
    f = oracle.film(render[..., :3].reshape(-1, 3), film.luminance_exposure_scalar, WP, BP)
    assert np.array_equal(out[..., :3].reshape(-1, 3), f)
    # bloom: mip chain over the modulated render, blended before exposure
    e2 = abi.KptExposure.camera()
    film2 = oracle.set_exposure(e2)
    render2, out2, bloom, _ = oracle.post_frame(est, gb, _settings(bloom=1), e2, film2)
    mix = render2[..., :3] + np.float32(0.175) * (bloom[..., :3] - render2[..., :3])
    f2 = oracle.film(mix.reshape(-1, 3), film2.luminance_exposure_scalar, WP, BP)
    assert np.array_equal(out2[..., :3].reshape(-1, 3), f2)
    # auto exposure: histogram of the modulated render -> average -> program -> film
    e3 = abi.KptExposure.camera()
    film3 = oracle.set_exposure(e3)
    render3, out3, _, hist = oracle.post_frame(est, gb, _settings(ae=1), e3, film3, frame_delta_ms=16.0)
    assert np.array_equal(hist, oracle.histogram(render3, WP, BP))
    avg = oracle.histogram_average(hist, WP, BP, 0.0, 16.0)
    assert film3.scene_average_luminance == avg
    e_chk = abi.KptExposure.camera()
    oracle.exposure_program(e_chk, avg)
    assert (e3.iso, e3.shutter_speed_secs) == (e_chk.iso, e_chk.shutter_speed_secs)
    assert film3.luminance_exposure_scalar == oracle.set_exposure(e_chk).luminance_exposure_scalar


# ---------------------------------------------------------------- effects pass
# computeEffects (kernels.cu:508-528) with getFishEyeUV (kernels.cu:340-376):
# pinned by what the mapping fixes, not by retyping it.

def _ramp(H, W, along="x"):
    x = np.arange(W, dtype=np.float32)[None, :].repeat(H, 0)
    y = np.arange(H, dtype=np.float32)[:, None].repeat(W, 1)
    v = x if along == "x" else y
    return np.stack([v, v, v, np.ones_like(v)], -1).astype(np.float32)


def test_effects_constant_image_is_unchanged():
    b = np.full((37, 53, 4), 0.25, np.float32)
    o = oracle.post_effects(b)
    assert np.array_equal(o[..., :3], b[..., :3]) and (o[..., 3] == 1.0).all()


@pytest.mark.parametrize("W,H", [(16, 8), (64, 48), (97, 61)])
def test_effects_corner_maps_to_itself(W, H):
    # at the corner pixel d = -m, |d| = bind, so tan(r*k)*bind/tan(bind*k) = bind
    # and uv = m + d = (0, 0): the green tap reads the source corner
    b = np.random.default_rng(1).random((H, W, 4)).astype(np.float32)
    o = oracle.post_effects(b)
    assert o[0, 0, 1] == b[0, 0, 1]


def test_effects_centre_pixel_reads_the_origin():
    # W = 2H: the centre pixel has p == m exactly, d = 0, normalize(0) is NaN,
    # and a NaN coordinate converts to pixel 0 as the device converts it
    H, W = 8, 16
    b = np.random.default_rng(2).random((H, W, 4)).astype(np.float32)
    o = oracle.post_effects(b)
    assert np.array_equal(o[H // 2, W // 2, :3], b[0, 0, :3])


@pytest.mark.parametrize("along", ["x", "y"])
def test_effects_chromatic_order_on_a_ramp(along):
    # R samples at uv + offset, B at uv - offset, offset >= 0: on an image that
    # grows along x (or y) the channels come out ordered R >= G >= B, and the
    # spread is widest at the image border, zero at the centre column/row
    H, W = 160, 200
    b = _ramp(H, W, along)
    o = oracle.post_effects(b)
    assert (o[..., 0] >= o[..., 1]).all() and (o[..., 1] >= o[..., 2]).all()
    spread = o[..., 0] - o[..., 2]
    # at the border the offset is ~0.01 of the size (2 px each way here)
    if along == "x":
        assert spread[:, :3].mean() >= 3.0 and spread[:, W // 2 - 1:W // 2 + 2].mean() <= 1.0
    else:
        assert spread[:3, :].mean() >= 2.5 and spread[H // 2 - 1:H // 2 + 2, :].mean() <= 1.0


def test_effects_fisheye_is_radially_monotone():
    # along the centre row the green source column grows with the pixel column
    # (the mapping tan(r k) / tan(bind k) is increasing on [0, bind])
    H, W = 48, 64
    b = _ramp(H, W, "x")
    o = oracle.post_effects(b)
    g = o[H // 2 - 1, :, 1]            # (the centre row's centre pixel is the NaN case)
    assert (np.diff(g) >= 0).all()
    # the corner distance is the radius kept fixed: along the centre row the
    # edges pull inward (source columns inside the frame)
    assert g[0] > 0 and g[-1] < W - 1


# ---------------------------------------------------------------- upscale
# upscale (upscale.cuh:287-292): output (x, y) reads the source bilinearly at
# (x / OW * W, y / OH * H).

def test_upscale_constant_and_corner():
    src = np.full((5, 7, 4), 0.5, np.float32)
    assert np.array_equal(oracle.upscale_bilinear(src, 17, 12), np.full((12, 17, 4), 0.5, np.float32))
    src = np.random.default_rng(3).random((5, 7, 4)).astype(np.float32)
    out = oracle.upscale_bilinear(src, 17, 12)
    assert np.array_equal(out[0, 0], src[0, 0])        # sample point (0, 0): weights 0


def test_upscale_reproduces_a_linear_ramp():
    # bilinear interpolation of a linear function is the function (up to
    # rounding) until the right-hand tap is clamped to the last column
    W, H, OW, OH = 8, 3, 20, 7
    x = np.arange(W, dtype=np.float32)
    src = np.stack([np.tile(x, (H, 1))] * 3 + [np.ones((H, W), np.float32)], -1)
    out = oracle.upscale_bilinear(src, OW, OH)
    s = np.arange(OW, dtype=np.float64) / OW * W
    want = np.where(s <= W - 1, s, W - 1)
    assert np.allclose(out[:, :, 0], want[None, :], atol=1e-5)
    assert np.allclose(out[..., 3], 1.0)


# ---------------------------------------------------------------- FSR 1.0
# The reference's USE_FSR build (upscale.cuh:13-408): EASU then RCAS with
# sharpness 0.  Pinned by what the algorithm fixes in closed form.

@pytest.mark.parametrize("size", [(7, 5, 17, 12), (16, 9, 40, 22), (8, 8, 8, 8)])
def test_fsr_constant_image_is_unchanged(size):
    # EASU normalises its weights and clamps to the 4 nearest taps; RCAS on a
    # constant ring: lobe = max(-3/16, min(max(-1/4, -1/4), 0)) = -3/16 and
    # (lobe * 4c + c) / (4 lobe + 1) = c
    W, H, OW, OH = size
    src = np.full((H, W, 4), 0.375, np.float32)
    out, easu = oracle.upscale_fsr(src, OW, OH, easu=True)
    assert np.array_equal(easu[..., :3], np.full((OH, OW, 3), 0.375, np.float32))
    assert np.array_equal(out[..., :3], np.full((OH, OW, 3), 0.375, np.float32))
    assert np.array_equal(out[..., 3], np.ones((OH, OW), np.float32))


def test_fsr_rcas_impulse_closed_form():
    # a pixel of value v with a black ring: mn4 = mx4 = 0 -> hitMin = 0/0 (NaN,
    # dropped by fmaxf), hitMax = (1 - 0) / (0 - 4) = -1/4, lobe = -3/16, out =
    # v / (1 - 3/4) = 4v; the pixel above sees mn4 = 0, mx4 = v -> hitMin = 0,
    # hitMax = (1 - v) / -4 < 0, lobe = 0, out = its own 0.  8 x 8 keeps the
    # reads' p / 8 * 8 exact.
    v = 0.1875
    src = np.zeros((8, 8, 4), np.float32)
    src[3, 3, :3] = v
    src[..., 3] = 1.0
    out = oracle.fsr_rcas(src)
    assert np.array_equal(out[3, 3, :3], np.full(3, 4 * v, np.float32))
    assert np.array_equal(out[2, 3, :3], np.zeros(3, np.float32))
    assert np.array_equal(out[4, 3, :3], np.zeros(3, np.float32))
    assert np.array_equal(out[0, 0, :3], np.zeros(3, np.float32))


def test_fsr_easu_at_one_to_one_centres_on_the_half_pixel():
    # at 1:1, con0 = (1, 1, 0, 0): pp = (x, y) has no fraction, and the four
    # nearest taps read textureReadBilinear at (x + 1/2, y + 1/2) ('f'), (x +
    # 3/2, y + 1/2) ('g'), (x + 1/2, y + 3/2) ('j'), (x + 3/2, y + 3/2) ('k')
    # (p0 = ((x+1)/W, (y-1)/H), p1 = p0 + (-1/W, 2/H), p2 = p0 + (1/W, 2/H),
    # the gather offsets -+1/2W, -+1/2H), i.e. with the truncating bilinear
    # the pixels x..x+2, y..y+2.  An image constant on 4 x 4 blocks therefore
    # has f = g = j = k at x, y = 0 mod 4, and the de-ring clamp returns that
    # block's value exactly; one block further right it is the next block's.
    W = H = 16
    rng = np.random.default_rng(8)
    blocks = rng.random((H // 4, W // 4)).astype(np.float32) * 0.5 + 0.25
    img = np.repeat(np.repeat(blocks, 4, 0), 4, 1)
    src = np.stack([img] * 3 + [np.ones_like(img)], -1)
    _, easu = oracle.upscale_fsr(src, W, H, easu=True)
    for y in range(0, H - 3, 4):
        for x in range(0, W - 3, 4):
            assert easu[y, x, 0] == img[y, x], (y, x)
            assert easu[y, x + 1, 0] == img[y, x], (y, x)       # pixels x+1..x+3 still inside


def test_fsr_easu_stays_within_the_nearest_taps_and_keeps_a_step_edge():
    # a vertical step edge 0 | 1 upscaled 2.5x: every output lies in [0, 1]
    # (dering clamp to the 4 nearest taps), every row is monotone across the
    # edge, the plateaus are exact, and the edge stays sharp: at most 4 output
    # pixels per row lie strictly between 0 and 1 (a bilinear upscale makes
    # 2.5 x 1 + 1 of them; a Gaussian-like kernel of EASU's support, more)
    W, H, OW, OH = 12, 6, 30, 15
    src = np.zeros((H, W, 4), np.float32)
    src[:, W // 2:, :3] = 1.0
    src[..., 3] = 1.0
    _, easu = oracle.upscale_fsr(src, OW, OH, easu=True)
    r = easu[..., 0]
    assert r.min() >= 0.0 and r.max() <= 1.0
    assert np.all(np.diff(r.astype(np.float64), axis=1) >= 0)
    assert np.all(r[:, :8] == 0.0) and np.all(r[:, -8:] == 1.0)
    assert np.all(((r > 0) & (r < 1)).sum(axis=1) <= 4)


def test_fsr_rcas_sharpens_the_knees_of_a_ramp():
    # a ramp 0.25 -> 0.75 along x (16 x 16: exact reads).  On its linear
    # interior the ring averages to the centre, so (lobe * 4e + e) / (4 lobe
    # + 1) = e; at the dark knee (e = d = b = h = 0.25, f = 0.375): mn4 = 0.25,
    # mx4 = 0.375, hitMin = 1/6, hitMax = 0.625 / -3, lobe = -1/6, out =
    # (-1/6 * 1.125 + 0.25) / (1/3) = 0.1875; at the bright knee (e = f = b = h
    # = 0.75, d = 0.625): mn4 = mx4 = 0.75, hitMin = 1/4, hitMax = 0.25 / -1,
    # lobe = -3/16, out = (-3/16 * 2.875 + 0.75) / (1/4) = 0.84375; plateaus
    # unchanged.
    W = H = 16
    x = np.clip((np.arange(W, dtype=np.float32) - 6.0) / 4.0, 0.0, 1.0) * 0.5 + 0.25
    src = np.zeros((H, W, 4), np.float32)
    src[..., :3] = x[None, :, None]
    src[..., 3] = 1.0
    row = oracle.fsr_rcas(src)[8, :, 0].astype(np.float64)
    assert np.isclose(row[6], 0.1875, rtol=1e-6) and np.isclose(row[10], 0.84375, rtol=1e-6)
    assert np.allclose(row[7:10], x[7:10], rtol=1e-6)
    assert np.all(row[:6] == 0.25) and np.all(row[11:] == 0.75)


def test_post_frame_at_the_render_size_is_unchanged_by_the_sized_form():
    # out_size equal to the render size and no upscale: the same frame as before
    H, W = 12, 20
    rng = np.random.default_rng(4)
    est = rng.random((H, W, 4)).astype(np.float32)
    gb = np.zeros((H, W, 4), np.uint32)
    gb[..., 2] = rng.integers(0, 1 << 24, (H, W), dtype=np.uint32)
    from common import settings
    st = settings()
    st.bloom_generate_bloom = 1
    outs = []
    for size in (None, (W, H)):
        e = abi.KptExposure.camera(aperture=8.0, iso=100, shutter=1 / 250)
        film = oracle.set_exposure(e)
        outs.append(oracle.post_frame(est, gb, st, e, film, 16.0, out_size=size))
    assert all(np.array_equal(a.view(np.uint32), b.view(np.uint32)) for a, b in zip(outs[0][:3], outs[1][:3]))
