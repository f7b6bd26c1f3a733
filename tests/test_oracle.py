"""Pins the CPU oracle to what the mathematics fixes (no GPU needed).

The reference is CUDA-only (nvcc/NVIDIA headers absent here) and ships no
tests or golden vectors, so each oracle routine is checked against closed
forms, independent formulations, invariants and brute force, chosen so a
dropped term, wrong sign/index or transposed operand fails.
"""
import math

import numpy as np
import pytest

import oracle
from kittlespt_amd import abi, scenes

from common import host_camera, random_rays


# ----------------------------------------------------------------- dmath
@pytest.mark.parametrize("fn,lo,hi,ref,max_ulp", [
    ("exp", -87.0, 88.0, np.exp, 2.0),
    ("exp2", -125.0, 127.0, np.exp2, 2.0),
    ("sin", -20.0, 20.0, np.sin, 1.0),
    ("cos", -20.0, 20.0, np.cos, 1.0),
    ("tan", -1.5, 1.5, np.tan, 1.0),
    ("acos", -1.0, 1.0, np.arccos, 1.0),
])
def test_dmath_ulp(fn, lo, hi, ref, max_ulp):
    x = np.random.default_rng(3).uniform(lo, hi, 300_000).astype(np.float32)
    y = oracle.dmath(fn, x).astype(np.float64)
    r = ref(x.astype(np.float64))
    ulp = np.abs(y - r) / np.spacing(np.abs(r).astype(np.float32)).astype(np.float64)
    assert ulp.max() <= max_ulp


@pytest.mark.parametrize("fn,ref", [("log2", np.log2), ("log", np.log)])
def test_dmath_log(fn, ref):
    x = np.exp(np.random.default_rng(4).uniform(-80, 80, 300_000)).astype(np.float32)
    x = np.concatenate([x, np.array([1e-45, 1e-40, 1.17549435e-38, 1.0, 2.0, 3.4e38], np.float32)])
    y = oracle.dmath(fn, x).astype(np.float64)
    r = ref(x.astype(np.float64))
    ulp = np.abs(y - r) / np.spacing(np.abs(r).astype(np.float32)).astype(np.float64)
    ulp[r == 0] = np.abs(y[r == 0])
    assert ulp.max() <= 1.0


def test_dmath_pow():
    rng = np.random.default_rng(5)
    x = rng.uniform(0, 8, 200_000).astype(np.float32)
    y = rng.uniform(-6, 6, 200_000).astype(np.float32)
    p = oracle.dmath("pow", x, y).astype(np.float64)
    r = np.power(x.astype(np.float64), y.astype(np.float64))
    ok = np.isfinite(r) & (r < 3e38) & (r > 1e-37)
    ulp = np.abs(p[ok] - r[ok]) / np.spacing(np.abs(r[ok]).astype(np.float32))
    assert ulp.max() <= 1.0
    # C99 special cases
    sx = np.array([2.0, 0.0, 0.0, -2.0, -2.0, 1.0, np.nan, -0.0, np.inf, 0.5], np.float32)
    sy = np.array([0.0, 2.0, -1.0, 3.0, 0.5, np.nan, 0.0, 3.0, -1.0, np.inf], np.float32)
    got = oracle.dmath("pow", sx, sy)
    want = np.power(sx.astype(np.float64), sy.astype(np.float64))
    assert np.array_equal(np.isnan(got), np.isnan(want))
    f = ~np.isnan(want)
    assert np.array_equal(got[f], want[f].astype(np.float32))


def test_dmath_exp_special():
    x = np.array([0.0, -0.0, 1.0, -200.0, 100.0, np.inf, -np.inf, np.nan], np.float32)
    e = oracle.dmath("exp", x)
    assert e[0] == 1.0 and e[1] == 1.0 and e[2] == np.float32(math.e)
    assert e[3] == 0.0 and np.isinf(e[4]) and np.isinf(e[5]) and e[6] == 0.0 and np.isnan(e[7])


# ----------------------------------------------------------------- RNG
def _pcg_inverse(h):
    """Undoes pcg_hash step by step: xorshift, odd multiply, random xorshift,
    LCG (each step of the PCG-RXS-M-XS output function is a bijection)."""
    M32 = 0xFFFFFFFF
    word = h ^ (h >> 22)                               # invert (w >> 22) ^ w
    word &= M32
    word = (word * pow(277803737, -1, 1 << 32)) & M32  # invert the odd multiply
    # invert ((state >> ((state >> 28) + 4)) ^ state): top 4 bits are unchanged
    shift = (word >> 28) + 4
    state = word
    for _ in range(8):
        state = word ^ (state >> shift)
    return ((state - 2891336453) * pow(747796405, -1, 1 << 32)) & M32


def test_pcg_hash_bijective_and_invertible():
    xs = np.random.default_rng(6).integers(0, 2**32, 20000, dtype=np.uint64)
    hs = [oracle.pcg_hash(int(x)) for x in xs]
    assert len(set(hs)) == len(set(int(x) for x in xs))
    for x, h in zip(xs[:2000], hs[:2000]):
        assert _pcg_inverse(h) == int(x)


def test_pcg_hash_uniform():
    hs = np.array([oracle.pcg_hash(i) for i in range(100000)], np.float64)
    u = hs / 2**32
    counts, _ = np.histogram(u, bins=64, range=(0, 1))
    chi2 = ((counts - 100000 / 64) ** 2 / (100000 / 64)).sum()
    assert chi2 < 120      # 63 dof: p ~ 1e-5 at 120


# ----------------------------------------------------------------- matrices
def test_matrix_inverses():
    rng = np.random.default_rng(7)
    for _ in range(50):
        m = np.eye(4)
        m[:3, :3] = scenes.random_rotation(rng) * rng.uniform(0.5, 4)
        m[:3, 3] = rng.uniform(-50, 50, 3)
        m16 = scenes.mat_colmajor(m)
        ref = np.linalg.inv(m)
        for inv in (oracle.glm_inverse(m16), oracle.mat4_inverse(m16)):
            got = inv.reshape(4, 4).T
            assert np.allclose(got, ref, rtol=1e-5, atol=1e-5)
    # Matrix4x4::inverse returns identity for |det| < 1e-6 (matrix_maths.cuh:323-326)
    sing = scenes.mat_colmajor(np.diag([1e-3, 1e-3, 1e-3, 1.0]))
    assert np.array_equal(oracle.mat4_inverse(sing), np.eye(4, dtype=np.float32).reshape(16))


# ----------------------------------------------------------------- BSDF
def _hemi_grid(n):
    """Midpoint quadrature over the upper hemisphere: directions and solid angles."""
    th = (np.arange(n) + 0.5) * (np.pi / 2) / n
    ph = (np.arange(4 * n) + 0.5) * (2 * np.pi) / (4 * n)
    T, P = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    dw = (np.sin(T) * (np.pi / 2 / n) * (2 * np.pi / (4 * n))).reshape(-1)
    return d, dw


@pytest.mark.parametrize("alpha", [0.05, 0.2, 0.5, 1.0])
def test_ggx_normalisation(alpha):
    """Integral of D(m) cos(theta_m) over the hemisphere is 1."""
    d, dw = _hemi_grid(1200 if alpha < 0.1 else 400)
    D = oracle.ggx_d(alpha, d, double=True).astype(np.float64)
    assert abs((D * d[:, 2] * dw).sum() - 1.0) < 2e-3


def test_fresnel_against_textbook():
    rng = np.random.default_rng(8)
    c = rng.uniform(-1, 1, 5000)
    eta = rng.uniform(1.1, 2.0, 5000)
    got = oracle.fresnel_dielectric(c, eta, double=True).astype(np.float64)
    # Fresnel equations for unpolarised light, incident side chosen by sign(cos)
    n1 = np.where(c > 0, 1.0, eta)
    n2 = np.where(c > 0, eta, 1.0)
    ci = np.abs(c)
    st = n1 / n2 * np.sqrt(np.maximum(0, 1 - ci * ci))
    tir = st >= 1
    ct = np.sqrt(np.maximum(0, 1 - st * st))
    rs = ((n1 * ci - n2 * ct) / (n1 * ci + n2 * ct)) ** 2
    rp = ((n2 * ci - n1 * ct) / (n2 * ci + n1 * ct)) ** 2
    want = np.where(tir, 1.0, 0.5 * (rs + rp))
    assert np.allclose(got, want, atol=1e-6)


def _sphere_grid(n):
    d, dw = _hemi_grid(n)
    lower = d * np.array([1, 1, -1])
    return np.concatenate([d, lower]), np.concatenate([dw, dw])


@pytest.mark.parametrize("rough", [0.3, 0.7, 1.0])
def test_pdf_normalisation(rough):
    """The glossy (half-vector) pdf integrates to 1 over the sphere; the opaque
    mixture 0.5*(glossy + cos/pi) integrates to 0.5 over the sphere (its
    diffuse term is odd in wi.z) and to 0.5*(1 + upper glossy mass) over the
    upper hemisphere."""
    d, dw = _sphere_grid(400)
    w0 = np.array([0.3, 0.0, math.sqrt(1 - 0.09)])
    wo = np.tile(w0, (len(d), 1))
    _, pc = oracle.bsdf_eval([1, 1, 1, 1.0, rough, 0.0, 1.45, 0.0], wo, d, double=True)
    pc = pc.astype(np.float64)
    # change of variables wi -> m = normalize(wo + wi) (Jacobian 1/(4 wo.m)):
    # the wi-space integral equals the D(m) cos(theta_m) mass of the
    # microfacets visible from wo
    hm, hw = _hemi_grid(400)
    Dm = oracle.ggx_d(rough * rough, hm, double=True).astype(np.float64)
    mass = (Dm * hm[:, 2] * hw * (hm @ w0 > 0)).sum()
    assert abs((pc * dw).sum() - mass) < 5e-3
    _, po = oracle.bsdf_eval([1, 1, 1, 0.0, rough, 0.0, 1.45, 0.0], wo, d, double=True)
    po = po.astype(np.float64)
    assert abs((po * dw).sum() - 0.5 * mass) < 5e-3
    up = d[:, 2] > 0
    glossy_up = (pc[up] * dw[up]).sum()
    assert abs((po[up] * dw[up]).sum() - 0.5 * (1.0 + glossy_up)) < 5e-3


def test_opaque_albedo_near_one():
    """White opaque dielectric: directional albedo of the reference's
    (1 - F*M)*diffuse + F*M mixture stays within a percent of one (the model
    is not exactly energy conserving)."""
    d, dw = _hemi_grid(300)
    for c in (1.0, 0.7, 0.3):
        wo = np.tile([math.sqrt(1 - c * c), 0.0, c], (len(d), 1))
        f, _ = oracle.bsdf_eval([1.0, 1.0, 1.0, 0.0, 1.0, 0.0, 1.45, 0.0], wo, d, double=True)
        alb = (f[:, 0].astype(np.float64) * d[:, 2] * dw).sum()
        assert 0.95 < alb < 1.02


def test_sampling_matches_pdf_diffuse():
    """Cosine-weighted diffuse samples follow pdf = cos/pi (chi-square over
    cos-theta bins)."""
    n = 200_000
    rng = np.random.default_rng(9)
    u = rng.random((n, 4)).astype(np.float32)
    u[:, 3] = 0.75                         # X.y >= 0.5 selects the diffuse lobe
    u[:, 2] = 0.99                         # not metallic, not transmissive
    wo = np.tile([0.0, 0.0, 1.0], (n, 1))
    wi, f, pdf, fl = oracle.bsdf_sample([0.8, 0.8, 0.8, 0.0, 1.0, 0.0, 1.45, 0.0], wo, u, double=True)
    assert np.all(fl == (8 | 2))
    ct = wi[:, 2].astype(np.float64)
    counts, edges = np.histogram(ct, bins=20, range=(0, 1))
    expect = n * (edges[1:] ** 2 - edges[:-1] ** 2)    # P(cos < c) = c^2
    chi2 = ((counts - expect) ** 2 / expect).sum()
    assert chi2 < 50
    assert np.allclose(pdf, ct / np.pi, rtol=1e-5, atol=1e-7)


def test_ggx_sampling_distribution():
    """Half vectors of conductor samples follow D(m) cos(theta_m)."""
    n = 200_000
    rng = np.random.default_rng(10)
    u = rng.random((n, 4)).astype(np.float32)
    u[:, 2] = 0.0                          # X.x < metallic -> conductor
    wo = np.tile([0.0, 0.0, 1.0], (n, 1))
    rough = 0.5
    wi, f, pdf, fl = oracle.bsdf_sample([0.9, 0.9, 0.9, 1.0, rough, 0.0, 1.5, 0.0], wo, u, double=True)
    h = wi.astype(np.float64) + wo
    h /= np.linalg.norm(h, axis=1, keepdims=True)
    alpha = rough ** 2
    a2 = alpha * alpha
    # CDF of cos(theta_m) under D cos: F(c) = (1 - c^2) / (1 + (a2 - 1) c^2) complement
    c = h[:, 2]
    edges = np.linspace(0, 1, 21)
    counts, _ = np.histogram(c, bins=edges)
    F = lambda x: 1 - (1 - x ** 2) / (1 + (a2 - 1) * x ** 2)  # noqa: E731
    expect = n * (F(edges[1:]) - F(edges[:-1]))
    m = expect > 50
    chi2 = ((counts[m] - expect[m]) ** 2 / expect[m]).sum()
    assert chi2 < 60


def test_lobe_selection_non_cumulative():
    """sampleF picks transmission when X.x < (1-m)*t, not m + (1-m)*t
    (bsdf.cuh:256): with m = 0.5, t = 1, X.x in [0.5, 0.75) is *not*
    transmissive in the reference."""
    n = 1000
    u = np.column_stack([np.full(n, 0.3), np.full(n, 0.6), np.linspace(0.5, 0.749, n), np.full(n, 0.9)])
    wo = np.tile([0.0, 0.0, 1.0], (n, 1))
    wi, f, pdf, fl = oracle.bsdf_sample([0.9, 0.9, 0.9, 0.5, 0.4, 1.0, 1.5, 0.0], wo, u.astype(np.float32))
    assert not np.any(fl & 4)


# ----------------------------------------------------------------- sky
def _sun(theta_deg, phi=4.18879):
    th = math.radians(theta_deg)
    v = np.array([math.cos(phi) * math.cos(th), math.sin(th), math.sin(phi) * math.cos(th)])
    return v / np.linalg.norm(v)


def test_sky_invariants():
    obs = np.array([0.0, 6378138.0, 0.0])
    sun = _sun(30.0)
    rng = np.random.default_rng(11)
    # rotational symmetry about the vertical through the observer does not hold
    # (the sun breaks it) but symmetry about the sun-zenith plane does: mirror
    # a direction across that plane and the radiance is unchanged
    nrm = np.cross(sun, [0.0, 1.0, 0.0])
    nrm /= np.linalg.norm(nrm)
    d = rng.normal(size=(500, 3))
    d[:, 1] = np.abs(d[:, 1]) + 0.05
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    dm = d - 2 * (d @ nrm)[:, None] * nrm[None, :]
    a = oracle.sky_li(np.tile(obs, (500, 1)), d, sun, double=True)
    b = oracle.sky_li(np.tile(obs, (500, 1)), dm, sun, double=True)
    assert np.allclose(a, b, rtol=1e-5, atol=0)       # float32 inputs/outputs
    # linear in the emission factor
    c = oracle.sky_li(np.tile(obs, (500, 1)), d, sun, emission=2.5, double=True)
    assert np.allclose(c, 2.5 * a, rtol=1e-6)
    # blue sky at zenith (Rayleigh dominates), redder toward the horizon at low sun
    z = oracle.sky_li([obs], [[0, 1, 0]], sun, double=True)[0]
    assert z[2] > z[1] > z[0] > 0
    # a ray from outside pointing away from the atmosphere sees nothing
    out = oracle.sky_li([[0, 7.0e6, 0]], [[0, 1, 0]], sun)
    assert np.all(out == 0)
    # sun below the horizon: the light march hits the ground for low points
    dark = oracle.sky_li([obs], [[1, 0.01, 0]], _sun(-20.0), double=True)[0]
    assert np.all(dark < 1e-3 * z)


def test_sky_against_converged_single_scattering():
    """The 16x8 midpoint march approximates the physical single-scattering
    integral (same constants, 4096x512 steps in numpy) within 15%."""
    obs = np.array([0.0, 6378138.0, 0.0])
    sun = _sun(30.0)
    dirs = np.array([[0, 1, 0], [0.5, 0.5, 0.0], [0.0, 0.2, 1.0]], np.float64)
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    got = oracle.sky_li(np.tile(obs, (3, 1)), dirs, sun, double=True).astype(np.float64)
    R, Ra, Hr, Hm = 6378137.0, 6438137.0, 7994.0, 1200.0
    bR = np.array([3.8e-6, 13.5e-6, 33.1e-6])
    bM = 21e-6
    bMe = bM + 2.1e-6

    def exit_t(o, d):
        b = 2 * o @ d
        c = o @ o - Ra * Ra
        return (-b + math.sqrt(b * b - 4 * c)) / 2

    for k, d in enumerate(dirs):
        T = exit_t(obs, d)
        n = 4096
        ts = (np.arange(n) + 0.5) * T / n
        pts = obs[None, :] + ts[:, None] * d[None, :]
        h = np.linalg.norm(pts, axis=1) - R
        odR = np.cumsum(np.exp(-h / Hr)) * T / n
        odM = np.cumsum(np.exp(-h / Hm)) * T / n
        LR = np.zeros(3)
        LM = np.zeros(3)
        for i in range(0, n, 8):
            Ts = exit_t(pts[i], sun)
            m = 512
            ls = (np.arange(m) + 0.5) * Ts / m
            lp = pts[i][None, :] + ls[:, None] * sun[None, :]
            lh = np.linalg.norm(lp, axis=1) - R
            lR = np.exp(-lh / Hr).sum() * Ts / m
            lM = np.exp(-lh / Hm).sum() * Ts / m
            tau = bR * (odR[i] + lR) + bMe * (odM[i] + lM)
            tr = np.exp(-tau)
            LR += tr * math.exp(-h[i] / Hr) * T / n * 8
            LM += tr * math.exp(-h[i] / Hm) * T / n * 8
        mu = d @ sun
        g = 0.76
        pr = 3 * (1 + mu * mu) / (16 * math.pi)
        pm = 3 * ((1 - g * g) * (1 + mu * mu)) / ((8 * math.pi) * (2 + g * g) * (1 + g * g - 2 * g * mu) ** 1.5)
        want = (bR * LR * pr + bM * LM * pm) * 6.0e5
        assert np.allclose(got[k], want, rtol=0.15), (k, got[k], want)


def test_phase_functions_normalised():
    mu = np.linspace(-1, 1, 200001)
    pr = 3 * (1 + mu ** 2) / (16 * np.pi)
    g = 0.76
    pm = 3 * ((1 - g * g) * (1 + mu ** 2)) / ((8 * np.pi) * (2 + g * g) * (1 + g * g - 2 * g * mu) ** 1.5)
    assert abs(np.trapezoid(pr, mu) * 2 * np.pi - 1) < 1e-6
    assert abs(np.trapezoid(pm, mu) * 2 * np.pi - 1) < 1e-3


# ----------------------------------------------------------------- camera
def test_camera_closed_form():
    sc = scenes.scene_plain(subdiv=1)
    W, H = 64, 48
    cam = host_camera(sc, W, H)
    px = np.array([(H // 2) * W + W // 2, 0, W - 1, (H - 1) * W], np.int32)
    o, d = oracle.camera_rays(cam, W, H, px)
    fwd = np.array(sc.camera_target) - np.array(sc.camera_eye)
    fwd /= np.linalg.norm(fwd)
    assert np.allclose(o[0], sc.camera_eye, atol=1e-5)
    # the pixel right of centre is half a pixel off axis
    assert math.degrees(math.acos(np.clip(d[0] @ fwd, -1, 1))) < 1.0
    # corner rays sit at the frustum corners: tan(vfov/2) vertically
    t = math.tan(sc.camera_fov / 2)
    for k in range(1, 4):
        local_z = d[k] @ fwd
        off = d[k] - local_z * fwd
        up = np.cross(np.cross(fwd, [0, 1, 0]), fwd)
        up /= np.linalg.norm(up)
        assert abs(abs(off @ up) / local_z - t * (1 - 1 / H)) < 2e-3


# ----------------------------------------------------------------- traversal
def test_bvh_matches_brute_force_plain():
    sc = scenes.scene_plain(subdiv=4)
    orc = oracle.OracleScene(sc)
    rays = random_rays(20000, (-3, 0.05, -3), (3, 3, 3), (0, 1, 0), 0.8, seed=12)
    h, _, _ = orc.trace(rays, oracle.MODE_CLOSEST)
    b, _, _ = orc.trace(rays, oracle.MODE_BRUTE_CLOSEST)
    assert np.array_equal(h["prim_id"], b["prim_id"])
    assert np.array_equal(h["t"], b["t"])
    a, _, _ = orc.trace(rays, oracle.MODE_ANY)
    ab, _, _ = orc.trace(rays, oracle.MODE_BRUTE_ANY)
    assert np.array_equal(a, ab)


def test_bvh_matches_brute_force_instanced():
    sc = scenes.scene_instanced_small()
    orc = oracle.OracleScene(sc)
    rays = random_rays(4000, (-10, -10, -10), (110, 110, 110), (50, 50, 50), 30.0, seed=13)
    # aim half of the rays at instance centres so many of them hit
    centres = sc.instance_model[:, 12:15].astype(np.float64)
    rng = np.random.default_rng(15)
    tgt = centres[rng.integers(0, len(centres), 2000)] + rng.normal(size=(2000, 3))
    d = tgt - rays["org"][:2000]
    rays["dir"][:2000] = d / np.linalg.norm(d, axis=1, keepdims=True)
    h, _, _ = orc.trace(rays, oracle.MODE_CLOSEST)
    b, _, _ = orc.trace(rays, oracle.MODE_BRUTE_CLOSEST)
    assert np.array_equal(h["prim_id"], b["prim_id"]) and np.array_equal(h["inst_id"], b["inst_id"])
    assert (h["prim_id"] >= 0).mean() > 0.2


def test_triangle_hits_against_float64():
    """Moller-Trumbore in float agrees with an exact float64 ray/plane
    intersection away from triangle edges."""
    sc = scenes.scene_plain(subdiv=3)
    orc = oracle.OracleScene(sc)
    rays = random_rays(5000, (-3, 0.05, -3), (3, 3, 3), (0, 1, 0), 0.5, seed=14)
    h, _, _ = orc.trace(rays, oracle.MODE_CLOSEST)
    hit = h["prim_id"] >= 0
    P = sc.positions.astype(np.float64)
    models = sc.mesh_model.reshape(-1, 4, 4).transpose(0, 2, 1).astype(np.float64)
    for i in np.nonzero(hit)[0][:2000]:
        pid, inst = h["prim_id"][i], h["inst_id"][i]
        Mi = np.linalg.inv(models[inst])
        o = Mi[:3, :3] @ rays["org"][i] + Mi[:3, 3]
        d = Mi[:3, :3] @ rays["dir"][i]
        v0, v1, v2 = P[pid]
        n = np.cross(v1 - v0, v2 - v0)
        t = ((v0 - o) @ n) / (d @ n)
        assert abs(t - h["t"][i]) <= 1e-4 * max(1.0, t)
        b = h["bary"][i]
        assert abs(b.sum() - 1) < 1e-5 and (b >= -1e-6).all()


def test_traversal_edge_rays():
    """Axis-parallel directions (1/d = +-inf, including -0.0 components) agree
    with brute force on distance when the ray does not lie in a box face
    plane."""
    sc = scenes.scene_plain(subdiv=2)
    orc = oracle.OracleScene(sc)
    dirs = np.array([[0, -1, 0], [0, 0, 1], [1, 0, 0], [-0.0, -1, -0.0], [0, -1, -0.0], [-1, 0, 0]], np.float32)
    orgs = np.array([[0.0123, 3, 0.0311], [0.0217, 1.0131, -3], [-3, 1.0171, 0.0213], [0.5, 3, 0.5],
                     [0.0131, 3, 0.777], [3, 1.0313, 0.1]], np.float32)
    rays = np.zeros(len(dirs), abi.RAY_DTYPE)
    rays["org"] = orgs
    rays["dir"] = dirs
    rays["tmax"] = np.inf
    h, _, _ = orc.trace(rays, oracle.MODE_CLOSEST)
    b, _, _ = orc.trace(rays, oracle.MODE_BRUTE_CLOSEST)
    assert np.array_equal(h["t"], b["t"])
    assert (h["prim_id"] >= 0).sum() >= 5


def test_slab_nan_quirk_preserved():
    """A ray lying exactly in a box face plane makes (pmin - o) * (1/d) =
    0 * inf = NaN; fminf/fmaxf drop the NaN and the box is missed
    (bounds.cuh:40-47).  The oracle keeps this reference behaviour, so the BVH
    can miss a vertex that brute force hits."""
    sc = scenes.scene_plain(subdiv=2)
    orc = oracle.OracleScene(sc)
    rays = np.zeros(1, abi.RAY_DTYPE)
    rays["org"] = [[-3.0, 1.0, 0.0]]        # y = 1, z = 0: the sphere's equatorial vertex plane
    rays["dir"] = [[1.0, 0.0, 0.0]]
    rays["tmax"] = np.inf
    h, _, _ = orc.trace(rays, oracle.MODE_CLOSEST)
    b, _, _ = orc.trace(rays, oracle.MODE_BRUTE_CLOSEST)
    assert b["prim_id"][0] >= 0 and h["prim_id"][0] == -1
