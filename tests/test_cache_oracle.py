"""Pins the oracle's radiance-cache network (oracle/kpt_oracle_cache.cpp):
the encoding against its closed form, the backward pass against central
finite differences of the loss (exact FP64 mode, denominators held at the
current outputs as the stop-gradient loss defines), the BF16 mode against
FP64, the loss at zero weights, and the Adam step against its formulas.
"""
import math

import numpy as np
import pytest

import oracle
from kittlespt_amd import abi

BMIN, BMAX = (-2.0, -1.0, -3.0), (2.0, 3.0, 1.0)


def he_weights(seed=0):
    rng = np.random.default_rng(seed)
    parts = []
    for out, inp in abi.CACHE_LAYERS:
        a = math.sqrt(6.0 / inp)
        parts.append(rng.uniform(-a, a, out * inp))
    w = np.concatenate(parts).astype(np.float32)
    assert len(w) == abi.CACHE_PARAMS
    return w


def queries(n, seed=1):
    rng = np.random.default_rng(seed)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(BMIN, BMAX, (n, 3))
    pos[:, 3] = rng.uniform(0, 1, n)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    dirs = np.zeros((n, 4), np.float32)
    dirs[:, :3] = d
    return pos, dirs


def test_encoding_closed_form():
    pos, dirs = queries(200)
    f = oracle.cache_encode(pos, dirs, BMIN, BMAX).astype(np.float64)
    ph = (pos[:, :3].astype(np.float64) - BMIN) / (np.array(BMAX) - BMIN)
    for k in range(4):
        for a in range(3):
            arg = math.pi * 2 ** k * ph[:, a]
            assert np.allclose(f[:, 6 * k + 2 * a], np.sin(arg), atol=3e-6)
            assert np.allclose(f[:, 6 * k + 2 * a + 1], np.cos(arg), atol=3e-6)
    for a in range(3):
        assert np.allclose(f[:, 24 + 2 * a], np.sin(math.pi / 2 * dirs[:, a]), atol=2e-6)
        assert np.allclose(f[:, 25 + 2 * a], np.cos(math.pi / 2 * dirs[:, a]), atol=2e-6)
    assert np.array_equal(f[:, 30], pos[:, 3]) and (f[:, 31] == 1).all()


def test_backward_matches_finite_differences():
    w = he_weights(2)
    pos, dirs = queries(24, 3)
    t = np.zeros((24, 4), np.float32)
    t[:, :3] = np.random.default_rng(4).uniform(0, 2, (24, 3))
    loss, g = oracle.cache_loss_grad(w, pos, dirs, t, BMIN, BMAX, mode=1)
    assert loss > 0
    rng = np.random.default_rng(5)
    idx = np.concatenate([rng.integers(off, off + out * inp, 6) for off, (out, inp) in
                          zip(np.cumsum([0] + [o * i for o, i in abi.CACHE_LAYERS])[:-1], abi.CACHE_LAYERS)])
    def fd(i, h):
        wp, wm = w.copy(), w.copy()
        wp[i] += np.float32(h)
        wm[i] -= np.float32(h)
        step = float(wp[i]) - float(wm[i])
        lp, _ = oracle.cache_loss_grad(wp, pos, dirs, t, BMIN, BMAX, mode=1, w_den=w, want_grad=False)
        lm, _ = oracle.cache_loss_grad(wm, pos, dirs, t, BMIN, BMAX, mode=1, w_den=w, want_grad=False)
        return (lp - lm) / step

    # the frozen-denominator loss is piecewise quadratic in one weight, so a
    # central difference is exact unless a ReLU switches inside the step;
    # such kinks show as a mismatch between step h and h/2 and are skipped
    checked = 0
    for i in idx:
        if abs(g[i]) < 1e-7:
            continue
        a, b = fd(i, 6e-5), fd(i, 3e-5)
        if abs(a - b) > 1e-6 * max(abs(a), abs(b)):
            continue
        assert b == pytest.approx(g[i], rel=1e-6, abs=1e-10), (i, b, g[i])
        checked += 1
    assert checked >= 20


def test_bf16_mode_tracks_fp64():
    w = he_weights(6)
    pos, dirs = queries(256, 7)
    o0 = oracle.cache_forward(w, pos, dirs, BMIN, BMAX, mode=0)
    o1 = oracle.cache_forward(w, pos, dirs, BMIN, BMAX, mode=1)
    assert np.abs(o0 - o1).max() <= 0.05 * np.abs(o1).max()
    t = np.zeros((256, 4), np.float32)
    t[:, :3] = 1.0
    _, g0 = oracle.cache_loss_grad(w, pos, dirs, t, BMIN, BMAX, mode=0)
    _, g1 = oracle.cache_loss_grad(w, pos, dirs, t, BMIN, BMAX, mode=1)
    cos = g0 @ g1 / (np.linalg.norm(g0) * np.linalg.norm(g1))
    assert cos > 0.99


def test_zero_weights_loss():
    w = np.zeros(abi.CACHE_PARAMS, np.float32)
    pos, dirs = queries(10)
    t = np.zeros((10, 4), np.float32)
    t[:, :3] = np.linspace(0.1, 1.0, 30).reshape(10, 3)
    loss, g = oracle.cache_loss_grad(w, pos, dirs, t, BMIN, BMAX, mode=1)
    assert loss == pytest.approx((t[:, :3].astype(np.float64) ** 2 / 0.01).mean(), rel=1e-12)
    assert not g.any()       # dead ReLUs: no gradient reaches any weight


def test_adam_step_formulas():
    rng = np.random.default_rng(8)
    w = rng.normal(size=abi.CACHE_PARAMS).astype(np.float32)
    m = rng.normal(size=abi.CACHE_PARAMS).astype(np.float32) * 0.1
    v = rng.uniform(0, 1, abi.CACHE_PARAMS).astype(np.float32) * 0.1
    g = rng.normal(size=abi.CACHE_PARAMS).astype(np.float32)
    w2, m2, v2 = oracle.cache_adam(w, m, v, g, 3, 1e-3)
    md = 0.9 * m.astype(np.float64) + 0.1 * g
    vd = 0.99 * v.astype(np.float64) + 0.01 * g.astype(np.float64) ** 2
    wd = w - 1e-3 * (md / (1 - 0.9 ** 3)) / (np.sqrt(vd / (1 - 0.99 ** 3)) + 1e-15)
    assert np.allclose(m2, md, rtol=1e-5, atol=1e-7)
    assert np.allclose(v2, vd, rtol=1e-5, atol=1e-9)
    assert np.allclose(w2, wd, rtol=1e-5, atol=1e-7)
