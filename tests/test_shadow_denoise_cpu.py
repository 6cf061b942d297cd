"""CPU checks pinning the oracle's sun-shadow denoise chain
(ShadowRayTraceFrag.glsl:505-511, ShadowTemporalFilter.glsl:98-267,
ShadowFilter.glsl:69-165) and the shared vpt_logf / vpt_powf, each against
a closed form or an independent computation."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from voxelpathtracer_amd.pipeline import viewproj

W, H = 96, 64


def f16(x):
    return np.float32(x).astype(np.float16).view(np.uint16)


def h2f(bits):
    return np.asarray(bits, np.uint16).view(np.float16).astype(np.float32)


def test_logf_powf_error_bounds():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(1e-30, 1e30, 20000), rng.uniform(0.5, 2.0, 20000),
                        np.float32(2.0) ** rng.integers(-140, 127, 2000)]).astype(np.float32)
    got = oracle.logf(x).astype(np.float64)
    ref = np.log(x.astype(np.float64))
    rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-3)
    assert rel.max() < 3e-7, rel.max()
    assert oracle.logf(np.float32([1.0]))[0] == 0.0
    xs = rng.uniform(2.0 / 3.0, 1.0, 50000).astype(np.float32)
    es = rng.uniform(0.0, 60.0, 50000).astype(np.float32)
    p = oracle.powf(xs, es).astype(np.float64)
    pr = np.power(xs.astype(np.float64), es.astype(np.float64))
    assert (np.abs(p - pr) / pr).max() < 5e-6
    assert oracle.powf(np.float32([0.0, 0.7, 1.0]), np.float32([2.0, 0.0, 33.0])).tolist() == [0.0, 1.0, 1.0]


def _slab_world():
    """Solid floor y < 10 (ids 3) and a slab 30 <= y < 34 over x, z in [16, 40)."""
    g = np.zeros((64, 48, 64), np.uint8)
    g[:, :10, :] = 3
    g[16:40, 30:34, 16:40] = 5
    return g


def _gbuffer(g, df, pos, yaw, pitch):
    cam = oracle.camera(pos, yaw, pitch, aspect=W / H)
    p = oracle.trace_primary(g, df, cam["inv_view"], cam["inv_proj"], W, H, 350)
    return cam, p


def test_transversal_closed_forms():
    g = _slab_world()
    df = oracle.df_build(g)
    cam, p = _gbuffer(g, df, (28.5, 24.5, 4.5), 90.0, -35.0)   # looking along +z, down at the floor under the slab
    sun = np.float32([0.0, 1.0, 0.0])
    r = oracle.trace_shadow2(g, df, cam["inv_view"], cam["inv_proj"], p["t"], p["face"], sun)
    r0 = oracle.trace_shadow(g, df, cam["inv_view"], cam["inv_proj"], p["t"], p["face"], sun)
    np.testing.assert_array_equal(r["mask"], r0["mask"])
    tv = h2f(r["transversal"])
    sky = p["t"] < 0
    assert sky.any() and (tv[sky] == h2f(f16(64.0))).all()
    top = (p["face"] == 2) & ~sky                                # +y faces: N.L = 1
    side = (p["face"] != 2) & ~sky                               # N.L = 0 <= 0.01
    assert (tv[side] == h2f(f16(0.01))).all() and (r["mask"][side] == 1).all()
    vox = p["voxel"][top]
    x, y, z = vox & 1023, (vox >> 10) & 1023, (vox >> 20) & 1023
    floor = top.copy()
    floor[top] = y == 9
    under = floor.copy()
    under[floor] = (x[y == 9] >= 16) & (x[y == 9] < 40) & (z[y == 9] >= 16) & (z[y == 9] < 40)
    open_ = floor & ~under
    assert under.sum() > 50 and open_.sum() > 50
    # floor top at y = 10, origin 10.06, slab bottom at y = 30: T = 19.94 (+1e-4 step epsilon)
    assert np.abs(tv[under] - 0.1994).max() < 2e-4 and (r["mask"][under] == 1).all()
    assert (tv[open_] == h2f(f16(0.0425))).all() and (r["mask"][open_] == 0).all()


def _static_inputs(mask_value, tv_value):
    g = np.zeros((48, 48, 64), np.uint8)
    g[:, :10, :] = 3
    df = oracle.df_build(g)
    cam, p = _gbuffer(g, df, (32.5, 30.5, 8.5), 90.0, -40.0)
    vp = viewproj(cam["view"], cam["proj"])
    mask = np.full((H, W), mask_value, np.uint8)
    mask[p["t"] < 0] = 0
    tv = np.full((H, W), f16(tv_value), np.uint16)
    return cam, p, vp, mask, tv


@pytest.mark.parametrize("mask_value", [0, 1])
def test_temporal_static_camera_ladder_and_blend(mask_value):
    """Static camera, uniform mask c and transversal 20 (no history clip):
    the frame count climbs by one per frame from empty history and every
    output is mix(c, history, clamp((1 - 1/(f + 1)) 1.2, 0.01, 0.97))."""
    cam, p, vp, mask, tv = _static_inputs(mask_value, 0.2)
    hist_c = np.full((H, W), 128, np.uint8)
    hist_f = np.zeros((H, W), np.uint16)
    lit = p["t"] > 0
    inner = np.zeros_like(lit)
    inner[8:-8, 8:-8] = True
    inner &= lit
    for frame in range(5):
        col, frm = oracle.shadow_temporal(cam["inv_view"], cam["inv_proj"], vp, p["t"], p["face"], p["t"], mask, tv,
                                          hist_c, hist_f)
        assert (h2f(frm)[inner] == frame + 1).all(), frame
        fi = np.float32(frame) + np.float32(1.0)
        a = np.clip((np.float32(1.0) - np.float32(1.0) / fi) * np.float32(1.2), np.float32(0.01), np.float32(0.97))
        hprev = hist_c[inner].astype(np.float32) / np.float32(255.0)
        c = np.float32(mask_value)
        expect = c * (np.float32(1.0) - a) + hprev * a
        np.testing.assert_array_equal(col[inner], np.round(np.clip(expect, 0, 1) * 255.0).astype(np.uint8))
        sky = p["t"] < 0
        assert (col[sky] == 0).all() and (h2f(frm)[sky] == 0).all()
        hist_c, hist_f = col, frm


@pytest.mark.parametrize("value", [0, 255])
def test_filter_uniform_identity_and_passthrough(value):
    cam, p, vp, mask, tv = _static_inputs(1, 0.2)
    color = np.full((H, W), value, np.uint8)
    frames = np.full((H, W), f16(9.0), np.uint16)
    out = oracle.shadow_filter(cam["inv_view"], cam["inv_proj"], p["t"], p["face"], color, tv, frames)
    np.testing.assert_array_equal(out, color)
    # sharp shadows (transversal * 100 below sqrt 2) and sky pass through unchanged
    rng = np.random.default_rng(2)
    color = rng.integers(0, 256, (H, W), dtype=np.uint8)
    sharp = np.full((H, W), f16(0.01), np.uint16)
    out = oracle.shadow_filter(cam["inv_view"], cam["inv_proj"], p["t"], p["face"], color, sharp, frames)
    np.testing.assert_array_equal(out, color)
    out = oracle.shadow_filter(cam["inv_view"], cam["inv_proj"], p["t"], p["face"], color, tv, frames)
    sky = p["t"] < 0
    np.testing.assert_array_equal(out[sky], color[sky])
    assert (out[~sky] != color[~sky]).any()
