"""CPU tests: the oracle pinned against closed forms, textbook algorithms and
libm/numpy, independently of the transcription itself.

The reference ships no tests or golden vectors, so each oracle function is
pinned by what the mathematics fixes (see the docstring of each test), and
the committed fixtures under tests/golden/ pin it against drift.
"""
from __future__ import annotations

import numpy as np
import pytest
from scipy.ndimage import distance_transform_cdt

import oracle
from tests.helpers import camera, f16, random_world, terrain


def brute_df(g):
    """min(cap, L1 distance to the nearest solid voxel), cap = min(254,
    nx + ny + nz) (MaxDistance, ManhattanDistanceX.comp:51)."""
    nz, ny, nx = g.shape
    cap = min(254, nx + ny + nz)
    pts = np.argwhere(g > 0)
    out = np.full(g.shape, cap, np.int64)
    if len(pts):
        z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        for (pz, py, px) in pts:
            out = np.minimum(out, np.abs(z - pz) + np.abs(y - py) + np.abs(x - px))
    return np.minimum(out, cap).astype(np.uint8)


def closed_form_df(g):
    cap = min(254, sum(g.shape))
    if not (g > 0).any():
        return np.full(g.shape, cap, np.uint8)
    return np.minimum(cap, distance_transform_cdt(g == 0, metric="taxicab")).astype(np.uint8)


# ---------------------------------------------------------------- DF

@pytest.mark.parametrize("seed,density", [(1, 0.01), (2, 0.1), (3, 0.0), (4, 1.0)])
def test_df_brute_force_tiny(seed, density):
    """min(cap, L1 to nearest solid) by brute force on tiny grids."""
    g = random_world(np.random.default_rng(seed), 12, 7, 9, density)
    np.testing.assert_array_equal(oracle.df_build(g), brute_df(g))


def test_df_closed_form_terrain():
    g = terrain(128, 64, 128, h_min=10, h_max=40)
    np.testing.assert_array_equal(oracle.df_build(g), closed_form_df(g))


def test_df_cap_254():
    """A single voxel in a long world: distances saturate at 254
    (MaxDistance = min(254, X+Y+Z), ManhattanDistanceX.comp:51)."""
    g = np.zeros((8, 8, 512), np.uint8)
    g[3, 4, 0] = 1
    df = oracle.df_build(g)
    x = np.arange(512)
    assert df[4, 4, 300] == 254
    np.testing.assert_array_equal(df[3, 4, :], np.minimum(254, x).astype(np.uint8))
    assert df[4, 3, 10] == min(254, 10 + 1 + 1)


def test_df_step_table():
    """E = floor(M == 1 ? 1 : M * 0.57735026918), with M read back exactly
    through the unorm8 path, for every byte."""
    tab = oracle.df_step_table()
    k = np.arange(256, dtype=np.float32)
    # unorm8 round trip is exact
    assert np.array_equal((k / np.float32(255)).astype(np.float32) * np.float32(255), k)
    expect = np.floor(k * np.float32(0.57735026918)).astype(np.int32)
    expect[1] = 1
    np.testing.assert_array_equal(tab, expect)
    assert tab[0] == 0 and tab[3] == 1 and tab[4] == 2 and tab[254] == 146


# ---------------------------------------------------------------- numerics

def test_expf_accuracy():
    x = np.concatenate([np.linspace(-87, 88, 2_000_001, dtype=np.float32),
                        np.float32([-0.0, 0.0, 1e-30, -1e-30])])
    y = oracle.expf(x).astype(np.float64)
    ref = np.exp(x.astype(np.float64))
    rel = np.abs(y - ref) / ref
    assert rel.max() < 4e-7, rel.max()
    assert oracle.expf(np.float32([-100.0]))[0] == 0.0
    assert np.isinf(oracle.expf(np.float32([100.0]))[0])
    assert np.isnan(oracle.expf(np.float32([np.nan]))[0])


def test_sincos_2pi_accuracy():
    u = np.concatenate([np.linspace(0, 1, 1_000_000, endpoint=False, dtype=np.float32),
                        np.float32([0.0, 0.25, 0.5, 0.75])])
    s, c = oracle.sincos_2pi(u)
    a = 2 * np.pi * u.astype(np.float64)
    assert np.abs(s - np.sin(a)).max() < 5e-7
    assert np.abs(c - np.cos(a)).max() < 5e-7
    assert s[-4] == 0.0 and c[-4] == 1.0 and s[-3] == 1.0 and c[-1] == 0.0


def test_f16_conversion_matches_numpy_rne():
    h = np.arange(65536, dtype=np.uint16)
    x = oracle.f16_bits_to_f32(h)
    ref = h.view(np.float16).astype(np.float32)
    finite = np.isfinite(ref)
    np.testing.assert_array_equal(x[finite], ref[finite])
    assert np.isnan(x[~finite & np.isnan(ref)]).all()
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.normal(0, 10, 1_000_000), rng.normal(0, 1e-5, 200_000), rng.uniform(-7e4, 7e4, 200_000),
                         [65504, 65519.99, 65520, -65520, 2**-25, 2**-24, 1.5 * 2**-24, 0.0, -0.0]]).astype(np.float32)
    got = oracle.f32_to_f16_bits(xs)
    np.testing.assert_array_equal(got, xs.astype(np.float16).view(np.uint16))


def test_unorm8_and_powi():
    # decode: the fused-correction form equals the IEEE quotient c / 255
    c8 = np.arange(256, dtype=np.float32)
    np.testing.assert_array_equal(oracle.unorm8_decode_table(), c8 / np.float32(255))
    x = np.linspace(-0.5, 1.5, 100_001, dtype=np.float32)
    c = oracle.unorm8_encode(x)
    ref = np.floor(np.clip(x, 0, 1) * np.float32(255) + np.float32(0.5)).astype(np.uint8)
    np.testing.assert_array_equal(c, ref)
    v = np.float32([0.0, 1.0, 3.0, 0.5, 0.9999])
    np.testing.assert_allclose(oracle.powi(v, 16), v.astype(np.float64) ** 16, rtol=2e-6)
    assert oracle.powi(np.float32([3.0]), 16)[0] == 43046721.0


def test_pcg_hash_known_values():
    """pcg_hash of Jarzynski & Olano (2020) restated in numpy."""
    v = np.arange(0, 1 << 20, 7, dtype=np.uint64)
    state = (v * 747796405 + 2891336453) & 0xFFFFFFFF
    word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & 0xFFFFFFFF
    ref = ((word >> 22) ^ word).astype(np.uint32)
    np.testing.assert_array_equal(oracle.pcg_hash(v.astype(np.uint32)), ref)


# ---------------------------------------------------------------- traversal

def _rays(org, d):
    d = np.asarray(d, np.float64)
    d = d / np.linalg.norm(d, axis=-1, keepdims=True)
    return np.concatenate([np.asarray(org, np.float64), d], -1).astype(np.float32)


def test_traversal_wall_known_answer():
    """A ray slightly off the +x axis toward a wall at x = 20 hits cell
    (20, y, z) on its -x face at distance ~14.5."""
    g = np.zeros((32, 32, 48), np.uint8)
    g[:, :, 20] = 7
    df = oracle.df_build(g)
    r = oracle.trace_rays(g, df, _rays([[5.5, 10.5, 10.5]], [[1.0, 1e-3, 2e-3]]), 350)
    assert r["face"][0] == 4 and r["block"][0] == 7
    assert tuple(r["cell"][0]) == (20, 10, 10)
    assert abs(r["t"][0] - 14.5) < 2e-3


def test_traversal_straight_down_and_axis_quirk():
    """Straight down onto the floor: top face (+y, id 2).  An exactly +x ray
    has zero y/z components: the DDA step then divides by zero and the
    reference's MinIdx rule picks the y axis (InitialRayTraceFrag.glsl:
    345-347), pushing the ray out of the volume -- a miss.  The transcription
    keeps that behaviour."""
    g = np.zeros((32, 32, 48), np.uint8)
    g[:, 0:3, :] = 3
    g[:, :, 40] = 5
    df = oracle.df_build(g)
    r = oracle.trace_rays(g, df, _rays([[10.5, 20.5, 10.5], [5.5, 10.5, 10.5]], [[0, -1, 0], [1, 0, 0]]), 350)
    assert r["face"][0] == 2 and tuple(r["cell"][0]) == (10, 2, 10) and abs(r["t"][0] - 17.5) < 2e-3
    assert r["face"][1] == 255 and r["t"][1] == -1.0


def test_traversal_invariants_and_textbook_agreement():
    """Against Amanatides & Woo in double precision: the DF traversal takes
    conservative skips and single-cell DDA steps, so its first solid cell
    agrees with the textbook DDA for all but a few rays (the reference's skip
    can land inside a solid cell and stop there).  Hits must be solid cells,
    t >= 0, and the face must oppose the ray on its axis."""
    g = terrain(128, 64, 128, h_min=10, h_max=40)
    df = oracle.df_build(g)
    rng = np.random.default_rng(5)
    n = 20000
    org = np.stack([rng.uniform(2, 126, n), rng.uniform(45, 62, n), rng.uniform(2, 126, n)], 1)
    d = rng.normal(size=(n, 3))
    d[:, 1] = -np.abs(d[:, 1]) - 0.2
    rays = _rays(org, d)
    r = oracle.trace_rays(g, df, rays, 350)
    hit = r["face"] != 255
    assert hit.mean() > 0.5
    c = r["cell"][hit]
    assert (g[c[:, 2], c[:, 1], c[:, 0]] > 0).all()
    assert (r["t"][hit] >= 0).all()
    normals = np.array([[0, 0, 1], [0, 0, -1], [0, 1, 0], [0, -1, 0], [-1, 0, 0], [1, 0, 0]], np.float64)
    nd = (normals[r["face"][hit]] * rays[hit, 3:6]).sum(1)
    assert (nd < 0).mean() > 0.999
    agree = same_face = both = 0
    for i in np.flatnonzero(hit)[:5000]:
        ref = oracle.dda_textbook(g, rays[i, 0:3].astype(np.float64), rays[i, 3:6].astype(np.float64), 1e4)
        if ref[0] < 0:
            continue
        both += 1
        if tuple(ref[0:3]) == tuple(r["cell"][i]):
            agree += 1
            same_face += int(ref[3] == r["face"][i])
    assert both > 4000
    assert agree / both > 0.998, agree / both
    assert same_face / agree > 0.998


def test_traversal_vs_raycast_detect():
    """World::RaycastDetect (CPU epsilon stepping, 48 steps) against the DF
    traversal on rays that hit within reach."""
    g = terrain(128, 64, 128, h_min=10, h_max=40)
    df = oracle.df_build(g)
    rng = np.random.default_rng(9)
    n = 3000
    org = np.stack([rng.uniform(10, 118, n), rng.uniform(42, 48, n), rng.uniform(10, 118, n)], 1)
    d = rng.normal(size=(n, 3))
    d[:, 1] = -np.abs(d[:, 1]) - 0.5
    rays = _rays(org, d)
    r = oracle.trace_rays(g, df, rays, 350)
    both = agree = 0
    for i in range(n):
        pick = oracle.raycast_detect(g, rays[i, 0:3], rays[i, 3:6])
        if pick[3] < 0 or r["face"][i] == 255:
            continue
        both += 1
        agree += int(tuple(pick[0:3]) == tuple(r["cell"][i]) and pick[3] == r["block"][i])
    assert both > 1000
    assert agree / both > 0.99, agree / both


def test_primary_image_properties():
    g = terrain()
    df = oracle.df_build(g)
    W, H = 160, 90
    cam, view, proj, iv, ip = camera(W, H)
    p = oracle.trace_primary(g, df, iv, ip, W, H, 350)
    hit = p["face"] != 255
    assert 0.3 < hit.mean() < 0.9
    # the sky is in the top rows (row 0 is the bottom of the image)
    assert hit[:10].mean() > 0.95 and hit[-5:].mean() < 0.2
    assert (p["t"][~hit] == -1).all() and (p["block"][~hit] == 0).all()
    v = p["voxel"][hit]
    cx, cy, cz = v & 1023, (v >> 10) & 1023, (v >> 20) & 1023
    np.testing.assert_array_equal(g[cz, cy, cx], p["block"][hit])
    assert (p["samples"] > 0).all()


# ---------------------------------------------------------------- GI

def test_gi_sh_encoding_identities():
    """IrridianceToSH: SHToY(sh) = 3.544905 * 0.282095 * Y = Y; the band-1
    coefficients are 0.488603 * Y * dir with |dir| = 1; Co = R - B,
    Cg = G - (B + Co/2)."""
    g = terrain(128, 64, 128, h_min=10, h_max=40)
    df = oracle.df_build(g)
    W, H = 96, 54
    cam, view, proj, iv, ip = camera(W, H, pos=(64.5, 50.5, 64.5), yaw=30.0, pitch=-25.0)
    p = oracle.trace_primary(g, df, iv, ip, W, H, 350)
    from voxelpathtracer_amd import gi_params
    prm = oracle.gi_params_from(gi_params(50.0, 2))
    out = oracle.trace_gi(g, df, iv, ip, p["t"], p["face"], prm)
    rgb, sh = out["rgb"].astype(np.float64), out["sh"].astype(np.float64)
    co = rgb[..., 0] - rgb[..., 2]
    cg = rgb[..., 1] - (rgb[..., 2] + co / 2)
    Y = np.maximum(rgb[..., 2] + co / 2 + cg / 2, 0)
    np.testing.assert_allclose(sh[..., 4], co, atol=1e-6)
    np.testing.assert_allclose(sh[..., 5], cg, atol=1e-6)
    np.testing.assert_allclose(3.544905 * sh[..., 3], Y, rtol=1e-5, atol=1e-6)
    hit = p["face"] != 255
    band1 = np.linalg.norm(sh[..., 0:3], axis=-1)
    np.testing.assert_allclose(band1[hit], 0.488603 * Y[hit], rtol=1e-4, atol=1e-6)
    assert (rgb >= 0).all() and (rgb <= 8).all()
    lum = rgb @ np.array([0.299, 0.587, 0.114])
    np.testing.assert_allclose(out["luma"][hit], np.clip(np.maximum(lum[hit], 0.01), 0.001, 64), rtol=1e-5)


def test_gi_cosine_lobe_statistics():
    """The sampled directions are cosine-weighted about the normal:
    E[cos] = 2/3, and the directions are unit length.  Read back from the SH
    band 1 of sky-hit (Y known) pixels."""
    g = np.zeros((64, 32, 64), np.uint8)
    g[:, 0:4, :] = 1                                 # a flat floor, open sky above
    df = oracle.df_build(g)
    W, H = 128, 128
    cam, view, proj, iv, ip = camera(W, H, pos=(32.0, 20.0, 32.0), yaw=0.0, pitch=-89.0)
    p = oracle.trace_primary(g, df, iv, ip, W, H, 350)
    assert (p["face"] == 2).all()
    from voxelpathtracer_amd import gi_params
    prm = gi_params(50.0, 0)
    out = oracle.trace_gi(g, df, iv, ip, p["t"], p["face"], oracle.gi_params_from(prm))
    sh = out["sh"].astype(np.float64)
    Y = 3.544905 * sh[..., 3]
    dirs = sh[..., 0:3] / (0.488603 * Y[..., None])
    cosv = dirs[..., 1]                              # normal = +y
    assert abs(np.linalg.norm(dirs, axis=-1) - 1).max() < 1e-3
    assert (cosv > 0).all()
    assert abs(cosv.mean() - 2 / 3) < 0.01
    # escaped rays carry the constant sky radiance
    assert (out["aosky"][..., 1] == 1).all()
    np.testing.assert_allclose(out["rgb"], np.broadcast_to(np.float32(prm.sky_radiance), out["rgb"].shape),
                               rtol=1e-6)


# ---------------------------------------------------------------- SVGF

def _flat_frame(W, H, face=2, t=10.0):
    return dict(t=np.full((H, W), t, np.float32), face=np.full((H, W), face, np.uint8),
                block=np.full((H, W), 3, np.uint8))


def test_svgf_constant_image_identity():
    """A constant image over a flat G-buffer is a fixed point of the
    pre-spatial and a-trous passes (their weights are normalised)."""
    W, H = 48, 32
    cam, view, proj, iv, ip = camera(W, H, pos=(10.0, 20.0, 10.0), yaw=0.0, pitch=-89.0)
    gb = _flat_frame(W, H)
    sh = np.broadcast_to(f16([0.3, 0.1, -0.2, 0.5]), (H, W, 4)).copy()
    cc = np.broadcast_to(f16([0.05, -0.02]), (H, W, 2)).copy()
    ao = np.full((H, W, 2), np.float32(200 / 255))
    pre = oracle.svgf_prespatial(iv, ip, gb["t"], gb["face"], sh, cc, np.full((H, W), 0.5, np.float32), ao)
    np.testing.assert_allclose(pre["sh"], sh, rtol=2e-3)
    np.testing.assert_allclose(pre["cocg"], cc, rtol=2e-3)
    var = np.full((H, W), f16(0.02), np.float32)
    util = np.zeros((H, W, 3), np.float32)
    util[..., 0] = 20
    for step in (16, 8, 4, 2, 1):
        out = oracle.svgf_atrous(gb["t"], gb["face"], sh, cc, var, ao, util, step, time=0.3)
        np.testing.assert_allclose(out["sh"], sh, rtol=2e-3)
        np.testing.assert_allclose(out["cocg"], cc, rtol=2e-3, atol=1e-6)
        assert (out["var"] <= var + 1e-6).all()     # sum w^2 v / (sum w)^2 <= v


def test_svgf_temporal_running_mean():
    """Static camera, static G-buffer: the history is a running mean with
    blend = max(1 / (n + 1), 0.05) (TemporalFilter.glsl:315-345)."""
    W, H = 40, 24
    cam, view, proj, iv, ip = camera(W, H, pos=(10.0, 20.0, 10.0), yaw=0.0, pitch=-89.0)
    from voxelpathtracer_amd import viewproj
    vp = viewproj(view, proj)
    gb = _flat_frame(W, H)
    hist = dict(sh=np.zeros((H, W, 4), np.float32), cocg=np.zeros((H, W, 2), np.float32),
                util=np.zeros((H, W, 3), np.float32), ao=np.zeros((H, W, 2), np.float32))
    rng = np.random.default_rng(3)
    values = f16(rng.uniform(0.1, 1.0, 30))
    mean = 0.0
    for n, v in enumerate(values):
        cur = dict(sh=np.full((H, W, 4), v, np.float32), cocg=np.zeros((H, W, 2), np.float32),
                   luma=np.full((H, W), v, np.float32), ao=np.ones((H, W, 2), np.float32))
        hist = oracle.svgf_temporal(iv, ip, vp, gb, gb, cur, hist)
        # frame 0 sees a zero history with spp 0 -> blend 1 (the input)
        blend = max(1.0 / (n + 1), 0.05)
        mean = mean * (1 - blend) + float(v) * blend
        inner = hist["sh"][H // 2, W // 2, 3]
        assert abs(inner - mean) < 2e-3 * max(1.0, mean), (n, inner, mean)
        assert hist["util"][H // 2, W // 2, 0] == min(n + 1, 2048)


def test_svgf_disocclusion_passthrough():
    """When no history tap is valid (the previous G-buffer is sky where the
    current one is geometry) the temporal output is the current input."""
    W, H = 40, 24
    cam, view, proj, iv, ip = camera(W, H, pos=(10.0, 20.0, 10.0), yaw=0.0, pitch=-89.0)
    from voxelpathtracer_amd import viewproj
    vp = viewproj(view, proj)
    gb = _flat_frame(W, H)
    prev_gb = dict(t=np.full((H, W), -1.0, np.float32), face=np.full((H, W), 255, np.uint8),
                   block=np.zeros((H, W), np.uint8))
    rng = np.random.default_rng(4)
    cur = dict(sh=f16(rng.uniform(0, 1, (H, W, 4))), cocg=f16(rng.uniform(-0.1, 0.1, (H, W, 2))),
               luma=f16(rng.uniform(0, 1, (H, W))), ao=np.ones((H, W, 2), np.float32))
    hist = dict(sh=np.full((H, W, 4), 5.0, np.float32), cocg=np.zeros((H, W, 2), np.float32),
                util=np.full((H, W, 3), 7.0, np.float32), ao=np.zeros((H, W, 2), np.float32))
    out = oracle.svgf_temporal(iv, ip, vp, gb, prev_gb, cur, hist)
    np.testing.assert_array_equal(out["sh"], cur["sh"])
    np.testing.assert_array_equal(out["cocg"], cur["cocg"])
    assert (out["util"][..., 0] == 0).all()


def test_sh_decode_identity():
    """SHToIrridiance with v = 0: Y = 2 * 0.886226 * 0.282095 * Y0 = Y0 / 2
    ... for an isotropic SH (band 1 = 0) the decoded grey level is
    2 * 0.886226 * w, and CoCg rescale by Y * 0.282095 / w."""
    n = 5
    w = np.float32([0.1, 0.2, 0.5, 1.0, 2.0])
    sh = np.zeros((1, n, 4), np.float32)
    sh[0, :, 3] = w
    cc = np.zeros((1, n, 2), np.float32)
    face = np.full((1, n), 2, np.uint8)
    rgb = oracle.sh_decode(sh, cc, face)
    np.testing.assert_allclose(rgb[0, :, 0], 2 * 0.886226 * w, rtol=1e-6)
    np.testing.assert_allclose(rgb[0], np.repeat(rgb[0, :, :1], 3, 1), rtol=1e-6)


# ---------------------------------------------------------------- camera

def test_camera_closed_form():
    """glm lookAt/perspective (FpsCamera.cpp:23-24,152-160): the product's
    camera equals the oracle's closed form, and inv_view maps the eye-space
    forward axis to the yaw/pitch front vector."""
    W, H = 1920, 1080
    cam, view, proj, iv, ip = camera(W, H)
    oc = oracle.camera((192.5, 100.5, 192.5), 45.0, -20.0, 60.0, W / H)
    np.testing.assert_allclose(iv, oc["inv_view"], atol=1e-6)
    np.testing.assert_allclose(ip, oc["inv_proj"], rtol=1e-6, atol=1e-7)
    IV = iv.reshape(4, 4).T.astype(np.float64)     # row-major
    fwd = IV[:3, :3] @ np.array([0, 0, -1.0])
    p, y = np.radians(-20), np.radians(45)
    np.testing.assert_allclose(fwd, [np.cos(p) * np.cos(y), np.sin(p), np.cos(p) * np.sin(y)], atol=1e-6)
    np.testing.assert_allclose(IV[:3, 3], [192.5, 100.5, 192.5])
    V = view.reshape(4, 4).T.astype(np.float64)
    P = proj.reshape(4, 4).T.astype(np.float64)
    np.testing.assert_allclose(V @ IV, np.eye(4), atol=1e-5)
    np.testing.assert_allclose(P @ ip.reshape(4, 4).T.astype(np.float64), np.eye(4), atol=1e-5)
    assert abs(P[1, 1] - 1 / np.tan(np.radians(30))) < 1e-5
