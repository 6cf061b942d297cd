"""GPU parity of the shaded reflection pass (ReflectionTraceFrag.glsl:
582-593, 764-1035) and the colour composite (ColorPassFrag.glsl:773-1010)
with the oracle, bit for bit: the reflection's binary16 distance, hit
record (with the reprojection bit) and RGBA binary16 colour, its ray and DF
sample counts, and the composite's float32 RGB.  Inputs are a real frame's
denoised planes: a few frames of the camera path through the GI pass, the
SVGF chain and the sun-shadow denoiser on the GPU, whose outputs the oracle
reads as well."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from tests.helpers import camera, svgf_camera_path, terrain
from voxelpathtracer_amd import SVGF, ShadowDenoiser, World, decode_texels, gi_params, viewproj
from voxelpathtracer_amd.world import DeviceGBuffer, composite, composite_params, reflect_params, reflect_shading

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scene(ctx):
    g = terrain(384, 128, 384)
    w = World(ctx, 384, 128, 384)
    w.m_WorldData[...] = g
    w.Buffer()
    w.GenerateDistanceField()
    yield g, oracle.df_build(g), w
    w.close()


def _frame(ctx, w, W, H, frames=3):
    """The last of `frames` frames of the camera path with its denoised
    diffuse (SVGF final texels, also as vpt_svgf_final's pointer) and
    denoised sun shadow."""
    import torch
    sv, sd = SVGF(ctx, W, H), ShadowDenoiser(ctx, W, H)
    gbs = [DeviceGBuffer(W, H), DeviceGBuffer(W, H)]
    rgb = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
    tex = torch.empty((H, W, 8), dtype=torch.int16, device="cuda")
    out = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
    fin = torch.empty((H, W, 8), dtype=torch.int16, device="cuda")
    mask = torch.empty((H, W), dtype=torch.uint8, device="cuda")
    trans = torch.empty((H, W), dtype=torch.int16, device="cuda")
    shadow = torch.empty((H, W), dtype=torch.uint8, device="cuda")
    prev_vp = None
    for f in range(frames):
        pos, yaw, pitch = svgf_camera_path(f)
        cam, view, proj, iv, ip = camera(W, H, pos, yaw, pitch)
        vp = viewproj(view, proj)
        cur, prev = gbs[f % 2], gbs[(f + 1) % 2] if f else gbs[0]
        prm = gi_params(frame=f)
        w.trace_primary(cam, cur, 350)
        w.trace_gi(cam, cur, prm, rgb, tex)
        sv.frame(cam, vp if prev_vp is None else prev_vp, cur, prev, tex, out, time=f / 60.0, texels_out=fin)
        w.trace_shadow2(cam, cur, list(prm.sun_dir), mask, trans, 350)
        sd.frame(cam, vp if prev_vp is None else prev_vp, cur, prev, mask, trans, shadow)
        prev_vp = vp
    torch.cuda.synchronize()
    return dict(cam=cam, iv=iv, ip=ip, vp=vp, gb=cur, fin=fin, final_ptr=sv.final(), shadow=shadow, prm=prm,
                frame=frames - 1, keep=(sv, sd, gbs))


def _host_planes(fr):
    tx = decode_texels(fr["fin"].cpu().numpy().view(np.uint16))
    sh = fr["shadow"].cpu().numpy().astype(np.float32) / np.float32(255.0)
    return tx, sh


def _reflect(w, fr, W, H, div, spp, rough, roughness=0.425, max_iter=64, shadow_iter=150):
    import torch
    Wr, Hr = W // div, H // div
    dist = torch.full((Hr, Wr), 0x7E00, dtype=torch.int16, device="cuda")
    hit = torch.full((Hr, Wr), 7, dtype=torch.int32, device="cuda")
    color = torch.full((Hr, Wr, 4), 0x7E00, dtype=torch.int16, device="cuda")
    ctr = torch.zeros(2, dtype=torch.int64, device="cuda")
    prm = fr["prm"]
    shading = reflect_shading(prm, fr["vp"], fr["final_ptr"], fr["shadow"])
    w.set_sample_counter(ctr)
    w.trace_reflect_shaded(fr["cam"], fr["gb"], list(prm.sun_dir), shading, dist, color, hit, div=div,
                           max_iter=max_iter, shadow_iter=shadow_iter, spp=spp, rough=rough, roughness=roughness,
                           frame=fr["frame"])
    torch.cuda.synchronize()
    w.set_sample_counter(None)
    return dist, hit, color, ctr


CASES = [  # W, H, div, spp, rough
    (1920, 1080, 4, 2, 1),    # the reference's settings
    (1920, 1080, 4, 2, 0),
    (640, 360, 1, 2, 1),
    (640, 360, 1, 1, 0),
    (640, 360, 2, 4, 1),
    (333, 187, 3, 3, 1),      # ragged
]


@pytest.mark.parametrize("W,H,div,spp,rough", CASES)
def test_reflect_shaded_parity(ctx, scene, W, H, div, spp, rough):
    g, df, w = scene
    print(f"reflect shaded {W}x{H} div {div} spp {spp} rough {rough}", flush=True)
    fr = _frame(ctx, w, W, H)
    dist, hit, color, ctr = _reflect(w, fr, W, H, div, spp, rough)
    tx, shadow = _host_planes(fr)
    host = fr["gb"].numpy()
    prm = fr["prm"]
    ref = oracle.trace_reflect_shaded(g, df, fr["iv"], fr["ip"], host["t"], host["face"], np.float32(list(prm.sun_dir)),
                                      tx["sh"], tx["cocg"], tx["ao"][..., 0], shadow, fr["vp"],
                                      np.float32(list(prm.light_color)), np.float32(list(prm.sky_radiance)),
                                      np.float32(list(prm.albedo)), div=div, max_iter=64, shadow_iter=150, spp=spp,
                                      rough=rough, roughness=0.425, frame=fr["frame"])
    np.testing.assert_array_equal(dist.cpu().numpy().view(np.uint16), ref["dist"], err_msg="dist")
    np.testing.assert_array_equal(hit.cpu().numpy().view(np.uint32), ref["hit"], err_msg="hit")
    np.testing.assert_array_equal(color.cpu().numpy().view(np.uint16), ref["color"], err_msg="colour")
    c = ctr.cpu().numpy()
    assert c[0] == ref["samples"].sum() and c[1] == ref["rays"].sum()
    hits = ref["hit"] != 255
    reproj = ((ref["hit"] >> 17) & 1).astype(bool) & hits
    # both branches of the reprojection test are exercised
    assert 0.05 < reproj.sum() / max(hits.sum(), 1) < 0.98, (reproj.sum(), hits.sum())


def test_reflect_shaded_matches_unshaded_geometry(ctx, scene):
    """The shaded pass's distances equal the unshaded pass's, and its hit
    records agree outside the reprojection / shadow bits (rough: the same
    rays; the shadow ray of reprojected samples is not traced)."""
    import torch
    g, df, w = scene
    W, H = 640, 360
    fr = _frame(ctx, w, W, H, frames=2)
    dist, hit, color, _ = _reflect(w, fr, W, H, 2, 2, 1)
    d0 = torch.empty_like(dist)
    h0 = torch.empty_like(hit)
    w.trace_reflect(fr["cam"], fr["gb"], list(fr["prm"].sun_dir), d0, h0, div=2, spp=2, rough=1, roughness=0.425,
                    frame=fr["frame"])
    torch.cuda.synchronize()
    assert torch.equal(dist, d0)
    m = ~np.uint32((1 << 16) | (1 << 17))
    np.testing.assert_array_equal(hit.cpu().numpy().view(np.uint32) & m, h0.cpu().numpy().view(np.uint32) & m)


@pytest.mark.parametrize("W,H,rdiv", [(1920, 1080, 4), (640, 360, 1), (333, 187, 3)])
def test_composite_parity(ctx, scene, W, H, rdiv):
    import torch
    g, df, w = scene
    print(f"composite {W}x{H} reflection div {rdiv}", flush=True)
    fr = _frame(ctx, w, W, H)
    _, _, color, _ = _reflect(w, fr, W, H, rdiv, 2, 1)
    cp = composite_params(fr["prm"])
    rgb = torch.full((H, W, 3), float("nan"), dtype=torch.float32, device="cuda")
    composite(ctx, fr["cam"], fr["gb"], cp, fr["final_ptr"], fr["shadow"], color, rdiv, rgb)
    torch.cuda.synchronize()
    tx, shadow = _host_planes(fr)
    host = fr["gb"].numpy()
    refl = color.cpu().numpy().view(np.float16).astype(np.float32)
    ref = oracle.composite(fr["iv"], fr["ip"], host["t"], host["face"], host["block"], tx["sh"], tx["cocg"],
                           tx["ao"][..., 0], shadow, refl, rdiv, np.float32(list(cp.sun_dir)),
                           np.float32(list(cp.moon_dir)), np.float32(list(cp.sun_radiance)),
                           np.float32(list(cp.moon_radiance)), np.float32(list(cp.sky_pixel_radiance)),
                           np.float32(list(cp.albedo)), cp.roughness)
    got = rgb.cpu().numpy()
    np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32), err_msg="composite")
    sky = host["t"] < 0
    assert sky.any() and (~sky).mean() > 0.3


def test_composite_argument_errors(ctx, scene):
    import torch
    from voxelpathtracer_amd._lib import VptError
    _, _, w = scene
    W, H = 64, 32
    cam, *_ = camera(W, H)
    gb = DeviceGBuffer(W, H)
    cp = composite_params(gi_params())
    dif = torch.zeros((H, W, 8), dtype=torch.int16, device="cuda")
    sh = torch.zeros((H, W), dtype=torch.uint8, device="cuda")
    refl = torch.zeros((H // 4, W // 4, 4), dtype=torch.int16, device="cuda")
    rgb = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
    with pytest.raises(VptError):
        composite(ctx, cam, gb, cp, dif, sh, refl, 0, rgb)
    cp.roughness = 1.5
    with pytest.raises(VptError):
        composite(ctx, cam, gb, cp, dif, sh, refl, 4, rgb)
    with pytest.raises(VptError):   # shading without its planes
        shading = reflect_shading(gi_params(), np.eye(4, dtype=np.float32).ravel(), 0, sh)
        d = torch.empty((H // 4, W // 4), dtype=torch.int16, device="cuda")
        w.trace_reflect_shaded(cam, gb, (0.0, 1.0, 0.0), shading, d, refl, div=4)
    assert reflect_params((0.0, 1.0, 0.0))["spp"] == 2


@pytest.mark.parametrize("W,H,div,spp,rough", [(1920, 1080, 4, 2, 1), (1920, 1080, 4, 2, 0), (640, 360, 1, 3, 1),
                                               (333, 187, 3, 5, 1), (640, 360, 2, 4, 0)])
def test_reflect_deferred_equals_shaded(ctx, scene, W, H, div, spp, rough):
    """vpt_trace_reflect_deferred + vpt_shade_reflect equal one
    vpt_trace_reflect_shaded call bit for bit (distances, hit records,
    colours, ray and sample counts)."""
    import torch
    from voxelpathtracer_amd.world import RECORD_BYTES, shade_reflect
    g, df, w = scene
    print(f"reflect deferred {W}x{H} div {div} spp {spp} rough {rough}", flush=True)
    fr = _frame(ctx, w, W, H, frames=2)
    dist, hit, color, ctr = _reflect(w, fr, W, H, div, spp, rough)
    Wr, Hr = W // div, H // div
    d2 = torch.full_like(dist, 0x7E00)
    h2 = torch.full_like(hit, 7)
    c2 = torch.full_like(color, 0x7E00)
    recs = torch.full((spp * Hr * Wr * RECORD_BYTES // 4,), -1, dtype=torch.int32, device="cuda")
    ctr2 = torch.zeros(2, dtype=torch.int64, device="cuda")
    prm = fr["prm"]
    kw = dict(div=div, max_iter=64, shadow_iter=150, spp=spp, rough=rough, roughness=0.425, frame=fr["frame"])
    w.set_sample_counter(ctr2)
    w.trace_reflect_deferred(fr["cam"], fr["gb"], list(prm.sun_dir), fr["vp"], d2, recs, h2, **kw)
    torch.cuda.synchronize()
    w.set_sample_counter(None)
    shading = reflect_shading(prm, fr["vp"], fr["final_ptr"], fr["shadow"])
    shade_reflect(ctx, fr["cam"], fr["gb"], list(prm.sun_dir), shading, recs, c2, h2, **kw)
    torch.cuda.synchronize()
    assert torch.equal(d2, dist) and torch.equal(h2, hit) and torch.equal(c2, color)
    assert torch.equal(ctr2, ctr)
