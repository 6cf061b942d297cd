"""GPU parity of the mirror-reflection pass (ReflectionTraceFrag.glsl:
764-1035, traversal core) with the oracle: the binary16 hit distance and
the hit record (face, block, sun shadow at the hit) of every target pixel,
bit for bit, at full and reduced resolution, plus the traced-ray and DF
sample counts."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from tests.helpers import camera, random_world, terrain
from tests.test_reflect_cpu import _corridor
from voxelpathtracer_amd import World, gi_params
from voxelpathtracer_amd.world import DeviceGBuffer

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def terrain_world(ctx):
    g = terrain(384, 128, 384)
    w = World(ctx, 384, 128, 384)
    w.m_WorldData[...] = g
    w.Buffer()
    w.GenerateDistanceField()
    yield g, oracle.df_build(g), w
    w.close()


def _world(ctx, g):
    nz, ny, nx = g.shape
    w = World(ctx, nx, ny, nz)
    w.m_WorldData[...] = g
    w.Buffer()
    w.GenerateDistanceField()
    return w


def _check(w, g, df, W, H, div, light, pos=None, yaw=None, pitch=None, max_iter=64, shadow_iter=150, spp=1, rough=0,
           roughness=0.425, frame=0):
    import torch
    kw = {k: v for k, v in (("pos", pos), ("yaw", yaw), ("pitch", pitch)) if v is not None}
    cam, view, proj, iv, ip = camera(W, H, **kw)
    gb = DeviceGBuffer(W, H)
    Wr, Hr = W // div, H // div
    dist = torch.full((Hr, Wr), 0x7E00, dtype=torch.int16, device="cuda")
    hit = torch.full((Hr, Wr), 7, dtype=torch.int32, device="cuda")
    print(f"reflect {W}x{H} div {div} iter {max_iter}/{shadow_iter} spp {spp} rough {rough} {roughness}", flush=True)
    w.trace_primary(cam, gb, 350)
    ctr = torch.zeros(2, dtype=torch.int64, device="cuda")
    w.set_sample_counter(ctr)
    w.trace_reflect(cam, gb, light, dist, hit, div=div, max_iter=max_iter, shadow_iter=shadow_iter, spp=spp,
                    rough=rough, roughness=roughness, frame=frame)
    torch.cuda.synchronize()
    w.set_sample_counter(None)
    host = gb.numpy()
    ref = oracle.trace_reflect(g, df, iv, ip, host["t"], host["face"], light, div=div, max_iter=max_iter,
                               shadow_iter=shadow_iter, spp=spp, rough=rough, roughness=roughness, frame=frame)
    np.testing.assert_array_equal(dist.cpu().numpy().view(np.uint16), ref["dist"], err_msg=f"dist {W}x{H}/{div}")
    np.testing.assert_array_equal(hit.cpu().numpy().view(np.uint32), ref["hit"], err_msg=f"hit {W}x{H}/{div}")
    c = ctr.cpu().numpy()
    assert c[0] == ref["samples"].sum() and c[1] == ref["rays"].sum()
    return ref


def test_reflect_terrain_small(ctx, terrain_world):
    g, df, w = terrain_world
    sun = np.float32(list(gi_params(50.0, 0).sun_dir))
    for div in (1, 2, 4):
        ref = _check(w, g, df, 320, 180, div, sun)
    hits = ref["hit"] != 255
    assert hits.mean() > 0.05 and ((ref["hit"] >> 16) & 1)[hits].any()


def test_reflect_terrain_1080p(ctx, terrain_world):
    g, df, w = terrain_world
    sun = np.float32(list(gi_params(50.0, 0).sun_dir))
    _check(w, g, df, 1920, 1080, 4, sun)
    _check(w, g, df, 1920, 1080, 1, sun)


def test_reflect_reference_defaults_1080p(ctx, terrain_world):
    """The reference's settings: 1/4 resolution, 2 rough samples, 64 / 150
    steps (vpt_reflect_params_default)."""
    from voxelpathtracer_amd.world import reflect_params
    g, df, w = terrain_world
    sun = np.float32(list(gi_params(50.0, 0).sun_dir))
    kw = reflect_params(sun, frame=7)
    assert (kw["div"], kw["spp"], kw["rough"], kw["max_iter"], kw["shadow_iter"]) == (4, 2, 1, 64, 150)
    ref = _check(w, g, df, 1920, 1080, kw["div"], sun, max_iter=kw["max_iter"], shadow_iter=kw["shadow_iter"],
                 spp=kw["spp"], rough=kw["rough"], roughness=kw["roughness"], frame=kw["frame"])
    hits = ref["hit"] >> 24
    assert (hits == 1).any() and (hits == 2).any()      # samples of one pixel hit and miss independently


def test_reflect_rough_sweep(ctx, terrain_world):
    g, df, w = terrain_world
    sun = np.float32(list(gi_params(50.0, 0).sun_dir))
    # (rough, spp 2: a lane per sample, 16x8-pixel blocks, partial at these sizes)
    for spp, rough, roughness, div, frame in ((1, 1, 0.3, 1, 0), (2, 0, 0.5, 1, 0), (8, 1, 0.9, 2, 3),
                                              (16, 1, 0.01, 4, 11), (5, 1, 1.7, 1, 2), (2, 1, 0.6, 1, 5),
                                              (2, 1, 0.2, 3, 1)):
        _check(w, g, df, 320, 180, div, sun, spp=spp, rough=rough, roughness=roughness, frame=frame)


def test_reflect_iteration_caps(ctx, terrain_world):
    g, df, w = terrain_world
    sun = np.float32(list(gi_params(30.0, 0).sun_dir))
    for mi, si in ((1, 1), (8, 3), (128, 350)):
        _check(w, g, df, 256, 144, 4 if mi == 128 else 1, sun, max_iter=mi, shadow_iter=si)


def test_reflect_ragged_and_tiny(ctx, terrain_world):
    """Target sizes that do not divide the G-buffer, a one-pixel target,
    and divisors 3 and 16."""
    g, df, w = terrain_world
    sun = np.float32(list(gi_params(50.0, 0).sun_dir))
    for W, H, div in ((203, 117, 4), (203, 117, 3), (17, 16, 16), (5, 3, 1)):
        _check(w, g, df, W, H, div, sun)


def test_reflect_corridor(ctx):
    """The corridor of the closed-form CPU test: wall-to-wall, floor and
    short reflections, shadowed and lit hits."""
    g = _corridor()
    df = oracle.df_build(g)
    w = _world(ctx, g)
    sun = np.array([0.35, 1.0, 0.25])
    sun = (sun / np.linalg.norm(sun)).astype(np.float32)
    ref = _check(w, g, df, 96, 64, 1, sun, pos=(30.5, 17.5, 3.5), yaw=70.0, pitch=-28.0, max_iter=350,
                 shadow_iter=350)
    assert ((ref["hit"] >> 16) & 1).any() and (((ref["hit"] >> 16) & 1) == 0)[ref["hit"] != 255].any()
    w.close()


def test_reflect_random_world(ctx):
    rng = np.random.default_rng(11)
    g = random_world(rng, 96, 64, 80, 0.08)
    df = oracle.df_build(g)
    w = _world(ctx, g)
    light = np.float32([-0.3, 0.8, 0.52])
    light /= np.linalg.norm(light)
    _check(w, g, df, 160, 96, 1, light, pos=(48.5, 30.5, 2.5), yaw=80.0, pitch=-12.0)
    _check(w, g, df, 160, 96, 2, light, pos=(48.5, 30.5, 2.5), yaw=80.0, pitch=-12.0)
    w.close()


def test_reflect_argument_errors(ctx, terrain_world):
    import torch
    from voxelpathtracer_amd._lib import VptError
    g, df, w = terrain_world
    cam, view, proj, iv, ip = camera(64, 32)
    gb = DeviceGBuffer(64, 32)
    d = torch.zeros((32, 64), dtype=torch.int16, device="cuda")
    sun = (0.0, 1.0, 0.0)
    for kw in (dict(div=0), dict(div=17), dict(div=33), dict(max_iter=-1), dict(shadow_iter=-1), dict(spp=0),
               dict(spp=17), dict(rough=2), dict(roughness=float("nan"))):
        with pytest.raises(VptError):
            w.trace_reflect(cam, gb, sun, d, **kw)
    with pytest.raises(VptError):
        w.trace_reflect(cam, gb, sun, None)
    gb2 = DeviceGBuffer(32, 32)
    with pytest.raises(VptError):
        w.trace_reflect(cam, gb2, sun, d)
