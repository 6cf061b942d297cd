"""vpt_trace_gi_noalloc (the GI pass over a fixed scratch array, no device
allocation) against vpt_trace_gi through the C ABI: bit-identical radiance
and SVGF texels on two calls in a row, the second on a smaller image, and
sizes beyond the fixed scratch refused."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from tests.helpers import camera, terrain
from voxelpathtracer_amd import World, gi_params
from voxelpathtracer_amd._lib import VPT_EINVAL, VptError, call
from voxelpathtracer_amd.world import DeviceGBuffer

pytestmark = pytest.mark.gpu


def _gi(fn, w, cam, gb, prm, rgb, tex):
    call(fn, w.handle, C.byref(cam), C.byref(gb.struct()), C.byref(prm), C.c_void_p(rgb.data_ptr()),
         C.c_void_p(tex.data_ptr()), None)


@pytest.fixture(scope="module")
def world(ctx):
    g = terrain(384, 128, 384)
    w = World(ctx, 384, 128, 384)
    w.m_WorldData[...] = g
    w.Buffer()
    w.GenerateDistanceField()
    yield w
    w.close()


def test_noalloc_equals_allocating_entry_point(world):
    import torch
    for W, H, pos, frame, seed in ((1920, 1080, (192.5, 100.5, 192.5), 3, 0x5EED0003),
                                   (640, 360, (120.5, 105.5, 260.5), 4, 0x00C0FFEE)):
        print(f"noalloc GI {W}x{H}", flush=True)
        cam, *_ = camera(W, H, pos=pos)
        gb = DeviceGBuffer(W, H)
        world.trace_primary(cam, gb, 350, stream=0)
        prm = gi_params(50.0, frame, albedo_seed=seed)
        ref_rgb = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
        ref_tex = torch.empty((H, W, 8), dtype=torch.int16, device="cuda")
        _gi("vpt_trace_gi", world, cam, gb, prm, ref_rgb, ref_tex)
        rgb = torch.full((H, W, 3), -1.0, dtype=torch.float32, device="cuda")
        tex = torch.full((H, W, 8), -1, dtype=torch.int16, device="cuda")
        _gi("vpt_trace_gi_noalloc", world, cam, gb, prm, rgb, tex)
        torch.cuda.synchronize()
        assert torch.equal(rgb, ref_rgb), f"{W}x{H}: radiance differs"
        assert torch.equal(tex, ref_tex), f"{W}x{H}: SVGF texels differ"
        assert float(ref_rgb.abs().sum()) > 0


def test_noalloc_rejects_oversize(world):
    import torch
    W, H = 2048, 1088                      # 2,228,224 pixels > 2,097,152
    cam, *_ = camera(W, H)
    gb = DeviceGBuffer(W, H)
    prm = gi_params(50.0, 0)
    rgb = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
    tex = torch.empty((H, W, 8), dtype=torch.int16, device="cuda")
    with pytest.raises(VptError) as e:
        _gi("vpt_trace_gi_noalloc", world, cam, gb, prm, rgb, tex)
    assert e.value.code == VPT_EINVAL
    assert "exceeds" in str(e.value)
