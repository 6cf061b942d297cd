"""GPU tests of calls that share a context across streams and of the
edit-list rule: a GI pass on one stream never reads another stream's scratch
or albedo table, frame requests with different albedo tables run side by
side in the server, and a voxel named twice in one vpt_world_set_blocks call
keeps its last id."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from tests.helpers import camera, terrain
from voxelpathtracer_amd import Server, World, gi_params
from voxelpathtracer_amd._lib import call
from voxelpathtracer_amd.world import DeviceGBuffer

pytestmark = pytest.mark.gpu


def _world(ctx, grid):
    nz, ny, nx = grid.shape
    w = World(ctx, nx, ny, nz)
    w.m_WorldData[...] = grid
    w.Buffer()
    w.GenerateDistanceField()
    return w


@pytest.fixture(scope="module")
def two_worlds(ctx):
    a = _world(ctx, terrain(384, 128, 384))
    b = _world(ctx, terrain(256, 96, 256, seed=0x5EED0101, seed_random=0x5EED0102))
    yield a, b
    a.close()
    b.close()


def test_trace_gi_streams_do_not_share_scratch_or_albedo(ctx, two_worlds):
    """vpt_trace_gi on two streams of one context, on two worlds and on one
    world, with different albedo tables and resolutions, each repeated so
    that launches of both streams are in flight together: every output
    equals the same call made alone on the default stream."""
    import torch
    wa, wb = two_worlds
    cases = [  # (world, W, H, camera position, albedo seed)
        (wa, 1920, 1080, (192.5, 100.5, 192.5), 0x5EED0003),
        (wb, 1280, 720, (128.5, 80.5, 128.5), 0x00C0FFEE),
        (wa, 960, 540, (100.5, 110.5, 250.5), 0x0BADF00D),
    ]
    setup = []
    for w, W, H, pos, seed in cases:
        cam, *_ = camera(W, H, pos=pos)
        gb = DeviceGBuffer(W, H)
        w.trace_primary(cam, gb, 350)
        prm = gi_params(50.0, 7, albedo_seed=seed)
        setup.append((w, cam, gb, prm, W, H))
    torch.cuda.synchronize()
    refs = []
    for w, cam, gb, prm, W, H in setup:
        rgb = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
        tex = torch.empty((H, W, 8), dtype=torch.int16, device="cuda")
        w.trace_gi(cam, gb, prm, rgb, tex)
        torch.cuda.synchronize()
        refs.append((rgb.cpu(), tex.cpu()))
    # the albedo tables differ, so a mix-up would show in the radiance
    assert not torch.equal(refs[0][0][:540, :960], refs[2][0])
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    outs = [(torch.zeros((H, W, 3), dtype=torch.float32, device="cuda"),
             torch.zeros((H, W, 8), dtype=torch.int16, device="cuda")) for _, _, _, _, W, H in setup]
    torch.cuda.synchronize()
    for rep in range(4):
        for k, (w, cam, gb, prm, W, H) in enumerate(setup):
            s = streams[(k + rep) % 2]
            w.trace_gi(cam, gb, prm, outs[k][0], outs[k][1], stream=s.cuda_stream)
    torch.cuda.synchronize()
    for k, ((rgb, tex), (ref_rgb, ref_tex)) in enumerate(zip(outs, refs)):
        assert torch.equal(rgb.cpu(), ref_rgb), f"case {k}: radiance differs from the call alone"
        assert torch.equal(tex.cpu(), ref_tex), f"case {k}: SVGF texels differ from the call alone"


def test_server_frames_with_different_albedo(ctx, two_worlds):
    """Two frame requests with different albedo tables in batched mode with
    two launches in flight: each equals its one-launch passes."""
    import torch
    wa, _ = two_worlds
    W, H = 640, 360
    cam, *_ = camera(W, H)
    prms = [gi_params(50.0, 2, albedo_seed=s) for s in (0x5EED0003, 0x00C0FFEE, 0x0BADF00D, 0x12345678)]
    refs = []
    for prm in prms:
        gb = DeviceGBuffer(W, H)
        mask = torch.empty((H, W), dtype=torch.uint8, device="cuda")
        rgb = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
        wa.trace_primary(cam, gb, 350)
        wa.trace_shadow(cam, gb, list(prm.sun_dir), mask, 350)
        wa.trace_gi(cam, gb, prm, rgb)
        torch.cuda.synchronize()
        refs.append((mask.cpu().reshape(-1), rgb.cpu().reshape(-1)))
    srv = Server(ctx, mode="batched", max_worlds=2, arena_bytes=256 << 20, max_inflight=2, world_budget=1 << 40)
    slot = srv.add_world(wa)
    tickets = [srv.submit_frame(slot, cam, prm, 350, 350) for prm in prms]
    for k, t in enumerate(tickets):
        out = srv.fetch(t)
        srv.release(t)
        assert torch.equal(out["mask"].cpu(), refs[k][0]), f"frame {k}: shadow mask"
        assert torch.equal(out["rgb"].cpu(), refs[k][1]), f"frame {k}: radiance differs from its own albedo table"
    st = srv.stats()
    assert st["completed"] == len(prms)
    srv.close()


@pytest.mark.parametrize("n_unique", [5, 300])
def test_set_blocks_duplicates_keep_last(ctx, n_unique):
    """vpt_world_set_blocks with voxels named several times: the grid holds
    each voxel's last id, through both the argument-list kernel (<= 128
    edits after merging) and the device-list kernel."""
    nx, ny, nz = 64, 32, 64
    g = np.zeros((nz, ny, nx), np.uint8)
    w = _world(ctx, g)
    rng = np.random.default_rng(n_unique)
    xs, ys, zs = rng.integers(0, nx, n_unique), rng.integers(0, ny, n_unique), rng.integers(0, nz, n_unique)
    coords = (xs | (ys << 10) | (zs << 20)).astype(np.uint32)
    coords = np.unique(coords)
    edits, ids = [], []
    for rep in range(4):                  # each voxel four times, ids differing
        for c in coords:
            edits.append(c)
            ids.append(int(rng.integers(1, 256)))
    edits = np.asarray(edits, np.uint32)
    ids = np.asarray(ids, np.uint8)
    call("vpt_world_set_blocks", w.handle, edits.ctypes.data_as(C.c_void_p), ids.ctypes.data_as(C.c_void_p),
         len(edits), None)
    expect = g.copy()
    for e, i in zip(edits, ids):          # sequential SetBlock
        expect[(e >> 20) & 1023, (e >> 10) & 1023, e & 1023] = i
    np.testing.assert_array_equal(w.download_grid(), expect)
    w.close()
