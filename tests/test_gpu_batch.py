"""GPU tests of the persistent paths: ray batches over several resident
worlds and the persistent frame chain (with and without HIP-graph replay),
each bit-identical to the one-launch kernels and to the oracle."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from tests.helpers import camera, random_world, svgf_camera_path, terrain
from voxelpathtracer_amd import HIT_DTYPE, Batch, Frame, World, gi_params, make_qrays
from voxelpathtracer_amd._lib import LaunchCfg
from voxelpathtracer_amd.world import DeviceGBuffer

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def _world(ctx, grid):
    nz, ny, nx = grid.shape
    w = World(ctx, nx, ny, nz)
    w.m_WorldData[...] = grid
    w.Buffer()
    w.GenerateDistanceField()
    return w


def _random_rays(rng, n, dims):
    nx, ny, nz = dims
    org = np.stack([rng.uniform(-10, nx + 10, n), rng.uniform(-5, ny + 5, n), rng.uniform(-10, nz + 10, n)], 1)
    d = rng.normal(size=(n, 3))
    d[: n // 50] = 0.0
    d[np.arange(n // 50), rng.integers(0, 3, n // 50)] = 1.0
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return org.astype(np.float32), d.astype(np.float32)


def _plain_hits(w, org, d, max_iter):
    torch = _torch()
    n = org.shape[0]
    rays = np.zeros((n, 6), np.float32)
    rays[:, :3], rays[:, 3:] = org, d
    dev = torch.from_numpy(rays.view(np.uint8).reshape(-1).copy()).cuda()
    hits = torch.empty(n * 12, dtype=torch.uint8, device="cuda")
    w.trace_rays(dev, hits, n, max_iter)
    torch.cuda.synchronize()
    return hits.cpu().numpy().view(HIT_DTYPE)


@pytest.fixture(scope="module")
def worlds(ctx):
    rng = np.random.default_rng(21)
    grids = [terrain(128, 128, 128), terrain(), random_world(rng, 256, 64, 128, 0.01)]
    ws = [_world(ctx, g) for g in grids]
    yield grids, ws
    for w in ws:
        w.close()


def test_batch_multi_world_parity(ctx, worlds):
    torch = _torch()
    grids, ws = worlds
    b = Batch(ctx, max_worlds=8)
    for slot, w in zip((0, 3, 7), ws):
        b.set_world(slot, w)
    rng = np.random.default_rng(5)
    parts = []
    for slot, g, w, n, it in zip((0, 3, 7), grids, ws, (60_000, 150_000, 40_000), (350, 48, 128)):
        nz, ny, nx = g.shape
        org, d = _random_rays(rng, n, (nx, ny, nz))
        parts.append((slot, w, org, d, it, _plain_hits(w, org, d, it)))
    org = np.concatenate([p[2] for p in parts])
    d = np.concatenate([p[3] for p in parts])
    it = np.concatenate([np.full(len(p[2]), p[4]) for p in parts])
    slot = np.concatenate([np.full(len(p[2]), p[0]) for p in parts])
    expect = np.concatenate([p[5] for p in parts])
    n = len(org)
    perm = rng.permutation(n)                 # interleave worlds in the queue
    q = make_qrays(org[perm], d[perm], it[perm], slot[perm], dst=perm)
    # extra rays: an empty slot, a slot past the table, a dst out of range
    extra = make_qrays(np.float32([[10, 10, 10]] * 3), np.float32([[0, -1, 0]] * 3), [350] * 3, [1, 200, 0],
                       dst=[n, n + 1, 10 ** 9])
    q = np.concatenate([q, extra])
    total = len(q)
    rays = torch.from_numpy(q.view(np.uint8).copy()).cuda()
    hits = torch.full((total * 12,), 0xAB, dtype=torch.uint8, device="cuda")
    print(f"batch parity: {n} rays over 3 worlds", flush=True)
    b.trace(rays, hits, total)
    torch.cuda.synchronize()
    got = hits.cpu().numpy().view(HIT_DTYPE)
    for k in ("t", "voxel", "face", "block", "samples"):
        np.testing.assert_array_equal(got[k][:n], expect[k], err_msg=k)
    # empty slot / unknown world: miss without a DF sample
    for i in (n, n + 1):
        assert got["t"][i] == -1.0 and got["face"][i] == 255 and got["samples"][i] == 0
    # the dropped result left the sentinel bytes alone beyond the two misses
    assert got.size == total
    # autotune picks a valid shape and the results stay identical
    ms = b.autotune(rays, hits, total)
    assert ms > 0
    cfg = b.cfg
    assert cfg.block in (64, 128, 256) and 1 <= cfg.waves_per_cu <= 64
    hits.fill_(0)
    b.trace(rays, hits, total)
    torch.cuda.synchronize()
    got2 = hits.cpu().numpy().view(HIT_DTYPE)
    for k in ("t", "voxel", "face", "block", "samples"):
        np.testing.assert_array_equal(got2[k][:n], expect[k], err_msg=k)
    b.close()


def test_batch_matches_oracle(ctx, worlds):
    torch = _torch()
    grids, ws = worlds
    g, w = grids[1], ws[1]
    df = oracle.df_build(g)
    b = Batch(ctx, max_worlds=1)
    b.set_world(0, w)
    rng = np.random.default_rng(8)
    org, d = _random_rays(rng, 50_000, (384, 128, 384))
    q = make_qrays(org, d, 350, 0)
    rays = torch.from_numpy(q.view(np.uint8).copy()).cuda()
    hits = torch.empty(len(q) * 12, dtype=torch.uint8, device="cuda")
    b.trace(rays, hits, len(q))
    torch.cuda.synchronize()
    got = hits.cpu().numpy().view(HIT_DTYPE)
    ref = oracle.trace_rays(g, df, np.concatenate([org, d], 1), 350)
    np.testing.assert_array_equal(got["t"], ref["t"])
    np.testing.assert_array_equal(got["face"], ref["face"])
    np.testing.assert_array_equal(got["samples"], ref["samples"])
    b.close()


def _plain_frame(w, cam, prm, W, H):
    torch = _torch()
    gb = DeviceGBuffer(W, H)
    mask = torch.empty((H, W), dtype=torch.uint8, device="cuda")
    rgb = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
    tex = torch.empty((H, W, 8), dtype=torch.int16, device="cuda")
    w.trace_primary(cam, gb, 350)
    w.trace_shadow(cam, gb, list(prm.sun_dir), mask, 350)
    w.trace_gi(cam, gb, prm, rgb, tex)
    torch.cuda.synchronize()
    return gb.numpy(), mask.cpu().numpy(), rgb.cpu().numpy(), tex.cpu().numpy()


@pytest.mark.parametrize("W,H", [(1920, 1080), (333, 97)])
def test_frame_persistent_parity(ctx, worlds, W, H):
    torch = _torch()
    grids, ws = worlds
    w = ws[1]
    fr = Frame(ctx, W, H)
    gb = DeviceGBuffer(W, H)
    mask = torch.empty((H, W), dtype=torch.uint8, device="cuda")
    rgb = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
    tex = torch.empty((H, W, 8), dtype=torch.int16, device="cuda")
    for f, use_graph in ((0, False), (1, True), (2, True), (3, True)):
        pos, yaw, pitch = svgf_camera_path(7 * f)
        cam, view, proj, iv, ip = camera(W, H, pos, yaw, pitch)
        prm = gi_params(50.0, f)
        print(f"frame parity {W}x{H} frame {f} graph={use_graph}", flush=True)
        fr.trace(w, cam, prm, gb, mask, rgb, tex, use_graph=use_graph)
        torch.cuda.synchronize()
        pg, pm, prgb, ptex = _plain_frame(w, cam, prm, W, H)
        host = gb.numpy()
        for k in ("t", "face", "block", "voxel"):
            np.testing.assert_array_equal(host[k], pg[k], err_msg=f"frame {f} {k}")
        np.testing.assert_array_equal(mask.cpu().numpy(), pm)
        np.testing.assert_array_equal(rgb.cpu().numpy(), prgb)
        np.testing.assert_array_equal(tex.cpu().numpy(), ptex)
    fr.close()


def test_frame_autotune_keeps_results(ctx, worlds):
    torch = _torch()
    grids, ws = worlds
    w = ws[1]
    W, H = 960, 540
    fr = Frame(ctx, W, H)
    gb = DeviceGBuffer(W, H)
    mask = torch.empty((H, W), dtype=torch.uint8, device="cuda")
    rgb = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
    tex = torch.empty((H, W, 8), dtype=torch.int16, device="cuda")
    cam, view, proj, iv, ip = camera(W, H)
    prm = gi_params(50.0, 4)
    ms = fr.autotune(w, cam, prm, gb, mask, rgb, tex)
    assert ms > 0
    fr.trace(w, cam, prm, gb, mask, rgb, tex, use_graph=True)
    torch.cuda.synchronize()
    pg, pm, prgb, ptex = _plain_frame(w, cam, prm, W, H)
    np.testing.assert_array_equal(gb.numpy()["t"], pg["t"])
    np.testing.assert_array_equal(mask.cpu().numpy(), pm)
    np.testing.assert_array_equal(rgb.cpu().numpy(), prgb)
    fr.close()


def test_batch_rejects_bad_arguments(ctx):
    from voxelpathtracer_amd._lib import VptError
    b = Batch(ctx, max_worlds=2)
    with pytest.raises(VptError):
        b.set_world(2, None)
    with pytest.raises(VptError):
        b.cfg = LaunchCfg(100, 8, 64, 32, 0)
    with pytest.raises(VptError):
        b.cfg = LaunchCfg(256, 8, 64, 0, 0)
    b.close()
