"""GPU tests of the serving loop: every request's result is bit-identical
alone and in a mix of concurrent requests, in plain and batched mode, to
the one-launch kernels, and (for a small case) to the oracle; admission
respects the arena; large requests split over launches."""
from __future__ import annotations

import threading

import numpy as np
import pytest

import oracle
from tests.helpers import camera, random_world, terrain
from voxelpathtracer_amd import HIT_DTYPE, Server, World, gi_params
from voxelpathtracer_amd._lib import VPT_ENOMEM, VptError

pytestmark = pytest.mark.gpu

SUN = (0.3, 0.85, 0.42)


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


def _rays(rng, n, dims):
    nx, ny, nz = dims
    org = np.stack([rng.uniform(-4, nx + 4, n), rng.uniform(-2, ny + 2, n), rng.uniform(-4, nz + 4, n)], 1)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([org, d], 1).astype(np.float32)


def _dev(a):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).reshape(-1).copy()).cuda()


@pytest.fixture(scope="module")
def scene(ctx):
    rng = np.random.default_rng(77)
    grids = [terrain(128, 128, 128), random_world(rng, 256, 64, 128, 0.02), terrain(192, 96, 160)]
    worlds = [_world(ctx, g) for g in grids]
    reqs = []
    sizes = [1, 37, 1000, 5000, 70_000, 200_000, 3000, 150_000, 64, 90_000, 12_345, 400_000]
    for i, n in enumerate(sizes):
        wi = i % 3
        nz, ny, nx = grids[wi].shape
        kind = ("primary", "shadow", "gi", "gi")[i % 4]
        it = {"primary": 350, "shadow": 350, "gi": 48}[kind]
        rays = _rays(rng, n, (nx, ny, nz))
        reqs.append(dict(world=wi, kind=kind, n=n, max_iter=it, host=rays, dev=_dev(rays)))
    # reference: the one-launch kernels
    torch = _torch()
    for r in reqs:
        w = worlds[r["world"]]
        hits = torch.empty(r["n"] * 12, dtype=torch.uint8, device="cuda")
        if r["kind"] == "gi":
            vis = torch.empty(r["n"], dtype=torch.uint8, device="cuda")
            w.trace_gi_rays(r["dev"], hits, vis, r["n"], r["max_iter"], SUN, 128)
            torch.cuda.synchronize()
            r["ref_vis"] = vis.cpu().numpy()
        else:
            w.trace_rays(r["dev"], hits, r["n"], r["max_iter"])
            torch.cuda.synchronize()
        r["ref_hits"] = hits.cpu().numpy().view(HIT_DTYPE)
    yield grids, worlds, reqs
    for w in worlds:
        w.close()


def _check(r, out):
    got = out["hits"].cpu().numpy().view(HIT_DTYPE)
    for k in ("t", "voxel", "face", "block", "samples"):
        np.testing.assert_array_equal(got[k], r["ref_hits"][k], err_msg=f"{r['kind']} n={r['n']} {k}")
    if r["kind"] == "gi":
        np.testing.assert_array_equal(out["vis"].cpu().numpy(), r["ref_vis"])


def _submit(srv, slots, r):
    return srv.submit_rays(slots[r["world"]], r["kind"], r["dev"], r["n"], r["max_iter"], SUN, 128)


def test_gi_rays_match_oracle(ctx, scene):
    """vpt_trace_gi_rays against the oracle: bounce, continuation origin
    (hit + N * 0.045 toward the sun when N.L >= 0.001), shadow trace."""
    grids, worlds, reqs = scene
    r = reqs[3]   # gi, 5000 rays
    g = grids[r["world"]]
    df = oracle.df_build(g)
    b = oracle.trace_rays(g, df, r["host"], 48)
    np.testing.assert_array_equal(r["ref_hits"]["t"], b["t"])
    np.testing.assert_array_equal(r["ref_hits"]["face"], b["face"])
    L = np.asarray(SUN, np.float32)
    normals = np.array([[0, 0, 1], [0, 0, -1], [0, 1, 0], [0, -1, 0], [-1, 0, 0], [1, 0, 0]], np.float32)
    hit = (b["t"] > 0) & (b["block"] > 0)
    face = np.where(hit, b["face"], 0).astype(np.int64)
    hn = normals[face]
    ndotl = np.maximum((hn * L).sum(1, dtype=np.float32), np.float32(0))
    cont = hit & (ndotl >= np.float32(0.001))
    o, d, t = r["host"][:, :3], r["host"][:, 3:], b["t"][:, None]
    ip = o + d * t
    so = (ip + hn * np.float32(0.045)).astype(np.float32)
    expect = np.full(r["n"], 255, np.uint8)
    if cont.any():
        sh = oracle.trace_rays(g, df, np.concatenate([so[cont], np.broadcast_to(L, (cont.sum(), 3))], 1), 128)
        expect[cont] = (sh["t"] > 0).astype(np.uint8)
    assert cont.sum() > 200 and (expect == 0).sum() > 20 and (expect == 1).sum() > 5
    np.testing.assert_array_equal(r["ref_vis"], expect)


@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_alone_equals_mix(ctx, scene, mode):
    torch = _torch()
    grids, worlds, reqs = scene
    srv = Server(ctx, mode=mode, max_worlds=4, arena_bytes=256 << 20, max_launch_rays=65536, max_inflight=2)
    slots = [srv.add_world(w) for w in worlds]
    # alone: one request at a time
    for r in reqs:
        print(f"{mode} alone: {r['kind']} n={r['n']}", flush=True)
        t = _submit(srv, slots, r)
        _check(r, srv.fetch(t))
        srv.release(t)
    # mixed: all at once from two client threads, plus a frame request
    tickets = [None] * len(reqs)

    def client(idx):
        for i in idx:
            tickets[i] = _submit(srv, slots, reqs[i])

    th = [threading.Thread(target=client, args=(range(k, len(reqs), 2),)) for k in (0, 1)]
    W, H = 320, 180
    cam, *_ = camera(W, H, pos=(64.5, 100.5, 64.5))
    prm = gi_params(50.0, 3)
    for x in th:
        x.start()
    ft = srv.submit_frame(slots[0], cam, prm, 350, 350)
    for x in th:
        x.join()
    print(f"{mode} mix: {len(reqs)} requests + 1 frame", flush=True)
    for r, t in zip(reqs, tickets):
        out = srv.fetch(t)
        _check(r, out)
        assert out["t_arrival_ns"] <= out["t_admit_ns"] <= out["t_start_ns"] <= out["t_done_ns"]
        srv.release(t)
    fo = srv.fetch(ft)
    srv.release(ft)
    # the frame equals the one-launch passes
    from voxelpathtracer_amd.world import DeviceGBuffer
    gb = DeviceGBuffer(W, H)
    mask = torch.empty((H, W), dtype=torch.uint8, device="cuda")
    rgb = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
    worlds[0].trace_primary(cam, gb, 350)
    worlds[0].trace_shadow(cam, gb, list(prm.sun_dir), mask, 350)
    worlds[0].trace_gi(cam, gb, prm, rgb)
    torch.cuda.synchronize()
    g = gb.numpy()
    np.testing.assert_array_equal(fo["t"].cpu().numpy(), g["t"].reshape(-1))
    np.testing.assert_array_equal(fo["face"].cpu().numpy(), g["face"].reshape(-1))
    np.testing.assert_array_equal(fo["mask"].cpu().numpy(), mask.cpu().numpy().reshape(-1))
    np.testing.assert_array_equal(fo["rgb"].cpu().numpy(), rgb.cpu().numpy().reshape(-1))
    st = srv.stats()
    assert st["completed"] == st["submitted"] == 2 * len(reqs) + 1
    if mode == "batched":
        # 1.1M rays over launches of 64K: big requests were split, small ones shared launches
        assert st["launches"] < 2 * len(reqs) + 1 + 40
    srv.close()


def test_admission_waits_for_memory(ctx, scene):
    grids, worlds, reqs = scene
    r = reqs[11]                              # 400K GI rays: 13 bytes of results per ray
    assert r["kind"] == "gi" and r["n"] == 400_000
    per = r["n"] * 13 + 512
    srv = Server(ctx, mode="batched", max_worlds=4, arena_bytes=2 * per + 1024, max_launch_rays=1 << 20)
    slots = [srv.add_world(w) for w in worlds]
    # a request that can never fit is refused at submit
    with pytest.raises(VptError) as e:
        srv.submit_rays(slots[0], "primary", r["dev"], 3 * per, 350)
    assert e.value.code == VPT_ENOMEM
    tickets = [_submit(srv, slots, r) for _ in range(6)]
    done = []

    def collector():
        for t in tickets:
            done.append(srv.fetch(t))
            srv.release(t)

    c = threading.Thread(target=collector)
    c.start()
    c.join(timeout=120)
    assert not c.is_alive()
    for out in done:
        _check(r, out)
    st = srv.stats()
    assert st["arena_peak"] <= st["arena_bytes"]
    # two requests fit at once: the third was admitted only after the first
    # had completed and been released
    admits = sorted(o["t_admit_ns"] for o in done)
    dones = sorted(o["t_done_ns"] for o in done)
    assert admits[2] >= dones[0]
    srv.close()


def test_world_slots_and_errors(ctx, scene):
    grids, worlds, reqs = scene
    srv = Server(ctx, mode="batched", max_worlds=2, world_budget=1 << 40)
    s0 = srv.add_world(worlds[0])
    s1 = srv.add_world(worlds[1])
    with pytest.raises(VptError):
        srv.add_world(worlds[2])          # no free slot
    with pytest.raises(VptError):
        srv.submit_rays(7, "primary", reqs[0]["dev"], 1, 350)
    t = _submit(srv, [s0, s1, s0], reqs[1] | {"world": 1})
    srv.fetch(t)
    with pytest.raises(VptError):
        srv.remove_world(s1)              # the request is not released yet
    srv.release(t)
    # an empty request completes at once with no results
    t0 = srv.submit_rays(s1, "gi", reqs[0]["dev"], 0, 48, SUN, 128)
    out0 = srv.fetch(t0)
    assert out0["hits"].numel() == 0 and out0["vis"].numel() == 0
    srv.release(t0)
    srv.remove_world(s1)
    s2 = srv.add_world(worlds[2])
    assert s2 == s1
    srv.close()
    small = Server(ctx, mode="plain", max_worlds=2, world_budget=1 << 20)
    with pytest.raises(VptError):
        small.add_world(worlds[2])        # over the world budget
    small.close()
