"""The served workload at its shape: 16 resident worlds from 128^3 to
512x256x512, ray-query requests of log-uniform size in [1K, 2M] rays with
kinds primary : shadow : GI = 1 : 1 : 2, submitted from two client threads
with spaced arrivals to the batched (continuously batching) server.  Every
result equals the one-launch kernels' bit for bit, and a sample of the
2M-ray GI request equals the oracle (bounce hit, and the sun visibility of
its continuation)."""
from __future__ import annotations

import threading
import time

import numpy as np
import pytest

import oracle
from tests.helpers import random_world
from voxelpathtracer_amd import HIT_DTYPE, Server, World, generate_terrain

pytestmark = pytest.mark.gpu

SUN = (0.3, 0.85, 0.42)
DIMS = [(128, 128, 128), (512, 256, 512), (192, 128, 192), (256, 128, 256), (384, 128, 384), (256, 256, 256),
        (320, 128, 320), (448, 128, 448), (128, 64, 512), (512, 128, 128), (256, 192, 256), (384, 256, 384),
        (160, 96, 160), (224, 128, 224), (288, 128, 288), (512, 256, 256)]


def _grid(k, dims):
    nx, ny, nz = dims
    if k % 4 == 3:
        return random_world(np.random.default_rng(100 + k), nx, ny, nz, 0.01)
    return generate_terrain(nx, ny, nz, seed=0x5EED0001 + k, seed_random=0x5EED0002 + k, h_min=max(4, ny // 8),
                            h_max=min(90, ny - 8))


def _rays(rng, n, dims, kind):
    nx, ny, nz = dims
    org = np.stack([rng.uniform(-4, nx + 4, n), rng.uniform(0, ny, n), rng.uniform(-4, nz + 4, n)], 1)
    if kind == "shadow":
        d = np.broadcast_to(np.asarray(SUN, np.float64) / np.linalg.norm(SUN), (n, 3))
    else:
        d = rng.normal(size=(n, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([org, d], 1).astype(np.float32)


@pytest.mark.timeout(600)
def test_serve_config_shape(ctx):
    import torch
    grids = [_grid(k, d) for k, d in enumerate(DIMS)]
    worlds = []
    for g in grids:
        nz, ny, nx = g.shape
        w = World(ctx, nx, ny, nz)
        w.m_WorldData[...] = g
        w.Buffer()
        w.GenerateDistanceField()
        worlds.append(w)
    rng = np.random.default_rng(2024)
    n_req = 40
    sizes = np.exp(rng.uniform(np.log(1000), np.log(2_000_000), n_req)).astype(np.int64)
    sizes[0] = 2_000_000                                     # the largest size of the workload
    kinds = [("gi", "primary", "gi", "shadow")[i % 4] for i in range(n_req)]
    kinds[0] = "gi"
    reqs = []
    for i in range(n_req):
        wi = int(rng.integers(0, len(worlds))) if i else 1    # the 2M-ray request on the 512x256x512 world
        kind, n = kinds[i], int(sizes[i])
        host = _rays(rng, n, DIMS[wi], kind)
        it = {"primary": 350, "shadow": 350, "gi": 48}[kind]
        dev = torch.from_numpy(host.view(np.uint8).reshape(-1).copy()).cuda()
        reqs.append(dict(world=wi, kind=kind, n=n, max_iter=it, host=host, dev=dev))
    print(f"serve config: {len(worlds)} worlds, {n_req} requests, {int(sizes.sum())} rays", flush=True)
    # reference: one launch per request
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
    budget = sum(w.device_bytes() for w in worlds) + (64 << 20)
    srv = Server(ctx, mode="batched", max_worlds=16, arena_bytes=1 << 30, max_launch_rays=1 << 20, max_inflight=2,
                 world_budget=budget)
    slots = [srv.add_world(w) for w in worlds]
    tickets = [None] * n_req

    def client(idx):
        for i in idx:
            r = reqs[i]
            tickets[i] = srv.submit_rays(slots[r["world"]], r["kind"], r["dev"], r["n"], r["max_iter"], SUN, 128)
            time.sleep(0.002)                                  # spaced arrivals

    th = [threading.Thread(target=client, args=(range(k, n_req, 2),)) for k in (0, 1)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    for i, (r, t) in enumerate(zip(reqs, tickets)):
        out = srv.fetch(t)
        got = out["hits"].cpu().numpy().view(HIT_DTYPE)
        for k in ("t", "voxel", "face", "block", "samples"):
            np.testing.assert_array_equal(got[k], r["ref_hits"][k], err_msg=f"request {i} {r['kind']} n={r['n']} {k}")
        if r["kind"] == "gi":
            np.testing.assert_array_equal(out["vis"].cpu().numpy(), r["ref_vis"], err_msg=f"request {i} vis")
        assert out["t_arrival_ns"] <= out["t_admit_ns"] <= out["t_start_ns"] <= out["t_done_ns"]
        srv.release(t)
    st = srv.stats()
    assert st["completed"] == n_req and st["rays"] == int(sizes.sum())
    assert st["launches"] > 5                                 # the 2M-ray request alone spans two launches
    srv.close()
    # a sample of the 2M-ray GI request against the oracle
    r = reqs[0]
    g = grids[r["world"]]
    df = oracle.df_build(g)
    idx = np.random.default_rng(5).choice(r["n"], 20_000, replace=False)
    sample = np.ascontiguousarray(r["host"][idx])
    b = oracle.trace_rays(g, df, sample, 48)
    ref = r["ref_hits"][idx]
    for k in ("t", "face", "block", "samples"):
        np.testing.assert_array_equal(ref[k], b[k], err_msg=f"oracle sample {k}")
    L = np.asarray(SUN, np.float32)
    normals = np.array([[0, 0, 1], [0, 0, -1], [0, 1, 0], [0, -1, 0], [-1, 0, 0], [1, 0, 0]], np.float32)
    hit = (b["t"] > 0) & (b["block"] > 0)
    face = np.where(hit, b["face"], 0).astype(np.int64)
    hn = normals[face]
    ndotl = np.maximum((hn * L).sum(1, dtype=np.float32), np.float32(0))
    cont = hit & (ndotl >= np.float32(0.001))
    o, d, t = sample[:, :3], sample[:, 3:], b["t"][:, None]
    so = ((o + d * t) + hn * np.float32(0.045)).astype(np.float32)
    expect = np.full(sample.shape[0], 255, np.uint8)
    sh = oracle.trace_rays(g, df, np.concatenate([so[cont], np.broadcast_to(L, (int(cont.sum()), 3))], 1), 128)
    expect[cont] = (sh["t"] > 0).astype(np.uint8)
    assert cont.sum() > 200
    np.testing.assert_array_equal(r["ref_vis"][idx], expect)
    for w in worlds:
        w.close()
