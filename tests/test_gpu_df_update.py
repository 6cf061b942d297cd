"""GPU parity of the incremental DF update (vpt_df_update) with the oracle's
full rebuild (GenerateDistanceField, World.cpp:69-113) after block edits
(World::SetBlock, World.cpp:367-375): single and batched additions and
removals on random and terrain worlds, edits on the world's faces and
corners, the last block of a world removed, and a chain of updates applied
one after another.  Bit-exact."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from tests.helpers import random_world, terrain
from voxelpathtracer_amd import World
from voxelpathtracer_amd._lib import VptError

pytestmark = pytest.mark.gpu


def _world(ctx, g):
    nz, ny, nx = g.shape
    w = World(ctx, nx, ny, nz)
    w.m_WorldData[...] = g
    w.Buffer()
    w.GenerateDistanceField()
    return w


def _apply(w, rng, k, spread, mode):
    nz, ny, nx = w.m_WorldData.shape
    c = np.array([rng.integers(0, nx), rng.integers(0, ny), rng.integers(0, nz)])
    cells = np.clip(c + rng.integers(-spread, spread + 1, size=(k, 3)), 0, [nx - 1, ny - 1, nz - 1])
    for x, y, z in cells:
        add = mode == "add" or (mode == "mix" and rng.random() < 0.5)
        w.SetBlock(int(x), int(y), int(z), int(rng.integers(1, 128)) if add else 0)


@pytest.mark.parametrize("kind", ["random", "terrain"])
@pytest.mark.parametrize("mode", ["add", "remove", "mix"])
def test_update_matches_rebuild(ctx, kind, mode):
    rng = np.random.default_rng({"add": 1, "remove": 2, "mix": 3}[mode] + (10 if kind == "terrain" else 0))
    g = terrain(256, 96, 192) if kind == "terrain" else random_world(rng, 256, 96, 192, 0.03)
    w = _world(ctx, g)
    print(f"df update {kind} {mode}", flush=True)
    for step, (k, spread) in enumerate([(1, 0), (1, 0), (8, 2), (40, 12), (200, 40)]):
        _apply(w, rng, k, spread, mode)
        region = w.UpdateDistanceField(region=(step % 2 == 0))
        assert (region is not None) == (step % 2 == 0)
        np.testing.assert_array_equal(w.download_df(), oracle.df_build(w.m_WorldData), err_msg=f"step {step} {region}")
    w.close()


def test_single_edit_region_is_local(ctx):
    """On a terrain world a block placed on the surface recomputes a small
    box, not the world."""
    g = terrain(384, 128, 384)
    w = _world(ctx, g)
    y = int(np.argmin(g[100, :, 200] > 0))    # first empty cell of the column (x 200, z 100)
    assert g[100, y, 200] == 0
    w.SetBlock(200, y, 100, 5)
    x0, y0, z0, x1, y1, z1 = w.UpdateDistanceField(region=True)
    vol = (x1 - x0) * (y1 - y0) * (z1 - z0)
    assert vol < 384 * 128 * 384 // 1000, (x0, y0, z0, x1, y1, z1)
    np.testing.assert_array_equal(w.download_df(), oracle.df_build(w.m_WorldData))
    w.close()


def test_faces_corners_and_last_block(ctx):
    g = np.zeros((64, 32, 48), np.uint8)
    w = _world(ctx, g)
    corners = [(0, 0, 0), (47, 31, 63), (0, 31, 0), (47, 0, 63), (23, 0, 31), (0, 15, 40)]
    for x, y, z in corners:
        w.SetBlock(x, y, z, 9)
        w.UpdateDistanceField()
        np.testing.assert_array_equal(w.download_df(), oracle.df_build(w.m_WorldData), err_msg=str((x, y, z)))
    for x, y, z in corners:
        w.SetBlock(x, y, z, 0)
        w.UpdateDistanceField()
        np.testing.assert_array_equal(w.download_df(), oracle.df_build(w.m_WorldData), err_msg=str((x, y, z)))
    assert (w.download_df() == min(254, 48 + 32 + 64)).all()
    w.close()


def test_no_edits_and_bad_edits(ctx):
    rng = np.random.default_rng(4)
    g = random_world(rng, 64, 32, 64, 0.05)
    w = _world(ctx, g)
    assert w.UpdateDistanceField(region=True) is None
    import ctypes as C
    from voxelpathtracer_amd._lib import call
    bad = np.array([64 | (1 << 10)], np.uint32)        # x = 64 outside a 64-wide world
    with pytest.raises(VptError):
        call("vpt_df_update", w.handle, bad.ctypes.data_as(C.c_void_p), 1, None, C.c_void_p(0))
    np.testing.assert_array_equal(w.download_df(), oracle.df_build(g))
    w.close()


def test_scattered_edits_take_the_rebuild(ctx):
    """Edits at opposite corners: the box spans the world, so the update
    runs the full rebuild behind its gate."""
    rng = np.random.default_rng(8)
    g = random_world(rng, 128, 64, 128, 0.02)
    w = _world(ctx, g)
    w.SetBlock(0, 0, 0, 1)
    w.SetBlock(127, 63, 127, 0 if g[127, 63, 127] else 1)
    assert w.UpdateDistanceField(region=True) == (0, 0, 0, 128, 64, 128)
    np.testing.assert_array_equal(w.download_df(), oracle.df_build(w.m_WorldData))
    w.close()


def test_large_world_updates(ctx):
    """1024 x 256 x 128 terrain (x-y slabs too large for the LDS pass of the
    rebuild): a run of local single-block edits, checked at the end."""
    g = terrain(1024, 256, 128)
    w = _world(ctx, g)
    rng = np.random.default_rng(12)
    print("df update 1024x256x128", flush=True)
    for k in range(6):
        x, z = int(rng.integers(0, 1024)), int(rng.integers(0, 128))
        y = int(np.argmin(w.m_WorldData[z, :, x] > 0))
        w.SetBlock(x, y, z, 0 if k % 3 == 2 else 4)
        if k % 3 == 2:
            w.SetBlock(x, max(0, y - 1), z, 0)     # dig into the surface
        w.UpdateDistanceField()
    np.testing.assert_array_equal(w.download_df(), oracle.df_build(w.m_WorldData))
    w.close()
