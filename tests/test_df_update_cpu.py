"""CPU checks of the claim the incremental DF update rests on, against the
oracle's full rebuild (GenerateDistanceField, World.cpp:69-113): after a set
of block edits, every cell whose distance changes lies inside a box around
the edits whose every face cell q has an old distance D_old(q) below its L1
distance to the edits' bounding box.  Random and terrain worlds, additions,
removals and mixed batches."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from tests.helpers import random_world, terrain


def _grown_box(df, lo, hi):
    """Per-face growth: a face of the box is final once every cell q on it
    has D_old(q) < L1 distance from q to the edits' bounding box; a face
    that fails moves out by max(v + 1, its margin), v the largest
    D_old(q) - dist(q) on it.  Faces on the world's border are final."""
    nz, ny, nx = df.shape
    dims = (nx, ny, nz)
    lo, hi = np.asarray(lo), np.asarray(hi)
    m = np.zeros((3, 2), np.int64)               # margin below / above per axis
    while True:
        b0 = np.maximum(0, lo - m[:, 0])
        b1 = np.minimum(dims, hi + m[:, 1] + 1)
        box = df[b0[2]:b1[2], b0[1]:b1[1], b0[0]:b1[0]].astype(np.int64)
        z, y, x = np.meshgrid(np.arange(b0[2], b1[2]), np.arange(b0[1], b1[1]), np.arange(b0[0], b1[0]),
                              indexing="ij")
        dist = sum(np.maximum(0, np.maximum(lo[a] - c, c - hi[a])) for a, c in enumerate((x, y, z)))
        slack = box - dist
        done = True
        for a in range(3):
            for side in range(2):
                if (b0[a] == 0) if side == 0 else (b1[a] == dims[a]):
                    continue
                face = np.take(slack, 0 if side == 0 else -1, axis=2 - a)
                v = int(face.max())
                if v >= 0:
                    m[a, side] = max(m[a, side] + v + 1, 2 * m[a, side])
                    done = False
        if done:
            return list(b0), list(b1)


def _edit(rng, g, k, mode, spread):
    nz, ny, nx = g.shape
    c = np.array([rng.integers(0, nx), rng.integers(0, ny), rng.integers(0, nz)])
    cells = np.clip(c + rng.integers(-spread, spread + 1, size=(k, 3)), 0, [nx - 1, ny - 1, nz - 1])
    g2 = g.copy()
    for x, y, z in cells:
        if mode == "add" or (mode == "mix" and rng.random() < 0.5):
            g2[z, y, x] = 7
        else:
            g2[z, y, x] = 0
    return g2, cells


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("mode", ["add", "remove", "mix"])
def test_changed_cells_inside_grown_box(seed, mode):
    rng = np.random.default_rng(100 + seed)
    g = terrain(64, 48, 80, seed=0x5EED0001 + seed) if seed % 2 else random_world(rng, 64, 48, 80, 0.02)
    df = oracle.df_build(g)
    for k, spread in [(1, 0), (6, 3), (20, 10)]:
        g2, cells = _edit(rng, g, k, mode, spread)
        df2 = oracle.df_build(g2)
        b0, b1 = _grown_box(df, cells.min(0), cells.max(0))
        changed = np.argwhere(df2 != df)          # (z, y, x)
        if changed.size:
            assert (changed[:, ::-1] >= b0).all() and (changed[:, ::-1] < b1).all(), (mode, k, b0, b1)


def test_known_shift_of_a_single_block():
    """One block in an empty world: removing it sends every cell to the cap,
    adding it back restores min(cap, L1)."""
    g = np.zeros((40, 24, 32), np.uint8)
    g[10, 5, 7] = 3
    df = oracle.df_build(g)
    z, y, x = np.meshgrid(np.arange(40), np.arange(24), np.arange(32), indexing="ij")
    l1 = np.abs(x - 7) + np.abs(y - 5) + np.abs(z - 10)
    np.testing.assert_array_equal(df, np.minimum(l1, min(254, 32 + 24 + 40)))
    assert (oracle.df_build(np.zeros_like(g)) == 96).all()
