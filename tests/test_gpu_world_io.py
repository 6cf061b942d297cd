"""GPU tests of the raw world file format (SaveWorld / LoadWorld,
WorldFileHandler.cpp:10-83): a saved world reloads byte for byte and its
rebuilt distance field equals the oracle's; a file of the wrong size is
refused."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from tests.helpers import random_world, terrain
from voxelpathtracer_amd import World
from voxelpathtracer_amd._lib import VptError

pytestmark = pytest.mark.gpu


def test_save_load_round_trip(ctx, tmp_path):
    g = terrain(128, 64, 96)
    g[10:20, 40:50, 30:40] = 77      # a floating block of another id
    w = World(ctx, 128, 64, 96)
    w.m_WorldData[...] = g
    w.Buffer()
    path = tmp_path / "world.bin"
    w.SaveWorld(str(path))
    raw = np.fromfile(path, np.uint8)
    assert raw.size == 128 * 64 * 96
    np.testing.assert_array_equal(raw.reshape(96, 64, 128), g)   # x fastest, then y, then z
    w2 = World(ctx, 128, 64, 96)
    w2.LoadWorld(str(path))
    np.testing.assert_array_equal(w2.m_WorldData, g)
    w2.GenerateDistanceField()
    np.testing.assert_array_equal(w2.download_df(), oracle.df_build(g))
    w.close()
    w2.close()


def test_load_rejects_wrong_size(ctx, tmp_path):
    rng = np.random.default_rng(3)
    g = random_world(rng, 64, 32, 64, 0.05)
    path = tmp_path / "short.bin"
    g.reshape(-1)[:-1].tofile(path)
    w = World(ctx, 64, 32, 64)
    with pytest.raises(VptError):
        w.LoadWorld(str(path))
    with pytest.raises(VptError):
        w.LoadWorld(str(tmp_path / "missing.bin"))
    w.close()


def test_minecraft_import_then_df(ctx, tmp_path):
    """MCWorldImporter::ImportWorld then the DF build (Pipeline.cpp:1236,
    1293-1294): the imported grid equals the reference reader's (committed
    fixture) and the DF built on the GPU equals the oracle's of that grid."""
    from tests.golden import mca_synth as M
    from pathlib import Path
    z = np.load(Path(__file__).resolve().parent / "golden" / "mca_legacy.npz")
    d = tmp_path / "region"
    d.mkdir()
    for k in z.files:
        if k.startswith("file:"):
            (d / k[5:]).write_bytes(z[k].tobytes())
    nx, ny, nz = M.DIMS
    w = World(ctx, nx, ny, nz)
    print("minecraft import 48^3", flush=True)
    stats = w.ImportMinecraftWorld(str(d), M.ORIGIN, M.block_lut(), M.NAMES)
    assert stats[2] == int(z["sections"])
    np.testing.assert_array_equal(w.m_WorldData, z["expected"])
    w.GenerateDistanceField()
    np.testing.assert_array_equal(w.download_df(), oracle.df_build(z["expected"]))
    w.close()
