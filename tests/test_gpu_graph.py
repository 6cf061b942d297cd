"""HIP-graph replay of a frame's calls (vpt_graph_*): the Python Pipeline
with use_graph, and the one-launch frame (vpt_trace_frame) captured
directly.  Every replay is compared with the same frame run eagerly, with
eager work and synchronisations of the program between the replays, and the
graph is instantiated once and then updated in place."""
from __future__ import annotations

import numpy as np
import pytest

from tests.helpers import camera, terrain
from voxelpathtracer_amd import Graph, Pipeline, World, gi_params
from voxelpathtracer_amd.world import DeviceGBuffer

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def world(ctx):
    w = World(ctx, 384, 128, 384)
    w.m_WorldData[...] = terrain(384, 128, 384)
    w.Buffer()
    w.GenerateDistanceField()
    yield w
    w.close()


def test_pipeline_graph_replays_equal_eager(ctx, world):
    import torch
    W, H, frames = 640, 360, 7
    eager = Pipeline(ctx, world, W, H)
    graphed = Pipeline(ctx, world, W, H, use_graph=True)
    gb = DeviceGBuffer(W, H)
    busy = torch.zeros(1 << 20, device="cuda")
    for f in range(frames):
        pos, yaw = (192.5 + 0.1 * f, 100.5, 192.5), 45.0 + 0.2 * f
        print(f"graph pipeline frame {f}", flush=True)
        eager.render(pos, yaw, -20.0)
        graphed.render(pos, yaw, -20.0)
        # the program's own eager work between replays
        cam, *_ = camera(W, H, pos=(100.5, 110.5, 100.5))
        world.trace_primary(cam, gb, 350)
        busy.add_(1.0)
        torch.cuda.synchronize()
        for name in ("denoised", "rgb", "texels", "shadow", "transversal", "shadow_denoised", "reflect_dist",
                     "reflect_hit", "reflect_color", "color"):
            a, b = getattr(eager, name), getattr(graphed, name)
            assert torch.equal(a, b), f"frame {f}: {name} differs between graph replay and eager run"
        ga, gb_ = eager.gbuf[f % 2].numpy(), graphed.gbuf[f % 2].numpy()
        for k in ("t", "face", "block", "voxel"):
            np.testing.assert_array_equal(ga[k], gb_[k], err_msg=f"frame {f}: G-buffer {k}")
    for chain, g in zip(("side", "main", "tail"), graphed.graphs):
        st = g.stats()
        assert st["launches"] == frames, (chain, st)
        assert st["instantiations"] == 1 and st["updates"] == frames - 1, (chain, st)
        assert st["nodes"] >= 2, (chain, st)
    assert float(busy[0]) == frames


def test_trace_frame_graph_replays_equal_eager(ctx, world):
    import torch
    W, H = 960, 540
    g = Graph(ctx)
    outs = {}
    for mode in ("eager", "graph"):
        gb = DeviceGBuffer(W, H)
        mask = torch.empty((H, W), dtype=torch.uint8, device="cuda")
        rgb = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
        tex = torch.empty((H, W, 8), dtype=torch.int16, device="cuda")
        res = []
        for f in range(5):
            cam, *_ = camera(W, H, pos=(192.5 + 0.3 * f, 100.5, 192.5), yaw=45.0 + 0.5 * f)
            prm = gi_params(50.0, f)
            if mode == "eager":
                world.trace_frame(cam, gb, prm, mask, rgb, tex, 350, 350)
            else:
                with g.capture() as s:
                    world.trace_frame(cam, gb, prm, mask, rgb, tex, 350, 350, stream=s)
            torch.cuda.synchronize()
            res.append((gb.numpy(), mask.cpu(), rgb.cpu(), tex.cpu()))
        outs[mode] = res
    for f, (a, b) in enumerate(zip(outs["eager"], outs["graph"])):
        for k in ("t", "face", "block", "voxel"):
            np.testing.assert_array_equal(a[0][k], b[0][k], err_msg=f"frame {f} {k}")
        assert torch.equal(a[1], b[1]) and torch.equal(a[2], b[2]) and torch.equal(a[3], b[3]), f"frame {f}"
    st = g.stats()
    assert st["launches"] == 5 and st["instantiations"] == 1 and st["updates"] == 4, st
    g.close()
