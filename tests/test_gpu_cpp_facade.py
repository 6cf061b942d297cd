"""The C++ host façades (include/vpt/World.hpp, Pipeline.hpp) drive the same
passes as the Python host: the C++ example's denoised frame, denoised sun
shadow, reflection hit distances and colour and composite equal the Python
Pipeline's bit for bit,
with the C++ pipeline tracing through the persistent + graph engine, one
launch per pass, or one launch per pass replayed as a HIP graph each frame
(the Python pipeline: one launch per pass, eager)."""
from __future__ import annotations

import json
import subprocess

import numpy as np
import pytest

from voxelpathtracer_amd import Pipeline, World
from voxelpathtracer_amd.build import DEMO
from voxelpathtracer_amd.world import generate_terrain

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("path", ["persistent", "plain", "graph"])
def test_cpp_facade_matches_python_pipeline(ctx, tmp_path, path):
    import torch
    W, H, frames = 320, 180, 4
    dump = tmp_path / "rgb.bin"
    print(f"running the C++ façade example ({path})", flush=True)
    extra = {"plain": ["--plain"], "graph": ["--graph"]}.get(path, [])
    r = subprocess.run([str(DEMO), str(W), str(H), str(frames), *extra, "--dump", str(dump)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    assert info["path"] == path and info["frames"] == frames
    raw = np.fromfile(dump, np.uint8)
    n_rgb, n_sh = H * W * 3 * 4, H * W
    cpp = raw[:n_rgb].view(np.float32).reshape(H, W, 3)
    cpp_shadow = raw[n_rgb:n_rgb + n_sh].reshape(H, W)
    n_rd = (H // 4) * (W // 4) * 2
    cpp_refl = raw[n_rgb + n_sh:n_rgb + n_sh + n_rd].view(np.uint16).reshape(H // 4, W // 4)
    cpp_rcol = raw[n_rgb + n_sh + n_rd:n_rgb + n_sh + 5 * n_rd].view(np.uint16).reshape(H // 4, W // 4, 4)
    cpp_color = raw[n_rgb + n_sh + 5 * n_rd:].view(np.float32).reshape(H, W, 3)

    world = World(ctx, 384, 128, 384)
    world.m_WorldData[...] = generate_terrain(384, 128, 384)
    world.Buffer()
    world.GenerateDistanceField()
    pipe = Pipeline(ctx, world, W, H)
    for f in range(frames):
        pipe.render((192.5 + 0.1 * f, 100.5, 192.5), 45.0 + 0.2 * f, -20.0)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(cpp, pipe.denoised.cpu().numpy())
    np.testing.assert_array_equal(cpp_shadow, pipe.shadow_denoised.cpu().numpy())
    np.testing.assert_array_equal(cpp_refl, pipe.reflect_dist.cpu().numpy().view(np.uint16))
    np.testing.assert_array_equal(cpp_rcol, pipe.reflect_color.cpu().numpy().view(np.uint16))
    np.testing.assert_array_equal(cpp_color.view(np.uint32), pipe.color.cpu().numpy().view(np.uint32))
    assert (cpp_refl != 0xBC00).any()                    # some reflections hit (-1 is 0xBC00)
    # RaycastDetect through the façade agrees with the Python host
    assert tuple(info["pick"]) == world.RaycastDetect((192.5, 100.5, 192.5), (0.664463, -0.342020, 0.664463))
