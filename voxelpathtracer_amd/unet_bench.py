"""Timing of the learned denoiser on one MI355X: `vpt_unet_denoise` on a
1920x1080 frame of the 384x128x384 terrain world (assembly + network), and
training steps at batch 16 x 128 x 128.  Weights are freshly initialised
(timing does not depend on them).

    python -m voxelpathtracer_amd.unet_bench [--reps N]
"""
from __future__ import annotations

import argparse
import json
import time


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=10)
    args = ap.parse_args()
    import torch
    from . import Context, World, generate_terrain, gi_params, make_camera
    from .unet import UNet
    from .world import DeviceGBuffer
    W, H = 1920, 1080
    ctx = Context(0)
    w = World(ctx, 384, 128, 384)
    w.m_WorldData[...] = generate_terrain(384, 128, 384)
    w.Buffer()
    w.GenerateDistanceField()
    cam, _, _ = make_camera((192.5, 100.5, 192.5), 45.0, -20.0, W, H)
    gb = DeviceGBuffer(W, H)
    rgb = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
    tex = torch.empty((H, W, 8), dtype=torch.int16, device="cuda")
    w.trace_primary(cam, gb, 350)
    w.trace_gi(cam, gb, gi_params(50.0, 0), rgb, tex)
    u = UNet(ctx, 1, H, W)
    u.init(1)
    out = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
    for _ in range(3):
        u.denoise(gb, tex, out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.reps):
        u.denoise(gb, tex, out)
    torch.cuda.synchronize()
    denoise_ms = (time.perf_counter() - t0) * 1e3 / args.reps
    x = torch.zeros((16, 128, 128, 32), dtype=torch.int16, device="cuda")
    t = torch.zeros((16, 128, 128, 3), dtype=torch.float32, device="cuda")
    for _ in range(3):
        u.backward(x, t, 16, 128, 128)
        u.adam_step(1e-3)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.reps):
        u.backward(x, t, 16, 128, 128)
        u.adam_step(1e-3)
    torch.cuda.synchronize()
    step_ms = (time.perf_counter() - t0) * 1e3 / args.reps
    print(json.dumps({"denoise_ms_1080p": round(denoise_ms, 4), "train_step_ms_16x128x128": round(step_ms, 4)}))
    u.close()
    w.close()


if __name__ == "__main__":
    main()
