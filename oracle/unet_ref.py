"""CPU fp32 reference of the learned denoiser's network (test infrastructure).

The same U-Net as libvpt.so's vpt_unet, written with torch CPU operators in
float32: conv weights rounded to bf16 as the GPU reads them, every stored
activation rounded to bf16 (round to nearest even) in the forward pass with
a straight-through gradient, fp32 biases and head.  Its autograd gradients
are the reference for the GPU backward kernels; the GPU additionally rounds
the stored activation gradients to bf16, so the comparison is to a stated
tolerance.  Only tests import this module.
"""
from __future__ import annotations

import numpy as np

LAYERS = [("x", None, 0, "e0a", 32, 32), ("e0a", None, 0, "e0b", 32, 32), ("p0", None, 0, "e1a", 32, 64),
          ("e1a", None, 0, "e1b", 64, 64), ("p1", None, 0, "b0", 64, 128), ("b0", None, 0, "b1", 128, 128),
          ("b1", "e1b", 1, "d1a", 192, 64), ("d1a", None, 0, "d1b", 64, 64), ("d1b", "e0b", 1, "d0a", 96, 32),
          ("d0a", None, 0, "d0b", 32, 32)]
POOLED = {"e0b": "p0", "e1b": "p1"}


def _torch():
    import torch
    return torch


def bf16_round(t):
    torch = _torch()
    return t.to(torch.bfloat16).to(torch.float32)


def _ste(t):
    """bf16 rounding in the forward pass, identity in the backward pass."""
    return t + (bf16_round(t) - t).detach()


def unpack(params: np.ndarray):
    """Flat parameter vector -> {name: float32 tensor requiring grad}."""
    torch = _torch()
    out, off = {}, 0
    for _, _, _, name, cin, cout in LAYERS:
        n = 9 * cout * cin
        out[name + ".w"] = torch.tensor(params[off:off + n].reshape(9, cout, cin), requires_grad=True)
        off += n
        out[name + ".b"] = torch.tensor(params[off:off + cout], requires_grad=True)
        off += cout
    out["head.w"] = torch.tensor(params[off:off + 96].reshape(3, 32), requires_grad=True)
    off += 96
    out["head.b"] = torch.tensor(params[off:off + 3], requires_grad=True)
    off += 3
    assert off == params.size
    return out


def pack_grads(p: dict) -> np.ndarray:
    parts = []
    for _, _, _, name, _, _ in LAYERS:
        parts += [p[name + ".w"].grad.numpy().reshape(-1), p[name + ".b"].grad.numpy().reshape(-1)]
    parts += [p["head.w"].grad.numpy().reshape(-1), p["head.b"].grad.numpy().reshape(-1)]
    return np.concatenate(parts)


def forward(p: dict, x_nhwc):
    """x_nhwc: float32 (n, h, w, 32).  Returns (log-space prediction (n, h,
    w, 3), {activation name: NHWC tensor})."""
    torch = _torch()
    F = torch.nn.functional
    acts = {"x": torch.as_tensor(x_nhwc).permute(0, 3, 1, 2)}
    for a, b, up, name, cin, cout in LAYERS:
        src = acts[a]
        if up:
            src = F.interpolate(src, scale_factor=2, mode="nearest")
        if b is not None:
            src = torch.cat([src, acts[b]], 1)
        w = bf16_round(p[name + ".w"])                         # [9][cout][cin]
        W = w.permute(1, 2, 0).reshape(cout, cin, 3, 3)        # tap = ky * 3 + kx
        y = F.conv2d(src, W, padding=1) + p[name + ".b"][None, :, None, None]
        acts[name] = _ste(F.relu(y))
        if name in POOLED:
            acts[POOLED[name]] = _ste(F.avg_pool2d(acts[name], 2))
    d0b = acts["d0b"]
    out = torch.einsum("nchw,jc->njhw", d0b, p["head.w"]) + p["head.b"][None, :, None, None]
    return out.permute(0, 2, 3, 1), {k: v.permute(0, 2, 3, 1) for k, v in acts.items()}


def loss_and_grads(params: np.ndarray, x_nhwc, target_nhwc):
    """L1 loss (mean) and the flat gradient vector."""
    torch = _torch()
    p = unpack(params)
    out, _ = forward(p, x_nhwc)
    loss = (out - torch.as_tensor(target_nhwc)).abs().mean()
    loss.backward()
    return float(loss), pack_grads(p)
