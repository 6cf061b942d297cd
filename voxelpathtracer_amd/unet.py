"""Learned denoiser (vpt_unet in include/vpt/vpt.h): wrapper, training-pair
generation with the tracer itself, training loop and evaluation against
256-spp references.

Pairs: a random terrain world and camera, a 128x128 window of a 1920x1080
60-degree frustum (so features have their full-resolution pixel size), the
primary G-buffer, one 1-spp GI frame as the input and the irradiance decoded
from 256 accumulated GI frames as the target (log1p space).  The network,
its loss, gradients and optimiser all run in libvpt.so.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from ._lib import Camera, call
from .world import Context, DeviceGBuffer, World, _ptr, _stream, generate_terrain, gi_params, make_camera

ACT_NAMES = ["x", "e0a", "e0b", "p0", "e1a", "e1b", "p1", "b0", "b1", "d1a", "d1b", "d0a", "d0b"]
ACT_CHANNELS = [32, 32, 32, 32, 64, 64, 64, 128, 128, 64, 64, 32, 32]
ACT_LEVEL = [0, 0, 0, 1, 1, 1, 2, 2, 2, 1, 1, 0, 0]
# (input a, input b, a upsampled, output, cin, cout) per conv layer
LAYERS = [("x", None, 0, "e0a", 32, 32), ("e0a", None, 0, "e0b", 32, 32), ("p0", None, 0, "e1a", 32, 64),
          ("e1a", None, 0, "e1b", 64, 64), ("p1", None, 0, "b0", 64, 128), ("b0", None, 0, "b1", 128, 128),
          ("b1", "e1b", 1, "d1a", 192, 64), ("d1a", None, 0, "d1b", 64, 64), ("d1b", "e0b", 1, "d0a", 96, 32),
          ("d0a", None, 0, "d0b", 32, 32)]


def param_layout():
    """[(name, offset, shape)] of the flat parameter vector."""
    out, off = [], 0
    for i, (_, _, _, name, cin, cout) in enumerate(LAYERS):
        out.append((f"{name}.w", off, (9, cout, cin)))
        off += 9 * cout * cin
        out.append((f"{name}.b", off, (cout,)))
        off += cout
    out.append(("head.w", off, (3, 32)))
    off += 96
    out.append(("head.b", off, (3,)))
    off += 3
    return out, off


class UNet:
    def __init__(self, ctx: Context, max_n: int, max_h: int, max_w: int):
        h = C.c_void_p()
        call("vpt_unet_create", ctx.handle, max_n, max_h, max_w, C.byref(h))
        self.handle = h
        self.ctx = ctx
        n = C.c_size_t()
        call("vpt_unet_param_count", self.handle, C.byref(n))
        self.n_params = n.value
        assert self.n_params == param_layout()[1]

    def init(self, seed: int = 1) -> None:
        call("vpt_unet_init", self.handle, seed)

    def get_params(self) -> np.ndarray:
        a = np.empty(self.n_params, np.float32)
        call("vpt_unet_get_params", self.handle, a.ctypes.data_as(C.c_void_p), self.n_params)
        return a

    def set_params(self, a) -> None:
        a = np.ascontiguousarray(a, np.float32)
        call("vpt_unet_set_params", self.handle, a.ctypes.data_as(C.c_void_p), a.size)

    def get_grads(self) -> np.ndarray:
        a = np.empty(self.n_params, np.float32)
        call("vpt_unet_get_grads", self.handle, a.ctypes.data_as(C.c_void_p), self.n_params)
        return a

    def forward(self, x, n: int, h: int, w: int, out, mode: int = 0, stream=None) -> None:
        """x: cuda int16 tensor (n, h, w, 32) of bf16 bits; out: cuda float32 (n, h, w, 3)."""
        call("vpt_unet_forward", self.handle, _ptr(x), n, h, w, mode, _ptr(out), _stream(stream))

    def backward(self, x, target, n: int, h: int, w: int, stream=None) -> float:
        loss = C.c_float()
        call("vpt_unet_backward", self.handle, _ptr(x), _ptr(target), n, h, w, C.byref(loss), _stream(stream))
        return loss.value

    def adam_step(self, lr: float, stream=None) -> None:
        call("vpt_unet_adam_step", self.handle, C.c_float(lr), _stream(stream))

    def assemble(self, gb: DeviceGBuffer, tex, x, stream=None) -> None:
        call("vpt_unet_assemble", self.handle, C.byref(gb.struct()), _ptr(tex), _ptr(x), _stream(stream))

    def denoise(self, gb: DeviceGBuffer, tex, rgb, stream=None) -> None:
        call("vpt_unet_denoise", self.handle, C.byref(gb.struct()), _ptr(tex), _ptr(rgb), _stream(stream))

    def activation(self, index: int, n: int, h: int, w: int) -> np.ndarray:
        lv, ch = ACT_LEVEL[index], ACT_CHANNELS[index]
        a = np.empty((n, h >> lv, w >> lv, ch), np.uint16)
        call("vpt_unet_read_activation", self.handle, index, n, h, w, a.ctypes.data_as(C.c_void_p))
        return a

    def close(self) -> None:
        if self.handle:
            call("vpt_unet_destroy", self.handle)
            self.handle = None


def texel_accumulate(tex, npx: int, acc, reset: bool, stream=None) -> None:
    call("vpt_texel_accumulate", _ptr(tex), npx, _ptr(acc), int(reset), _stream(stream))


def irradiance_decode(gb: DeviceGBuffer, rgb, tex=None, acc=None, scale: float = 1.0, stream=None) -> None:
    call("vpt_irradiance_decode", C.byref(gb.struct()), _ptr(tex), _ptr(acc), C.c_float(scale), _ptr(rgb),
         _stream(stream))


def bf16_bits(a: np.ndarray) -> np.ndarray:
    """float32 -> bf16 bit patterns (round to nearest even)."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    return r


def bf16_to_f32(b: np.ndarray) -> np.ndarray:
    return (np.asarray(b, np.uint16).astype(np.uint32) << 16).view(np.float32)


def tonemap_psnr(a, ref) -> float:
    """PSNR of x / (1 + x) tonemapped irradiance (peak 1)."""
    a = np.asarray(a, np.float64)
    ref = np.asarray(ref, np.float64)
    ta, tr = a / (1.0 + a), ref / (1.0 + ref)
    mse = float(np.mean((ta - tr) ** 2))
    return float("inf") if mse == 0 else -10.0 * math.log10(mse)


def crop_camera(pos, yaw, pitch, ox: int, oy: int, w: int, h: int, full_w: int = 1920, full_h: int = 1080) -> Camera:
    """Camera of the w x h window at pixel (ox, oy) of a full_w x full_h
    60-degree camera: inv_proj = inv_proj_full * M, with M mapping the
    window's clip x/y onto the full frame's."""
    cam, view, proj = make_camera(pos, yaw, pitch, full_w, full_h)
    ip = np.array(cam.inv_proj, np.float64).reshape(4, 4).T      # row-major
    M = np.eye(4)
    M[0, 0] = w / full_w
    M[0, 3] = (2.0 * ox + w) / full_w - 1.0
    M[1, 1] = h / full_h
    M[1, 3] = (2.0 * oy + h) / full_h - 1.0
    ipc = (ip @ M).astype(np.float32)
    c = Camera()
    c.inv_view[:] = list(cam.inv_view)
    c.inv_proj[:] = [float(v) for v in ipc.T.reshape(-1)]
    c.width, c.height = w, h
    return c


class PairMaker:
    """Renders (1-spp input, 256-spp target) pairs with the tracer."""

    def __init__(self, ctx: Context, worlds: int = 4, size: int = 128, spp: int = 256, seed: int = 7,
                 dims=(256, 128, 256)):
        import torch
        self.ctx, self.size, self.spp = ctx, size, spp
        self.rng = np.random.default_rng(seed)
        self.dims = dims
        self.worlds = []
        for k in range(worlds):
            nx, ny, nz = dims
            w = World(ctx, nx, ny, nz)
            w.m_WorldData[...] = generate_terrain(nx, ny, nz, seed=0x5EED4000 + k + 97 * seed,
                                                  seed_random=0x5EED5000 + k, h_min=ny * 20 // 128,
                                                  h_max=ny * 90 // 128)
            w.Buffer()
            w.GenerateDistanceField()
            self.worlds.append(w)
        S = size
        self.gb = DeviceGBuffer(S, S)
        self.rgb = torch.empty((S, S, 3), dtype=torch.float32, device="cuda")
        self.tex = torch.empty((S, S, 8), dtype=torch.int16, device="cuda")
        self.acc = torch.empty((S, S, 6), dtype=torch.float32, device="cuda")
        self.ref = torch.empty((S, S, 3), dtype=torch.float32, device="cuda")

    def close(self):
        for w in self.worlds:
            w.close()

    def pair(self, unet: UNet, x_out, t_out, frame_base: int = 0):
        """Writes one pair: x_out (S, S, 32) int16 bf16 bits, t_out (S, S, 3)
        float32 log1p(target irradiance)."""
        import torch
        S, rng = self.size, self.rng
        w = self.worlds[int(rng.integers(len(self.worlds)))]
        nx, ny, nz = self.dims
        pos = (float(rng.uniform(0.15, 0.85) * nx), float(ny * rng.uniform(0.7, 0.95)),
               float(rng.uniform(0.15, 0.85) * nz))
        yaw, pitch = float(rng.uniform(0, 360)), float(rng.uniform(-45, -10))
        ox, oy = int(rng.integers(0, 1920 - S)), int(rng.integers(0, 1080 - S))
        cam = crop_camera(pos, yaw, pitch, ox, oy, S, S)
        tick = float(rng.uniform(35.0, 65.0))
        w.trace_primary(cam, self.gb, 350)
        base = int(rng.integers(1 << 20)) * 1024 + frame_base
        for f in range(self.spp):
            prm = gi_params(tick, base + f)
            w.trace_gi(cam, self.gb, prm, self.rgb, self.tex)
            texel_accumulate(self.tex, S * S, self.acc, f == 0)
        irradiance_decode(self.gb, self.ref, acc=self.acc, scale=1.0 / self.spp)
        prm = gi_params(tick, base + self.spp)   # an independent 1-spp frame as the input
        w.trace_gi(cam, self.gb, prm, self.rgb, self.tex)
        unet.assemble(self.gb, self.tex, x_out)
        t_out.copy_(torch.log1p(self.ref))


def make_dataset(ctx: Context, unet: UNet, count: int, size: int = 128, spp: int = 256, seed: int = 7, worlds: int = 4):
    """(x, t): cuda tensors (count, S, S, 32) int16 and (count, S, S, 3) float32."""
    import torch
    pm = PairMaker(ctx, worlds=worlds, size=size, spp=spp, seed=seed)
    x = torch.empty((count, size, size, 32), dtype=torch.int16, device="cuda")
    t = torch.empty((count, size, size, 3), dtype=torch.float32, device="cuda")
    for i in range(count):
        pm.pair(unet, x[i], t[i])
    torch.cuda.synchronize()
    pm.close()
    return x, t


def train(unet: UNet, x, t, steps: int, batch: int = 16, lr: float = 2e-3, seed: int = 3, log_every: int = 0):
    """Adam on random minibatches (with random flips); returns the loss curve."""
    import torch
    g = torch.Generator(device="cpu").manual_seed(seed)
    n, S = x.shape[0], x.shape[1]
    xb = torch.empty((batch, S, S, 32), dtype=torch.int16, device="cuda")
    tb = torch.empty((batch, S, S, 3), dtype=torch.float32, device="cuda")
    losses = []
    for step in range(steps):
        idx = torch.randint(0, n, (batch,), generator=g).cuda()
        xs, ts = x[idx], t[idx]
        if step % 2:
            xs, ts = xs.flip(2), ts.flip(2)   # horizontal flip
        xb.copy_(xs)
        tb.copy_(ts)
        # cosine learning-rate decay
        cur = lr * 0.5 * (1.0 + math.cos(math.pi * step / max(steps, 1)))
        loss = unet.backward(xb, tb, batch, S, S)
        unet.adam_step(max(cur, lr * 0.02))
        losses.append(loss)
        if log_every and step % log_every == 0:
            print(f"step {step} loss {loss:.5f}", flush=True)
    return losses
