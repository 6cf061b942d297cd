"""Frame orchestration of the hot path — the slice of the reference's
``MainPipeline::StartPipeline`` frame loop that this library replaces
(Core/Pipeline.cpp:1960-2852): primary trace into a ping-pong G-buffer
(:1963-2016), diffuse GI trace (:2171-2281), the SVGF chain (:2284-2615)
and the sun-shadow trace (:2790-2852).  Previous-frame matrices follow
:1816-1824.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import SvgfParams, call, lib
from .world import Context, DeviceGBuffer, World, _ptr, _stream, gi_params, make_camera


def svgf_params_default() -> SvgfParams:
    p = SvgfParams()
    call("vpt_svgf_params_default", C.byref(p))
    return p


def decode_texels(tex: np.ndarray) -> dict:
    """(H, W, 8) uint16 texels -> float planes (the storage values)."""
    tex = np.asarray(tex, np.uint16)
    h = tex[..., :7].view(np.float16).astype(np.float32)
    ao = (tex[..., 7] & 0xFF).astype(np.float32) / np.float32(255.0)
    sky = (tex[..., 7] >> 8).astype(np.float32) / np.float32(255.0)
    return dict(sh=h[..., 0:4], cocg=h[..., 4:6], x=h[..., 6], ao=np.stack([ao, sky], -1))


class SVGF:
    """vpt_svgf: history and ping-pong planes of one view."""

    def __init__(self, ctx: Context, width: int, height: int, params: SvgfParams | None = None):
        h = C.c_void_p()
        self.params = params or svgf_params_default()
        call("vpt_svgf_create", ctx.handle, width, height, C.byref(self.params), C.byref(h))
        self.handle = h
        self.width, self.height = width, height

    def reset(self, stream=None):
        call("vpt_svgf_reset", self.handle, _stream(stream))

    def frame(self, cam, prev_viewproj, cur: DeviceGBuffer, prev: DeviceGBuffer, texels_in, rgb_out, time: float,
              texels_out=None, stream=None):
        vp = (C.c_float * 16)(*map(float, np.asarray(prev_viewproj, np.float32).ravel()))
        call("vpt_svgf_frame", self.handle, C.byref(cam), vp, C.byref(cur.struct()), C.byref(prev.struct()),
             _ptr(texels_in), time, _ptr(rgb_out), _ptr(texels_out), _stream(stream))

    def read_plane(self, which: int, dst, stream=None):
        call("vpt_svgf_read_plane", self.handle, which, _ptr(dst), _stream(stream))

    def read_utility(self, dst, stream=None):
        call("vpt_svgf_read_utility", self.handle, _ptr(dst), _stream(stream))

    def final(self) -> int:
        """Device address of the last frame's final texels (vpt_svgf_final),
        valid until the next frame is enqueued."""
        p = C.c_void_p()
        call("vpt_svgf_final", self.handle, C.byref(p))
        return p.value or 0

    def close(self):
        if self.handle:
            lib().vpt_svgf_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ShadowDenoiser:
    """vpt_shadow_denoise: ShadowTemporalFilter + ShadowFilter with their
    history (Pipeline.cpp:2854-2948)."""

    def __init__(self, ctx: Context, width: int, height: int):
        h = C.c_void_p()
        call("vpt_shadow_denoise_create", ctx.handle, width, height, C.byref(h))
        self.handle = h
        self.width, self.height = width, height

    def reset(self, stream=None):
        call("vpt_shadow_denoise_reset", self.handle, _stream(stream))

    def frame(self, cam, prev_viewproj, cur: DeviceGBuffer, prev: DeviceGBuffer, mask, transversal, out,
              filter_scale: float = 1.0, stream=None):
        """mask / transversal: vpt_trace_shadow2's outputs (uint8, binary16
        bits); out: uint8 tensor for the filtered shadow."""
        vp = (C.c_float * 16)(*map(float, np.asarray(prev_viewproj, np.float32).ravel()))
        call("vpt_shadow_denoise_frame", self.handle, C.byref(cam), vp, C.byref(cur.struct()), C.byref(prev.struct()),
             _ptr(mask), _ptr(transversal), float(filter_scale), _ptr(out), _stream(stream))

    def read(self, which: int, dst, stream=None):
        """which 0: temporal colour (uint8), 1: frame count (binary16 bits)."""
        call("vpt_shadow_denoise_read", self.handle, which, _ptr(dst), _stream(stream))

    def close(self):
        if self.handle:
            lib().vpt_shadow_denoise_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """vpt_graph: HIP-graph replay of one frame's calls.  Use as

        with graph.capture(stream) as s:   # s: the stream to enqueue on
            ...calls on s...

    which captures the calls, instantiates (first frame) or updates the
    graph in place, and replays it ordered after `stream`'s earlier work."""

    def __init__(self, ctx: Context):
        h = C.c_void_p()
        call("vpt_graph_create", ctx.handle, C.byref(h))
        self.handle = h

    def capture(self, stream=None, join_stream=None):
        """The replay starts after `stream`'s earlier work; `join_stream`
        (default `stream`) waits for it to finish."""
        import contextlib

        @contextlib.contextmanager
        def cm():
            cap = C.c_void_p()
            call("vpt_graph_begin", self.handle, _stream(stream), C.byref(cap))
            try:
                yield int(cap.value or 0)
            except BaseException:
                lib().vpt_graph_abort(self.handle)
                raise
            call("vpt_graph_launch", self.handle, _stream(stream if join_stream is None else join_stream))

        return cm()

    def stats(self) -> dict:
        st = _lib.GraphStats()
        call("vpt_graph_stats_get", self.handle, C.byref(st))
        return {k: getattr(st, k) for k, _ in st._fields_ if k != "pad"}

    def close(self):
        if self.handle:
            lib().vpt_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def viewproj(view: np.ndarray, proj: np.ndarray) -> np.ndarray:
    """proj * view for column-major float32 matrices, summed in k order in
    float32 from 0 (u_PrevProjection * u_PrevView, TemporalFilter.glsl:64)."""
    V = np.asarray(view, np.float32).reshape(4, 4)   # V[col][row]
    P = np.asarray(proj, np.float32).reshape(4, 4)
    prod = V[:, :, None] * P[None, :, :]             # prod[c, k, r] = P[k][r] * V[c][k], rounded to float32
    out = np.float32(0.0) + prod[:, 0, :]
    for k in range(1, 4):
        out = out + prod[:, k, :]
    return out.astype(np.float32).ravel()


class Pipeline:
    """Per-frame pass chain for one view of one world, in the reference's
    order (MainPipeline::StartPipeline, Pipeline.cpp:1960-3700): primary,
    sun shadow (with its transversal), its denoise and the reflection rays
    at the reference's settings (1/4 resolution, 2 rough samples) beside
    diffuse GI and SVGF, then the reflection hits shaded from the denoised
    planes and the colour pass's composite into `color` (H, W, 3)
    float32."""

    def __init__(self, ctx: Context, world: World, width: int, height: int, sun_tick: float = 50.0,
                 max_iter: int = 350, shadow_iter: int = 350, use_graph: bool = False):
        """use_graph: replay each frame's three chains — the side chain
        (shadow with transversal, shadow denoise, reflection rays) and the
        main chain (GI, SVGF), launched on their own streams after the
        primary pass so that they still overlap, and the tail (reflection
        shading, composite) after both — as HIP graphs; each is captured
        every frame and its instantiated graph updated in place."""
        import torch
        self.ctx, self.world = ctx, world
        self.width, self.height = width, height
        self.gbuf = [DeviceGBuffer(width, height), DeviceGBuffer(width, height)]
        self.rgb = torch.empty((height, width, 3), dtype=torch.float32, device="cuda")
        self.texels = torch.empty((height, width, 8), dtype=torch.int16, device="cuda")
        self.denoised = torch.empty((height, width, 3), dtype=torch.float32, device="cuda")
        self.shadow = torch.empty((height, width), dtype=torch.uint8, device="cuda")
        self.transversal = torch.empty((height, width), dtype=torch.int16, device="cuda")
        self.shadow_denoised = torch.empty((height, width), dtype=torch.uint8, device="cuda")
        self.reflect_div = 4 if width >= 4 and height >= 4 else 1   # ReflectionTraceResolution 0.25
        hr, wr = height // self.reflect_div, width // self.reflect_div
        self.reflect_dist = torch.empty((hr, wr), dtype=torch.int16, device="cuda")
        self.reflect_hit = torch.empty((hr, wr), dtype=torch.int32, device="cuda")
        self.reflect_color = torch.empty((hr, wr, 4), dtype=torch.int16, device="cuda")
        from .world import RECORD_BYTES, reflect_params
        self.reflect_spp = reflect_params((0.0, 1.0, 0.0))["spp"]
        self.reflect_records = torch.empty((self.reflect_spp * hr * wr * RECORD_BYTES // 4,), dtype=torch.int32,
                                           device="cuda")
        self.color = torch.empty((height, width, 3), dtype=torch.float32, device="cuda")
        self.svgf = SVGF(ctx, width, height)
        self.shadow_denoiser = ShadowDenoiser(ctx, width, height)
        self.side = torch.cuda.Stream()   # shadow, shadow denoise, reflection rays (default priority)
        self.sun_tick = sun_tick
        self.max_iter, self.shadow_iter = max_iter, shadow_iter
        self.frame_index = 0
        self.prev_view = None
        self.prev_proj = None
        self.graphs = (Graph(ctx), Graph(ctx), Graph(ctx)) if use_graph else None   # side, main, tail
        self.fork = torch.cuda.Event()
        self.join = torch.cuda.Event()

    def render(self, pos, yaw: float, pitch: float, stream=None):
        """One frame: primary; then the sun shadow with its transversal, the
        shadow denoise and the reflection rays on a side stream, beside GI
        and SVGF on `stream`; then the reflection shading and the composite.
        Returns the camera."""
        import torch
        main = torch.cuda.current_stream() if stream is None else torch.cuda.ExternalStream(int(stream))
        cam, view, proj = make_camera(pos, yaw, pitch, self.width, self.height)
        cur = self.gbuf[self.frame_index % 2]
        prev = self.gbuf[(self.frame_index + 1) % 2]
        if self.prev_view is None:
            self.prev_view, self.prev_proj = view, proj
        prm = gi_params(self.sun_tick, self.frame_index)
        sun = list(prm.sun_dir)
        if self.frame_index == 0:
            prev = cur
        prev_vp = viewproj(self.prev_view, self.prev_proj)
        cur_vp = viewproj(view, proj)
        self.world.trace_primary(cam, cur, self.max_iter, stream=main.cuda_stream)

        from .world import composite, composite_params, reflect_params, reflect_shading, shade_reflect
        rp = reflect_params(sun, self.frame_index)
        rp["div"] = self.reflect_div

        # the side passes read only the G-buffer and the shadow pass
        def side_chain(s):
            self.world.trace_shadow2(cam, cur, sun, self.shadow, self.transversal, self.shadow_iter, stream=s)
            self.shadow_denoiser.frame(cam, prev_vp, cur, prev, self.shadow, self.transversal, self.shadow_denoised,
                                       stream=s)
            self.world.trace_reflect_deferred(cam, cur, sun, cur_vp, self.reflect_dist, self.reflect_records,
                                              self.reflect_hit, **rp, stream=s)

        def main_chain(s):
            self.world.trace_gi(cam, cur, prm, self.rgb, self.texels, stream=s)
            self.svgf.frame(cam, prev_vp, cur, prev, self.texels, self.denoised, time=self.frame_index / 60.0,
                            stream=s)

        # reflection hits shaded from both chains' outputs, then the composite
        def tail(s):
            diffuse = self.svgf.final()
            shading = reflect_shading(prm, cur_vp, diffuse, self.shadow_denoised)
            shade_reflect(self.ctx, cam, cur, sun, shading, self.reflect_records, self.reflect_color,
                          self.reflect_hit, **rp, stream=s)
            composite(self.ctx, cam, cur, composite_params(prm), diffuse, self.shadow_denoised, self.reflect_color,
                      self.reflect_div, self.color, stream=s)

        # the main chain (the longer one) is enqueued first; the side chain
        # starts from the fork event after the primary pass
        fork, join = self.fork, self.join
        fork.record(main)
        self.side.wait_event(fork)
        if self.graphs is None:
            main_chain(main.cuda_stream)
            side_chain(self.side.cuda_stream)
            join.record(self.side)
            main.wait_event(join)
            tail(main.cuda_stream)
        else:
            # each replay's completion is joined on the stream it started
            # from; the tail's replay waits for both chains
            with self.graphs[1].capture(main.cuda_stream) as s:
                main_chain(s)
            with self.graphs[0].capture(self.side.cuda_stream) as s:
                side_chain(s)
            join.record(self.side)
            main.wait_event(join)
            with self.graphs[2].capture(main.cuda_stream) as s:
                tail(s)
        self.prev_view, self.prev_proj = view, proj
        self.frame_index += 1
        return cam


__all__ = ["SVGF", "Pipeline", "Graph", "decode_texels", "svgf_params_default", "viewproj", "_lib"]
