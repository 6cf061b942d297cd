"""Persistent tracing: ray batches over resident worlds (vpt_batch) and the
persistent per-frame pass chain (vpt_frame), with the launch-shape
autotuner and HIP-graph replay (include/vpt/vpt.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import LaunchCfg, QRay, call, lib
from .world import Context, DeviceGBuffer, World, _ptr, _stream

QRAY_DTYPE = np.dtype([("origin", "<f4", 3), ("dir", "<f4", 3), ("meta", "<u4"), ("dst", "<u4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("voxel", "<u4"), ("face", "u1"), ("block", "u1"), ("samples", "<u2")])
assert QRAY_DTYPE.itemsize == C.sizeof(QRay) == 32 and HIT_DTYPE.itemsize == 12


def make_qrays(origins, dirs, max_iter, world_slot, dst=None) -> np.ndarray:
    """Packs rays into vpt_qray records (meta = max_iter | world << 16)."""
    origins = np.asarray(origins, np.float32)
    n = origins.shape[0]
    q = np.zeros(n, QRAY_DTYPE)
    q["origin"] = origins
    q["dir"] = np.asarray(dirs, np.float32)
    q["meta"] = (np.asarray(max_iter, np.uint32) & 0xFFFF) | (np.asarray(world_slot, np.uint32) << 16)
    q["dst"] = np.arange(n, dtype=np.uint32) if dst is None else np.asarray(dst, np.uint32)
    return q


class Batch:
    """vpt_batch: one persistent launch per ray batch, any mix of worlds."""

    def __init__(self, ctx: Context, max_worlds: int = 16):
        h = C.c_void_p()
        call("vpt_batch_create", ctx.handle, max_worlds, C.byref(h))
        self.handle = h
        self.max_worlds = max_worlds
        self._worlds: dict[int, World] = {}

    def set_world(self, slot: int, world: World | None) -> None:
        call("vpt_batch_set_world", self.handle, slot, world.handle if world else None)
        if world is None:
            self._worlds.pop(slot, None)
        else:
            self._worlds[slot] = world

    @property
    def cfg(self) -> LaunchCfg:
        c = LaunchCfg()
        call("vpt_batch_get_cfg", self.handle, C.byref(c))
        return c

    @cfg.setter
    def cfg(self, c: LaunchCfg) -> None:
        call("vpt_batch_set_cfg", self.handle, C.byref(c))

    def trace(self, rays, hits, n: int, stream=None) -> None:
        """rays / hits: cuda tensors holding n vpt_qray / vpt_hit records."""
        call("vpt_batch_trace", self.handle, _ptr(rays), n, _ptr(hits), _stream(stream))

    def autotune(self, rays, hits, n: int, stream=None) -> float:
        ms = C.c_float()
        call("vpt_batch_autotune", self.handle, _ptr(rays), n, _ptr(hits), C.byref(ms), _stream(stream))
        return ms.value

    def close(self):
        if self.handle:
            lib().vpt_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Frame:
    """vpt_frame: persistent primary / shadow / GI passes of one view."""

    def __init__(self, ctx: Context, width: int, height: int):
        h = C.c_void_p()
        call("vpt_frame_create", ctx.handle, width, height, C.byref(h))
        self.handle = h
        self.width, self.height = width, height

    @property
    def cfg(self) -> LaunchCfg:
        c = LaunchCfg()
        call("vpt_frame_get_cfg", self.handle, C.byref(c))
        return c

    @cfg.setter
    def cfg(self, c: LaunchCfg) -> None:
        call("vpt_frame_set_cfg", self.handle, C.byref(c))

    def trace(self, world: World, cam, prm, gbuf: DeviceGBuffer, mask, rgb, texels=None, primary_iter: int = 350,
              shadow_iter: int = 350, use_graph: bool = False, stream=None) -> None:
        call("vpt_frame_trace", self.handle, world.handle, C.byref(cam), C.byref(prm), primary_iter, shadow_iter,
             C.byref(gbuf.struct()), _ptr(mask), _ptr(rgb), _ptr(texels), int(use_graph), _stream(stream))

    def autotune(self, world: World, cam, prm, gbuf: DeviceGBuffer, mask, rgb, texels=None, primary_iter: int = 350,
                 shadow_iter: int = 350, stream=None) -> float:
        ms = C.c_float()
        call("vpt_frame_autotune", self.handle, world.handle, C.byref(cam), C.byref(prm), primary_iter, shadow_iter,
             C.byref(gbuf.struct()), _ptr(mask), _ptr(rgb), _ptr(texels), C.byref(ms), _stream(stream))
        return ms.value

    def close(self):
        if self.handle:
            lib().vpt_frame_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["Batch", "Frame", "make_qrays", "QRAY_DTYPE", "HIT_DTYPE", "LaunchCfg"]
