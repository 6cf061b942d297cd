"""Serving loop over resident worlds (vpt_server in include/vpt/vpt.h).

Requests are ray queries (primary, shadow or GI rays already in device
memory) or frames (primary + shadow + GI passes of one camera).  The server
admits them while their result memory fits in its arena and runs them one at
a time (``mode="plain"``, one launch per request) or continuously batched
(``mode="batched"``): each launch traces the rays of every admitted request
in one kernel, one lane per ray by default (the persistent refill kernel is
a launch-shape option).  Results stay in server
memory until :meth:`Server.release`; :meth:`Server.fetch` copies them into
torch tensors.
"""
from __future__ import annotations

import ctypes as C

from . import _lib
from ._lib import Camera, GiParams, LaunchCfg, Request, RequestResult, ServerCfg, ServerStats, call, lib
from .world import Context, World, _ptr

KINDS = {"primary": _lib.VPT_REQ_PRIMARY, "shadow": _lib.VPT_REQ_SHADOW, "gi": _lib.VPT_REQ_GI,
         "frame": _lib.VPT_REQ_FRAME}
_HIP_MEMCPY_DEFAULT = 4
_hip = None


def _copy(dst_tensor, src_ptr: int, nbytes: int) -> None:
    """Device-to-device copy of nbytes from a raw device pointer (the
    caller has synchronised with the producing work)."""
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMemcpy.restype = C.c_int
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    if nbytes == 0:
        return
    rc = _hip.hipMemcpy(C.c_void_p(dst_tensor.data_ptr()), C.c_void_p(src_ptr), C.c_size_t(nbytes),
                        _HIP_MEMCPY_DEFAULT)
    if rc != 0:
        raise RuntimeError(f"hipMemcpy failed with {rc}")


def now_ns() -> int:
    return int(lib().vpt_now_ns())


class Server:
    def __init__(self, ctx: Context, mode: str = "batched", max_worlds: int = 16, arena_bytes: int = 1 << 30,
                 world_budget: int = 16 << 30, max_launch_rays: int = 1 << 21, max_inflight: int = 2,
                 launch_cfg: LaunchCfg | None = None):
        cfg = ServerCfg(_lib.VPT_SERVE_BATCHED if mode == "batched" else _lib.VPT_SERVE_PLAIN, max_worlds,
                        arena_bytes, world_budget, max_launch_rays, max_inflight)
        if mode not in ("batched", "plain"):
            raise ValueError(mode)
        h = C.c_void_p()
        call("vpt_server_create", ctx.handle, C.byref(cfg), C.byref(h))
        self.handle = h
        self.mode = mode
        self._worlds: dict[int, World] = {}
        self._keep: dict[int, object] = {}   # ray tensors of in-flight requests
        if launch_cfg is not None:
            call("vpt_server_set_launch_cfg", self.handle, C.byref(launch_cfg))

    def add_world(self, world: World) -> int:
        slot = C.c_int()
        call("vpt_server_add_world", self.handle, world.handle, C.byref(slot))
        self._worlds[slot.value] = world
        return slot.value

    def remove_world(self, slot: int) -> None:
        call("vpt_server_remove_world", self.handle, slot)
        self._worlds.pop(slot, None)

    def submit_rays(self, slot: int, kind: str, rays, n: int, max_iter: int, sun=(0.0, 1.0, 0.0),
                    shadow_iter: int = 128, arrival_ns: int = 0) -> int:
        """rays: cuda tensor holding n vpt_ray records (origin, dir float32),
        written by work on the current stream (synchronised here)."""
        import torch
        torch.cuda.current_stream().synchronize()
        r = Request()
        r.world = slot
        r.kind = KINDS[kind]
        r.n = n
        r.max_iter = max_iter
        r.d_rays = _ptr(rays).value
        r.sun_dir[:] = [float(x) for x in sun]
        r.shadow_iter = shadow_iter
        r.arrival_ns = arrival_ns
        t = C.c_uint64()
        call("vpt_server_submit", self.handle, C.byref(r), C.byref(t))
        self._keep[t.value] = rays
        return t.value

    def submit_frame(self, slot: int, cam: Camera, gi: GiParams, primary_iter: int = 350, shadow_iter: int = 350,
                     arrival_ns: int = 0) -> int:
        r = Request()
        r.world = slot
        r.kind = _lib.VPT_REQ_FRAME
        r.cam = C.pointer(cam)
        r.gi = C.pointer(gi)
        r.primary_iter = primary_iter
        r.frame_shadow_iter = shadow_iter
        r.arrival_ns = arrival_ns
        t = C.c_uint64()
        call("vpt_server_submit", self.handle, C.byref(r), C.byref(t))
        return t.value

    def wait(self, ticket: int, timeout_ms: int = -1) -> RequestResult:
        res = RequestResult()
        call("vpt_server_wait", self.handle, C.c_uint64(ticket), timeout_ms, C.byref(res))
        return res

    def fetch(self, ticket: int) -> dict:
        """Waits for the request and copies its results into new cuda tensors."""
        import torch
        res = self.wait(ticket)
        if res.status != 0:
            raise _lib.VptError(res.status, f"request {ticket} failed")
        out = {"t_arrival_ns": res.t_arrival_ns, "t_admit_ns": res.t_admit_ns, "t_start_ns": res.t_start_ns,
               "t_done_ns": res.t_done_ns}
        n = res.n
        if res.kind == _lib.VPT_REQ_FRAME:
            for name, ptr, dt, per in (("t", res.gbuffer.d_t, torch.float32, 1), ("face", res.gbuffer.d_face,
                                       torch.uint8, 1), ("block", res.gbuffer.d_block, torch.uint8, 1),
                                       ("voxel", res.gbuffer.d_voxel, torch.int32, 1),
                                       ("mask", res.d_mask, torch.uint8, 1), ("rgb", res.d_rgb, torch.float32, 3)):
                t = torch.empty(n * per, dtype=dt, device="cuda")
                _copy(t, ptr, t.numel() * t.element_size())
                out[name] = t
        else:
            hits = torch.empty(n * 12, dtype=torch.uint8, device="cuda")
            _copy(hits, res.d_hits or 0, n * 12)
            out["hits"] = hits
            if res.kind == _lib.VPT_REQ_GI:
                vis = torch.empty(n, dtype=torch.uint8, device="cuda")
                _copy(vis, res.d_vis or 0, n)
                out["vis"] = vis
        return out

    def release(self, ticket: int) -> None:
        call("vpt_server_release", self.handle, C.c_uint64(ticket))
        self._keep.pop(ticket, None)

    def stats(self) -> dict:
        s = ServerStats()
        call("vpt_server_stats_get", self.handle, C.byref(s))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def close(self) -> None:
        if self.handle:
            call("vpt_server_destroy", self.handle)
            self.handle = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
