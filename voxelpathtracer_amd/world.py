"""Host-side mirror of the reference's ``VoxelRT::World`` (Core/World.h:35-222)
over the C ABI.

The reference keeps the block grid inline (``m_WorldData``), uploads it as a
3D texture (``Buffer``), builds the distance field with three compute
dispatches (``InitializeDistanceGenerator`` / ``GenerateDistanceField``) and
offers CPU block picking (``RaycastDetect``).  Here the grid copy stays on
the host as a numpy array (z, y, x), the device copy lives in HBM in bricked
layout, and every GPU step is a libvpt call.  Device buffers handed to the
ray kernels are torch tensors (memory and streams only).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import Camera, GBuffer, Hit, Ray, call, lib


def _stream(stream=None) -> C.c_void_p:
    if stream is None:
        import torch
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)
    return C.c_void_p(int(stream))


def _ptr(t) -> C.c_void_p:
    return C.c_void_p(0 if t is None else t.data_ptr())


class Context:
    """vpt_ctx: binds one HIP device."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        call("vpt_ctx_create", device, C.byref(h))
        self.handle = h
        self.device = device

    def close(self):
        if self.handle:
            lib().vpt_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class World:
    """VoxelRT::World: block grid + distance field resident in HBM."""

    def __init__(self, ctx: Context, nx: int = 384, ny: int = 128, nz: int = 384):
        h = C.c_void_p()
        call("vpt_world_create", ctx.handle, nx, ny, nz, C.byref(h))
        self.ctx = ctx
        self.handle = h
        self.nx, self.ny, self.nz = nx, ny, nz
        self.m_WorldData = np.zeros((nz, ny, nx), np.uint8)
        self._pending: dict[int, int] = {}
        self._counter = None

    # -- World.h:46-70 ---------------------------------------------------
    def GetBlock(self, x: int, y: int, z: int) -> int:
        return int(self.m_WorldData[z, y, x])

    def SetBlock(self, x: int, y: int, z: int, block: int) -> None:
        """Host edit; becomes visible on the device at the next
        ``Buffer`` / ``FlushEdits`` (the reference follows SetBlock with a
        one-voxel glTexSubImage3D, World.cpp:367-374)."""
        self.m_WorldData[z, y, x] = block
        self._pending[x | (y << 10) | (z << 20)] = block

    def FlushEdits(self, stream=None) -> None:
        if not self._pending:
            return
        edits = np.fromiter(self._pending.keys(), np.uint32)
        ids = np.fromiter(self._pending.values(), np.uint8)
        call("vpt_world_set_blocks", self.handle, edits.ctypes.data_as(C.c_void_p), ids.ctypes.data_as(C.c_void_p),
             len(edits), _stream(stream))
        self._pending.clear()

    # -- World.h:167-174 -------------------------------------------------
    def Buffer(self, stream=None) -> None:
        grid = np.ascontiguousarray(self.m_WorldData)
        call("vpt_world_upload", self.handle, grid.ctypes.data_as(C.c_void_p), _stream(stream))
        self._pending.clear()

    def InitializeDistanceGenerator(self) -> None:
        """The DF storage is allocated with the world; nothing to compile."""

    def GenerateDistanceField(self, stream=None) -> None:
        self.FlushEdits(stream)
        call("vpt_df_build", self.handle, _stream(stream))

    def UpdateDistanceField(self, stream=None, region: bool = False):
        """Flushes pending SetBlock edits and updates the DF around them
        (vpt_df_update, asynchronous).  With region=True waits for the
        stream and returns the recomputed box (x0, y0, z0, x1, y1, z1)."""
        if not self._pending:
            return None
        edits = np.fromiter(self._pending.keys(), np.uint32)
        self.FlushEdits(stream)
        box = (C.c_int32 * 6)() if region else None
        call("vpt_df_update", self.handle, edits.ctypes.data_as(C.c_void_p), len(edits), box, _stream(stream))
        return tuple(box) if region else None

    # -- World.h:186 -----------------------------------------------------
    def RaycastDetect(self, pos, direction) -> tuple[int, int, int, int]:
        out = (C.c_int32 * 4)()
        p = (C.c_float * 3)(*map(float, pos))
        d = (C.c_float * 3)(*map(float, direction))
        grid = np.ascontiguousarray(self.m_WorldData)
        call("vpt_raycast_detect", grid.ctypes.data_as(C.c_void_p), self.nx, self.ny, self.nz, p, d, out)
        return tuple(out)

    # -- downloads (parity) ------------------------------------------------
    def download_df(self, stream=None) -> np.ndarray:
        df = np.empty((self.nz, self.ny, self.nx), np.uint8)
        call("vpt_df_download", self.handle, df.ctypes.data_as(C.c_void_p), _stream(stream))
        return df

    def download_grid(self, stream=None) -> np.ndarray:
        g = np.empty((self.nz, self.ny, self.nx), np.uint8)
        call("vpt_grid_download", self.handle, g.ctypes.data_as(C.c_void_p), _stream(stream))
        return g

    def device_bytes(self) -> int:
        n = C.c_size_t()
        call("vpt_world_bytes", self.handle, C.byref(n))
        return n.value

    def set_sample_counter(self, counter_tensor) -> None:
        """counter_tensor: int64 cuda tensor of two elements (DF samples,
        traced rays), or None to stop counting."""
        self._counter = counter_tensor
        call("vpt_world_set_sample_counter", self.handle, _ptr(counter_tensor))

    # -- tracing ---------------------------------------------------------
    def trace_primary(self, cam: Camera, gbuf: "DeviceGBuffer", max_iter: int = 350, stream=None) -> None:
        call("vpt_trace_primary", self.handle, C.byref(cam), max_iter, C.byref(gbuf.struct()), _stream(stream))

    def trace_shadow(self, cam: Camera, gbuf: "DeviceGBuffer", sun_dir, mask, max_iter: int = 350,
                     stream=None) -> None:
        sd = (C.c_float * 3)(*map(float, sun_dir))
        call("vpt_trace_shadow", self.handle, C.byref(cam), C.byref(gbuf.struct()), sd, max_iter, _ptr(mask),
             _stream(stream))

    def trace_shadow2(self, cam: Camera, gbuf: "DeviceGBuffer", sun_dir, mask, transversal, max_iter: int = 350,
                      stream=None) -> None:
        """The shadow pass with its transversal output (binary16 bits in an int16/uint16 tensor)."""
        sd = (C.c_float * 3)(*map(float, sun_dir))
        call("vpt_trace_shadow2", self.handle, C.byref(cam), C.byref(gbuf.struct()), sd, max_iter, _ptr(mask),
             _ptr(transversal), _stream(stream))

    def trace_reflect(self, cam: Camera, gbuf: "DeviceGBuffer", light_dir, dist, hit=None, div: int = 4,
                      max_iter: int = 64, shadow_iter: int = 150, spp: int = 1, rough: int = 0,
                      roughness: float = 0.425, frame: int = 0, stream=None) -> None:
        """Reflections off the G-buffer (ReflectionTraceFrag.glsl:764-1035,
        traversal core) at (W // div) x (H // div): mirror (rough 0) or GGX
        sampled (rough 1), spp samples.  `dist` receives the binary16 bits of
        o_HitDistance (int16/uint16 tensor), `hit` (optional, int32/uint32)
        face | block << 8 | shadow << 16 | hits << 24, 255 without a hit.
        reflect_params() gives the reference's defaults."""
        prm = _lib.ReflectParams((C.c_float * 3)(*map(float, light_dir)), div, max_iter, shadow_iter, spp, rough,
                                 float(roughness), int(frame))
        call("vpt_trace_reflect", self.handle, C.byref(cam), C.byref(gbuf.struct()), C.byref(prm), _ptr(dist),
             _ptr(hit), _stream(stream))

    def trace_reflect_shaded(self, cam: Camera, gbuf: "DeviceGBuffer", light_dir, shading, dist, color, hit=None,
                             div: int = 4, max_iter: int = 64, shadow_iter: int = 150, spp: int = 1, rough: int = 0,
                             roughness: float = 0.425, frame: int = 0, stream=None) -> None:
        """trace_reflect plus the hit shading (vpt_trace_reflect_shaded):
        hits reprojected into the frame's denoised diffuse and shadow,
        shaded with the albedo table; `color` receives RGBA binary16 bits
        per target pixel ((H // div, W // div, 4) int16/uint16 tensor).
        shading: a ReflectShading (reflect_shading() builds one)."""
        prm = _lib.ReflectParams((C.c_float * 3)(*map(float, light_dir)), div, max_iter, shadow_iter, spp, rough,
                                 float(roughness), int(frame))
        call("vpt_trace_reflect_shaded", self.handle, C.byref(cam), C.byref(gbuf.struct()), C.byref(prm),
             C.byref(shading), _ptr(dist), _ptr(hit), _ptr(color), _stream(stream))

    def trace_reflect_deferred(self, cam: Camera, gbuf: "DeviceGBuffer", light_dir, viewproj, dist, records, hit=None,
                               div: int = 4, max_iter: int = 64, shadow_iter: int = 150, spp: int = 1, rough: int = 0,
                               roughness: float = 0.425, frame: int = 0, stream=None) -> None:
        """The traversal half of trace_reflect_shaded (vpt_trace_reflect_deferred):
        `records` receives spp records of RECORD_BYTES per target pixel
        (an int32 tensor of spp * (H // div) * (W // div) * 4 elements);
        shade_reflect() completes the pass."""
        prm = _lib.ReflectParams((C.c_float * 3)(*map(float, light_dir)), div, max_iter, shadow_iter, spp, rough,
                                 float(roughness), int(frame))
        vp = (C.c_float * 16)(*map(float, np.asarray(viewproj, np.float32).ravel()))
        call("vpt_trace_reflect_deferred", self.handle, C.byref(cam), C.byref(gbuf.struct()), C.byref(prm), vp,
             _ptr(dist), _ptr(hit), _ptr(records), _stream(stream))

    def trace_frame(self, cam: Camera, gbuf: "DeviceGBuffer", prm, mask, rgb, texels=None, primary_iter: int = 350,
                    shadow_iter: int = 350, transversal=None, stream=None) -> None:
        """Primary, shadow and GI passes of one frame; the shadow pass runs
        beside GI on a stream of the context (vpt_trace_frame).  With
        `transversal` the shadow pass also writes its binary16 transversal
        (vpt_trace_shadow2)."""
        call("vpt_trace_frame", self.handle, C.byref(cam), C.byref(prm), primary_iter, shadow_iter,
             C.byref(gbuf.struct()), _ptr(mask), _ptr(transversal), _ptr(rgb), _ptr(texels), _stream(stream))

    def trace_gi(self, cam: Camera, gbuf: "DeviceGBuffer", prm, rgb, texels=None, stream=None) -> None:
        call("vpt_trace_gi", self.handle, C.byref(cam), C.byref(gbuf.struct()), C.byref(prm), _ptr(rgb),
             _ptr(texels), _stream(stream))

    def trace_rays(self, rays, hits, n: int | None = None, max_iter: int = 350, stream=None) -> None:
        """rays: uint8/float cuda tensor holding vpt_ray records; hits: tensor for vpt_hit records."""
        if n is None:
            n = rays.numel() * rays.element_size() // C.sizeof(Ray)
        call("vpt_trace_rays", self.handle, _ptr(rays), n, max_iter, _ptr(hits), _stream(stream))

    def trace_gi_rays(self, rays, hits, vis, n: int, gi_iter: int = 48, sun_dir=(0.0, 1.0, 0.0),
                      shadow_iter: int = 128, stream=None) -> None:
        """GI ray queries: bounce hits into `hits`, sun visibility at the hit
        into `vis` (1 blocked, 0 visible, 255 no shadow ray)."""
        sd = (C.c_float * 3)(*map(float, sun_dir))
        call("vpt_trace_gi_rays", self.handle, _ptr(rays), n, gi_iter, sd, shadow_iter, _ptr(hits), _ptr(vis),
             _stream(stream))

    def ImportMinecraftWorld(self, directory: str, origin=(0.0, 0.0, 0.0), mc_to_block=None, names=None,
                             stream=None) -> list:
        """MCWorldImporter::ImportWorld (NBT/Importer.cpp:150-168): the
        region files of `directory` into m_WorldData, then Buffer().  See
        mc_import for the arguments; returns its stats.  Rebuild the DF
        afterwards (Pipeline.cpp:1293-1294)."""
        grid, stats = mc_import(directory, (self.nx, self.ny, self.nz), origin, mc_to_block, names)
        self.m_WorldData[...] = grid
        self.Buffer(stream)
        return stats

    def SaveWorld(self, path: str) -> None:
        """Raw world file: nx*ny*nz block ids, x fastest (WorldFileHandler.cpp:10-37)."""
        call("vpt_world_save", self.handle, str(path).encode())

    def LoadWorld(self, path: str, stream=None) -> None:
        """Loads a raw world file into the device grid and m_WorldData
        (WorldFileHandler.cpp:39-83); rebuild the DF afterwards."""
        call("vpt_world_load", self.handle, str(path).encode(), _stream(stream))
        call("vpt_grid_download", self.handle, self.m_WorldData.ctypes.data_as(C.c_void_p), _stream(stream))

    def close(self):
        if self.handle:
            lib().vpt_world_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceGBuffer:
    """InitialTraceFBO planes (t, face, block, voxel) as cuda tensors."""

    def __init__(self, width: int, height: int, device="cuda"):
        import torch
        self.width, self.height = width, height
        self.t = torch.empty((height, width), dtype=torch.float32, device=device)
        self.face = torch.empty((height, width), dtype=torch.uint8, device=device)
        self.block = torch.empty((height, width), dtype=torch.uint8, device=device)
        self.voxel = torch.empty((height, width), dtype=torch.int32, device=device)

    def struct(self) -> GBuffer:
        return GBuffer(self.t.data_ptr(), self.face.data_ptr(), self.block.data_ptr(), self.voxel.data_ptr(),
                       self.width, self.height)

    def numpy(self) -> dict:
        return dict(t=self.t.cpu().numpy(), face=self.face.cpu().numpy(), block=self.block.cpu().numpy(),
                    voxel=self.voxel.cpu().numpy().view(np.uint32))


def make_camera(pos, yaw_deg: float, pitch_deg: float, width: int, height: int, fov_deg: float = 60.0,
                znear: float = 0.1, zfar: float = 1000.0):
    """glm lookAt/perspective camera (FpsCamera.cpp).  Returns (Camera, view, proj)."""
    cam = Camera()
    view = (C.c_float * 16)()
    proj = (C.c_float * 16)()
    p = (C.c_float * 3)(*map(float, pos))
    call("vpt_camera_make", p, yaw_deg, pitch_deg, fov_deg, width, height, znear, zfar, C.byref(cam), view, proj)
    return cam, np.array(view, np.float32), np.array(proj, np.float32)


def generate_terrain(nx: int, ny: int, nz: int, seed: int = 0x5EED0001, seed_random: int = 0x5EED0002,
                     h_min: int = 20, h_max: int = 90, random_permille: int = 10) -> np.ndarray:
    grid = np.empty((nz, ny, nx), np.uint8)
    call("vpt_generate_terrain", nx, ny, nz, seed, seed_random, h_min, h_max, random_permille,
         grid.ctypes.data_as(C.c_void_p))
    return grid


def gi_params(sun_tick: float = 50.0, frame: int = 0, albedo_seed: int = 0x5EED0003):
    prm = _lib.GiParams()
    call("vpt_gi_params_default", sun_tick, frame, albedo_seed, C.byref(prm))
    return prm


__all__ = ["Context", "World", "DeviceGBuffer", "make_camera", "generate_terrain", "gi_params", "Ray", "Hit"]


def mc_import(directory: str, dims, origin=(0.0, 0.0, 0.0), mc_to_block=None, names=None):
    """Minecraft region files -> a (nz, ny, nx) uint8 block grid and stats
    [files, chunks, sections] (vpt_mc_import, host only).

    mc_to_block: 256 legacy-id -> block-id bytes (BlockDatabase::
    GetIDFromMCID); the identity when None.  names: {namespaced id: (legacy
    id, data value)} for palette-format chunks; a name not listed imports as
    stone (1:0)."""
    nx, ny, nz = map(int, dims)
    lut = np.arange(256, dtype=np.uint8) if mc_to_block is None else np.ascontiguousarray(mc_to_block, np.uint8)
    if lut.shape != (256,):
        raise ValueError("mc_to_block must hold 256 entries")
    items = list((names or {}).items())
    table = (_lib.McBlock * max(len(items), 1))()
    keep = []
    for k, (name, (bid, dv)) in enumerate(items):
        b = name.encode()
        keep.append(b)
        table[k] = _lib.McBlock(b, int(bid), int(dv))
    grid = np.empty((nz, ny, nx), np.uint8)
    stats = (C.c_int32 * 3)()
    org = (C.c_float * 3)(*map(float, origin))
    call("vpt_mc_import", str(directory).encode(), nx, ny, nz, org, lut.ctypes.data_as(C.c_void_p),
         C.cast(table, C.c_void_p), len(items), grid.ctypes.data_as(C.c_void_p), stats)
    return grid, list(stats)


def _addr(x) -> int:
    """Device address of a tensor or an integer pointer."""
    return x if isinstance(x, int) else _ptr(x).value or 0


def reflect_shading(gi: _lib.GiParams, viewproj, diffuse, shadow) -> _lib.ReflectShading:
    """Hit shading of the reflection pass from the frame's GI constants
    (vpt_reflect_shading_from_gi): diffuse: the denoised SVGF texels (a
    tensor or SVGF.final()), shadow: the denoised shadow (uint8 tensor),
    viewproj: the frame's projection * view."""
    out = _lib.ReflectShading()
    vp = (C.c_float * 16)(*map(float, np.asarray(viewproj, np.float32).ravel()))
    call("vpt_reflect_shading_from_gi", C.byref(gi), vp, C.c_void_p(_addr(diffuse)), C.c_void_p(_addr(shadow)),
         C.byref(out))
    return out


def composite_params(gi: _lib.GiParams) -> _lib.CompositeParams:
    """The colour composite's constants from the frame's GI constants
    (vpt_composite_params_from_gi)."""
    out = _lib.CompositeParams()
    call("vpt_composite_params_from_gi", C.byref(gi), C.byref(out))
    return out


def composite(ctx: Context, cam: Camera, gbuf: "DeviceGBuffer", prm: _lib.CompositeParams, diffuse, shadow, reflect,
              reflect_div: int, rgb, stream=None) -> None:
    """The colour pass's composite (vpt_composite): rgb (H, W, 3) float32
    from the G-buffer, the denoised diffuse texels (tensor or
    SVGF.final()), the denoised shadow and the shaded reflection colour at
    (H // reflect_div, W // reflect_div)."""
    call("vpt_composite", ctx.handle, C.byref(cam), C.byref(gbuf.struct()), C.byref(prm), C.c_void_p(_addr(diffuse)),
         _ptr(shadow), _ptr(reflect), int(reflect_div), _ptr(rgb), _stream(stream))


RECORD_BYTES = 16   # VPT_REFLECT_RECORD_BYTES


def shade_reflect(ctx: Context, cam: Camera, gbuf: "DeviceGBuffer", light_dir, shading, records, color, hit=None,
                  div: int = 4, max_iter: int = 64, shadow_iter: int = 150, spp: int = 1, rough: int = 0,
                  roughness: float = 0.425, frame: int = 0, stream=None) -> None:
    """The shading half of the deferred reflection pass (vpt_shade_reflect):
    the same parameters as the World.trace_reflect_deferred call whose
    records it shades."""
    prm = _lib.ReflectParams((C.c_float * 3)(*map(float, light_dir)), div, max_iter, shadow_iter, spp, rough,
                             float(roughness), int(frame))
    call("vpt_shade_reflect", ctx.handle, C.byref(cam), C.byref(gbuf.struct()), C.byref(prm), C.byref(shading),
         _ptr(records), _ptr(hit), _ptr(color), _stream(stream))


def reflect_params(light_dir, frame: int = 0) -> dict:
    """The reference's reflection settings (vpt_reflect_params_default) as
    World.trace_reflect keyword arguments."""
    p = _lib.ReflectParams()
    call("vpt_reflect_params_default", (C.c_float * 3)(*map(float, light_dir)), int(frame), C.byref(p))
    return dict(div=p.div, max_iter=p.max_iter, shadow_iter=p.shadow_iter, spp=p.spp, rough=p.rough,
                roughness=p.roughness, frame=p.frame)
