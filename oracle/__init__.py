"""CPU oracle — TEST INFRASTRUCTURE ONLY.

ctypes wrapper of ``oracle/_build/libvpt_oracle.so`` (built from
``oracle/vpt_oracle.cpp`` by ``oracle/Makefile``), a line-for-line C++
restatement of the reference shaders.  Only ``tests/``,
``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` may use
it, and only as the checker: the product path never calls it.

All arrays are numpy; images are row-major (H, W, ...).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

_DIR = Path(__file__).resolve().parent
LIB_PATH = _DIR / "_build" / "libvpt_oracle.so"


class OGiParams(C.Structure):
    _fields_ = [("sun_dir", C.c_float * 3), ("light_color", C.c_float * 3), ("sky_radiance", C.c_float * 3),
                ("sky_pixel_radiance", C.c_float * 3), ("albedo", C.c_float * 384), ("frame", C.c_uint32),
                ("gi_iter", C.c_int32), ("shadow_iter", C.c_int32)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            import subprocess
            subprocess.run(["make", "-s", "-C", str(_DIR)], check=True)
        _lib = C.CDLL(str(LIB_PATH))
    return _lib


def _p(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def df_build(grid: np.ndarray) -> np.ndarray:
    """grid: (nz, ny, nx) uint8 -> DF (nz, ny, nx) uint8."""
    grid = np.ascontiguousarray(grid, dtype=np.uint8)
    nz, ny, nx = grid.shape
    df = np.empty_like(grid)
    lib().oracle_df_build(_p(grid), nx, ny, nz, _p(df))
    return df


def trace_rays(grid, df, rays: np.ndarray, max_iter: int):
    """rays: (n, 6) float32 origin+dir.  Returns dict of t, cell (n,3), face, block, samples."""
    nz, ny, nx = grid.shape
    rays = _f32(rays)
    n = rays.shape[0]
    t = np.empty(n, np.float32)
    cell = np.empty((n, 3), np.int32)
    face = np.empty(n, np.uint8)
    block = np.empty(n, np.uint8)
    samples = np.empty(n, np.int32)
    lib().oracle_trace_rays(_p(grid), _p(df), nx, ny, nz, _p(rays), C.c_int(n), C.c_int(max_iter), _p(t), _p(cell),
                            _p(face), _p(block), _p(samples))
    return dict(t=t, cell=cell, face=face, block=block, samples=samples)


def trace_primary(grid, df, inv_view, inv_proj, W: int, H: int, max_iter: int = 350):
    nz, ny, nx = grid.shape
    iv, ip = _f32(inv_view), _f32(inv_proj)
    t = np.empty((H, W), np.float32)
    face = np.empty((H, W), np.uint8)
    block = np.empty((H, W), np.uint8)
    voxel = np.empty((H, W), np.uint32)
    samples = np.empty((H, W), np.int32)
    lib().oracle_trace_primary(_p(grid), _p(df), nx, ny, nz, _p(iv), _p(ip), W, H, max_iter, _p(t), _p(face),
                               _p(block), _p(voxel), _p(samples))
    return dict(t=t, face=face, block=block, voxel=voxel, samples=samples)


def trace_shadow(grid, df, inv_view, inv_proj, gb_t, gb_face, sun_dir, max_iter: int = 350):
    nz, ny, nx = grid.shape
    H, W = gb_t.shape
    mask = np.empty((H, W), np.uint8)
    samples = np.empty((H, W), np.int32)
    rays = np.empty((H, W), np.int32)
    sun = _f32(sun_dir)
    lib().oracle_trace_shadow(_p(grid), _p(df), nx, ny, nz, _p(_f32(inv_view)), _p(_f32(inv_proj)), W, H,
                              _p(_f32(gb_t)), _p(np.ascontiguousarray(gb_face, np.uint8)), _p(sun), max_iter,
                              _p(mask), _p(samples), _p(rays))
    return dict(mask=mask, samples=samples, rays=rays)


def trace_shadow2(grid, df, inv_view, inv_proj, gb_t, gb_face, sun_dir, max_iter: int = 350):
    """Shadow pass with its transversal output (binary16 bits), ShadowRayTraceFrag.glsl:414-513."""
    nz, ny, nx = grid.shape
    H, W = gb_t.shape
    mask = np.empty((H, W), np.uint8)
    tv = np.empty((H, W), np.uint16)
    samples = np.empty((H, W), np.int32)
    rays = np.empty((H, W), np.int32)
    lib().oracle_trace_shadow2(_p(grid), _p(df), nx, ny, nz, _p(_f32(inv_view)), _p(_f32(inv_proj)), W, H,
                               _p(_f32(gb_t)), _p(np.ascontiguousarray(gb_face, np.uint8)), _p(_f32(sun_dir)), max_iter,
                               _p(mask), _p(tv), _p(samples), _p(rays))
    return dict(mask=mask, transversal=tv, samples=samples, rays=rays)


def trace_reflect(grid, df, inv_view, inv_proj, gb_t, gb_face, light_dir, div: int = 4, max_iter: int = 64,
                  shadow_iter: int = 150, spp: int = 1, rough: int = 0, roughness: float = 0.425, frame: int = 0):
    """Reflections (ReflectionTraceFrag.glsl:764-1035, traversal core) at
    (W // div) x (H // div), mirror (rough 0) or GGX-sampled (rough 1), spp
    samples: hit distance (binary16 bits), hit record face | block << 8 |
    shadow << 16 | hits << 24 (255 without a hit), samples, rays."""
    nz, ny, nx = grid.shape
    H, W = gb_t.shape
    Hr, Wr = H // div, W // div
    dist = np.empty((Hr, Wr), np.uint16)
    hit = np.empty((Hr, Wr), np.uint32)
    samples = np.empty((Hr, Wr), np.int32)
    rays = np.empty((Hr, Wr), np.int32)
    lib().oracle_trace_reflect(_p(grid), _p(df), nx, ny, nz, _p(_f32(inv_view)), _p(_f32(inv_proj)), W, H,
                               _p(_f32(gb_t)), _p(np.ascontiguousarray(gb_face, np.uint8)), _p(_f32(light_dir)), div,
                               max_iter, shadow_iter, spp, rough, C.c_float(roughness), C.c_uint32(frame), _p(dist),
                               _p(hit), _p(samples), _p(rays))
    return dict(dist=dist, hit=hit, samples=samples, rays=rays)


def trace_reflect_shaded(grid, df, inv_view, inv_proj, gb_t, gb_face, light_dir, sh, cocg, ao, shadow, viewproj,
                         sun_rad, sky_rad, albedo, div: int = 4, max_iter: int = 64, shadow_iter: int = 150,
                         spp: int = 1, rough: int = 0, roughness: float = 0.425, frame: int = 0):
    """The reflection pass with reprojection into the frame's denoised
    diffuse and hit shading (ReflectionTraceFrag.glsl:582-593, 764-1035).
    sh (H, W, 4), cocg (H, W, 2), ao (H, W): the denoised diffuse planes;
    shadow (H, W): the denoised sun shadow in [0, 1]; viewproj: the frame's
    u_Projection * u_View.  Adds colour: (Hr, Wr, 4) binary16 bits."""
    nz, ny, nx = grid.shape
    H, W = gb_t.shape
    Hr, Wr = H // div, W // div
    dist = np.empty((Hr, Wr), np.uint16)
    hit = np.empty((Hr, Wr), np.uint32)
    color = np.empty((Hr, Wr, 4), np.uint16)
    samples = np.empty((Hr, Wr), np.int32)
    rays = np.empty((Hr, Wr), np.int32)
    lib().oracle_trace_reflect_shaded(_p(grid), _p(df), nx, ny, nz, _p(_f32(inv_view)), _p(_f32(inv_proj)), W, H,
                                      _p(_f32(gb_t)), _p(np.ascontiguousarray(gb_face, np.uint8)), _p(_f32(light_dir)),
                                      div, max_iter, shadow_iter, spp, rough, C.c_float(roughness), C.c_uint32(frame),
                                      _p(_f32(sh)), _p(_f32(cocg)), _p(_f32(ao)), _p(_f32(shadow)), _p(_f32(viewproj)),
                                      _p(_f32(sun_rad)), _p(_f32(sky_rad)), _p(_f32(albedo)), _p(dist), _p(hit),
                                      _p(color), _p(samples), _p(rays))
    return dict(dist=dist, hit=hit, color=color, samples=samples, rays=rays)


def composite(inv_view, inv_proj, gb_t, gb_face, gb_block, sh, cocg, ao, shadow, refl, rdiv, sun_dir, moon_dir,
              sun_rad, moon_rad, sky_pixel, albedo, roughness: float):
    """The colour pass's composite (ColorPassFrag.glsl:773-946, DFG branch)
    for flat dielectric materials: (H, W, 3) float32.  refl: (Hr, Wr, 4)
    reflection colour values (floats of its binary16 texels)."""
    H, W = gb_t.shape
    out = np.empty((H, W, 3), np.float32)
    lib().oracle_composite(_p(_f32(inv_view)), _p(_f32(inv_proj)), W, H, _p(_f32(gb_t)),
                           _p(np.ascontiguousarray(gb_face, np.uint8)), _p(np.ascontiguousarray(gb_block, np.uint8)),
                           _p(_f32(sh)), _p(_f32(cocg)), _p(_f32(ao)), _p(_f32(shadow)), _p(_f32(refl)), int(rdiv),
                           _p(_f32(sun_dir)), _p(_f32(moon_dir)), _p(_f32(sun_rad)), _p(_f32(moon_rad)),
                           _p(_f32(sky_pixel)), _p(_f32(albedo)), C.c_float(roughness), _p(out))
    return out


def shadow_temporal(inv_view, inv_proj, prev_viewproj, cur_t, cur_face, prev_t, mask, transversal, prev_color,
                    prev_frames):
    """ShadowTemporalFilter.main: (colour u8, frame count binary16 bits)."""
    H, W = cur_t.shape
    col = np.empty((H, W), np.uint8)
    frm = np.empty((H, W), np.uint16)
    u8 = lambda a: np.ascontiguousarray(a, np.uint8)
    u16 = lambda a: np.ascontiguousarray(a, np.uint16)
    lib().oracle_shadow_temporal(_p(_f32(inv_view)), _p(_f32(inv_proj)), _p(_f32(prev_viewproj)), W, H,
                                 _p(_f32(cur_t)), _p(u8(cur_face)), _p(_f32(prev_t)), _p(u8(mask)), _p(u16(transversal)),
                                 _p(u8(prev_color)), _p(u16(prev_frames)), _p(col), _p(frm))
    return col, frm


def shadow_filter(inv_view, inv_proj, gb_t, gb_face, color, transversal, frames, scale: float = 1.0):
    """ShadowFilter.main: filtered shadow, u8 unorm."""
    H, W = gb_t.shape
    out = np.empty((H, W), np.uint8)
    lib().oracle_shadow_filter(_p(_f32(inv_view)), _p(_f32(inv_proj)), W, H, _p(_f32(gb_t)),
                               _p(np.ascontiguousarray(gb_face, np.uint8)), _p(np.ascontiguousarray(color, np.uint8)),
                               _p(np.ascontiguousarray(transversal, np.uint16)),
                               _p(np.ascontiguousarray(frames, np.uint16)), C.c_float(scale), _p(out))
    return out


def logf(x):
    x = _f32(x)
    y = np.empty_like(x)
    lib().oracle_logf_n(_p(x), _p(y), C.c_int(x.size))
    return y


def powf(x, e):
    x, e = np.broadcast_arrays(_f32(x), _f32(e))
    x, e = np.ascontiguousarray(x), np.ascontiguousarray(e)
    y = np.empty_like(x)
    lib().oracle_powf_n(_p(x), _p(e), _p(y), C.c_int(x.size))
    return y


def gi_params_from(prm) -> OGiParams:
    """Copies a product-side GiParams (same layout) into the oracle struct."""
    o = OGiParams()
    C.memmove(C.addressof(o), C.addressof(prm), C.sizeof(OGiParams))
    return o


def trace_gi(grid, df, inv_view, inv_proj, gb_t, gb_face, prm: OGiParams):
    nz, ny, nx = grid.shape
    H, W = gb_t.shape
    rgb = np.empty((H, W, 3), np.float32)
    sh = np.empty((H, W, 6), np.float32)
    luma = np.empty((H, W), np.float32)
    aosky = np.empty((H, W, 2), np.float32)
    samples = np.empty((H, W), np.int32)
    rays = np.empty((H, W), np.int32)
    lib().oracle_trace_gi(_p(grid), _p(df), nx, ny, nz, _p(_f32(inv_view)), _p(_f32(inv_proj)), W, H,
                          _p(_f32(gb_t)), _p(np.ascontiguousarray(gb_face, np.uint8)), C.byref(prm), _p(rgb), _p(sh),
                          _p(luma), _p(aosky), _p(samples), _p(rays))
    return dict(rgb=rgb, sh=sh, luma=luma, aosky=aosky, samples=samples, rays=rays)


# ---- SVGF passes (planes are float arrays of storage-rounded values) -------

def svgf_prespatial(inv_view, inv_proj, gb_t, gb_face, sh, cocg, util, ao):
    H, W = gb_t.shape
    o_sh = np.empty((H, W, 4), np.float32)
    o_cc = np.empty((H, W, 2), np.float32)
    o_ut = np.empty((H, W), np.float32)
    o_ao = np.empty((H, W, 2), np.float32)
    lib().oracle_svgf_prespatial(_p(_f32(inv_view)), _p(_f32(inv_proj)), W, H, _p(_f32(gb_t)),
                                 _p(np.ascontiguousarray(gb_face, np.uint8)), _p(_f32(sh)), _p(_f32(cocg)),
                                 _p(_f32(util)), _p(_f32(ao)), _p(o_sh), _p(o_cc), _p(o_ut), _p(o_ao))
    return dict(sh=o_sh, cocg=o_cc, util=o_ut, ao=o_ao)


def svgf_temporal(inv_view, inv_proj, prev_vp, cur_gb, prev_gb, cur, prev):
    """cur_gb / prev_gb: dicts with t, face, block.  cur: dict sh, cocg, luma, ao.
    prev: dict sh, cocg, util (H,W,3), ao."""
    H, W = cur_gb["t"].shape
    o_sh = np.empty((H, W, 4), np.float32)
    o_cc = np.empty((H, W, 2), np.float32)
    o_ut = np.empty((H, W, 3), np.float32)
    o_ao = np.empty((H, W, 2), np.float32)
    u8 = lambda a: np.ascontiguousarray(a, np.uint8)  # noqa: E731
    lib().oracle_svgf_temporal(_p(_f32(inv_view)), _p(_f32(inv_proj)), _p(_f32(prev_vp)), W, H,
                               _p(_f32(cur_gb["t"])), _p(u8(cur_gb["face"])), _p(u8(cur_gb["block"])),
                               _p(_f32(prev_gb["t"])), _p(u8(prev_gb["face"])), _p(u8(prev_gb["block"])),
                               _p(_f32(cur["sh"])), _p(_f32(cur["cocg"])), _p(_f32(cur["luma"])), _p(_f32(cur["ao"])),
                               _p(_f32(prev["sh"])), _p(_f32(prev["cocg"])), _p(_f32(prev["util"])),
                               _p(_f32(prev["ao"])), _p(o_sh), _p(o_cc), _p(o_ut), _p(o_ao))
    return dict(sh=o_sh, cocg=o_cc, util=o_ut, ao=o_ao)


def svgf_variance(inv_view, inv_proj, gb_t, gb_face, sh, cocg, util):
    H, W = gb_t.shape
    o_sh = np.empty((H, W, 4), np.float32)
    o_cc = np.empty((H, W, 2), np.float32)
    o_var = np.empty((H, W), np.float32)
    lib().oracle_svgf_variance(_p(_f32(inv_view)), _p(_f32(inv_proj)), W, H, _p(_f32(gb_t)),
                               _p(np.ascontiguousarray(gb_face, np.uint8)), _p(_f32(sh)), _p(_f32(cocg)),
                               _p(_f32(util)), _p(o_sh), _p(o_cc), _p(o_var))
    return dict(sh=o_sh, cocg=o_cc, var=o_var)


def svgf_atrous(gb_t, gb_face, sh, cocg, var, ao, temporal_util, step: int, scale: float = 1.35,
                color_phi_bias: float = 3.325, time: float = 0.0):
    H, W = gb_t.shape
    o_sh = np.empty((H, W, 4), np.float32)
    o_cc = np.empty((H, W, 2), np.float32)
    o_var = np.empty((H, W), np.float32)
    o_ao = np.empty((H, W, 2), np.float32)
    lib().oracle_svgf_atrous(W, H, _p(_f32(gb_t)), _p(np.ascontiguousarray(gb_face, np.uint8)), _p(_f32(sh)),
                             _p(_f32(cocg)), _p(_f32(var)), _p(_f32(ao)), _p(_f32(temporal_util)), C.c_int(step),
                             C.c_float(scale), C.c_float(color_phi_bias), C.c_float(time), _p(o_sh), _p(o_cc),
                             _p(o_var), _p(o_ao))
    return dict(sh=o_sh, cocg=o_cc, var=o_var, ao=o_ao)


def sh_decode(sh, cocg, face):
    H, W = face.shape
    rgb = np.empty((H, W, 3), np.float32)
    lib().oracle_sh_decode(C.c_int(H * W), _p(_f32(sh)), _p(_f32(cocg)), _p(np.ascontiguousarray(face, np.uint8)),
                           _p(rgb))
    return rgb


def raycast_detect(grid, pos, direction):
    nz, ny, nx = grid.shape
    out = np.empty(4, np.int32)
    lib().oracle_raycast_detect(_p(grid), nx, ny, nz, _p(_f32(pos)), _p(_f32(direction)), _p(out))
    return out


def dda_textbook(grid, origin, direction, max_t: float):
    nz, ny, nx = grid.shape
    out = np.empty(4, np.int32)
    o = np.ascontiguousarray(origin, np.float64)
    d = np.ascontiguousarray(direction, np.float64)
    lib().oracle_dda_textbook(_p(grid), nx, ny, nz, _p(o), _p(d), C.c_double(max_t), _p(out))
    return out


def camera(pos, yaw, pitch, fov=60.0, aspect=16 / 9, znear=0.1, zfar=1000.0):
    view = np.empty(16, np.float32)
    proj = np.empty(16, np.float32)
    iv = np.empty(16, np.float32)
    ip = np.empty(16, np.float32)
    lib().oracle_camera(_p(_f32(pos)), C.c_float(yaw), C.c_float(pitch), C.c_float(fov), C.c_float(aspect),
                        C.c_float(znear), C.c_float(zfar), _p(view), _p(proj), _p(iv), _p(ip))
    return dict(view=view, proj=proj, inv_view=iv, inv_proj=ip)


def time_plain(grid, inv_view, inv_proj, W, H, max_iter=350, reps=5):
    """CPU baseline: DF build + primary trace, median of `reps` after a warm-up.
    Returns dict(df_ms, primary_ms, samples, threads)."""
    nz, ny, nx = grid.shape
    out = np.empty(4, np.float64)
    lib().oracle_time_plain(_p(grid), nx, ny, nz, _p(_f32(inv_view)), _p(_f32(inv_proj)), W, H, max_iter, reps,
                            _p(out))
    return dict(df_ms=float(out[0]), primary_ms=float(out[1]), samples=float(out[2]), threads=int(out[3]))


def threads() -> int:
    return int(lib().oracle_threads())


# ---- planted mistakes (the oracle's own pinning tests) ----------------------

def mutation_count() -> int:
    """Number of planted-mistake ids (0 = none)."""
    return int(lib().oracle_mutation_count())


def mutation_name(mid: int) -> str:
    f = lib().oracle_mutation_name
    f.restype = C.c_char_p
    return f(C.c_int(mid)).decode()


def set_mutation(mid: int) -> None:
    """Switches on planted mistake `mid` in the SVGF and shadow-filter
    passes (0 switches them off).  For the pinning tests only."""
    lib().oracle_set_mutation(C.c_int(mid))


# ---- shared numerics (array wrappers) ---------------------------------------

def expf(x):
    x = _f32(x)
    y = np.empty_like(x)
    lib().oracle_expf_n(_p(x), _p(y), C.c_int(x.size))
    return y


def sincos_2pi(u):
    u = _f32(u)
    s = np.empty_like(u)
    c = np.empty_like(u)
    lib().oracle_sincos2pi_n(_p(u), _p(s), _p(c), C.c_int(u.size))
    return s, c


def f32_to_f16_bits(x):
    x = _f32(x)
    h = np.empty(x.shape, np.uint16)
    lib().oracle_f32_to_f16_n(_p(x), _p(h), C.c_int(x.size))
    return h


def f16_bits_to_f32(h):
    h = np.ascontiguousarray(h, np.uint16)
    x = np.empty(h.shape, np.float32)
    lib().oracle_f16_to_f32_n(_p(h), _p(x), C.c_int(h.size))
    return x


def pcg_hash(v):
    v = np.ascontiguousarray(v, np.uint32)
    h = np.empty_like(v)
    lib().oracle_pcg_n(_p(v), _p(h), C.c_int(v.size))
    return h


def df_step_table():
    out = np.empty(256, np.int32)
    lib().oracle_df_step_table(_p(out))
    return out


def unorm8_encode(x):
    x = _f32(x)
    c = np.empty(x.shape, np.uint8)
    lib().oracle_unorm8_n(_p(x), _p(c), C.c_int(x.size))
    return c


def powi(x, e: int):
    x = _f32(x)
    y = np.empty_like(x)
    lib().oracle_powi_n(_p(x), C.c_int(e), _p(y), C.c_int(x.size))
    return y


def unorm8_decode_table():
    out = np.empty(256, np.float32)
    lib().oracle_unorm8_decode_table(_p(out))
    return out
