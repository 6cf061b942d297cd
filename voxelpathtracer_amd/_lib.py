"""ctypes binding of libvpt.so (include/vpt/vpt.h).

The library is loaded from ``voxelpathtracer_amd/_build/libvpt.so``; there
is no fallback implementation: if the library is missing, every entry point
raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("VPT_LIB", _PKG / "_build" / "libvpt.so"))

VPT_OK = 0
VPT_EINVAL = -1
VPT_ENOMEM = -2
VPT_EDEVICE = -3
VPT_ESTATE = -4
VPT_ETIMEOUT = -5
VPT_EBUSY = -6

VPT_SERVE_PLAIN = 0
VPT_SERVE_BATCHED = 1
VPT_REQ_PRIMARY = 0
VPT_REQ_SHADOW = 1
VPT_REQ_GI = 2
VPT_REQ_FRAME = 3


class VptError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"vpt error {code}: {msg}")
        self.code = code


class Camera(C.Structure):
    _fields_ = [("inv_view", C.c_float * 16), ("inv_proj", C.c_float * 16), ("width", C.c_int32),
                ("height", C.c_int32)]


class GBuffer(C.Structure):
    _fields_ = [("d_t", C.c_void_p), ("d_face", C.c_void_p), ("d_block", C.c_void_p), ("d_voxel", C.c_void_p),
                ("width", C.c_int32), ("height", C.c_int32)]


class GiParams(C.Structure):
    _fields_ = [("sun_dir", C.c_float * 3), ("light_color", C.c_float * 3), ("sky_radiance", C.c_float * 3),
                ("sky_pixel_radiance", C.c_float * 3), ("albedo", C.c_float * 384), ("frame", C.c_uint32),
                ("gi_iter", C.c_int32), ("shadow_iter", C.c_int32)]


class SvgfParams(C.Structure):
    _fields_ = [("color_phi_bias", C.c_float), ("resolution_scale", C.c_float), ("pre_spatial", C.c_int32),
                ("steps", C.c_int32 * 5)]


class ReflectParams(C.Structure):
    _fields_ = [("light_dir", C.c_float * 3), ("div", C.c_int32), ("max_iter", C.c_int32), ("shadow_iter", C.c_int32),
                ("spp", C.c_int32), ("rough", C.c_int32), ("roughness", C.c_float), ("frame", C.c_uint32)]


class ReflectShading(C.Structure):
    _fields_ = [("d_diffuse", C.c_void_p), ("d_shadow", C.c_void_p), ("viewproj", C.c_float * 16),
                ("sun_radiance", C.c_float * 3), ("sky_radiance", C.c_float * 3), ("albedo", C.c_float * 384)]


class CompositeParams(C.Structure):
    _fields_ = [("sun_dir", C.c_float * 3), ("moon_dir", C.c_float * 3), ("sun_radiance", C.c_float * 3),
                ("moon_radiance", C.c_float * 3), ("sky_pixel_radiance", C.c_float * 3), ("roughness", C.c_float),
                ("albedo", C.c_float * 384)]


class McBlock(C.Structure):
    _fields_ = [("name", C.c_char_p), ("id", C.c_uint8), ("data", C.c_uint8)]


class Ray(C.Structure):
    _fields_ = [("origin", C.c_float * 3), ("dir", C.c_float * 3)]


class Hit(C.Structure):
    _fields_ = [("t", C.c_float), ("voxel", C.c_uint32), ("face", C.c_uint8), ("block", C.c_uint8),
                ("samples", C.c_uint16)]


class QRay(C.Structure):
    _fields_ = [("origin", C.c_float * 3), ("dir", C.c_float * 3), ("meta", C.c_uint32), ("dst", C.c_uint32)]


class LaunchCfg(C.Structure):
    _fields_ = [("block", C.c_int32), ("waves_per_cu", C.c_int32), ("chunk", C.c_int32), ("refill", C.c_int32),
                ("flags", C.c_int32)]


class ServerCfg(C.Structure):
    _fields_ = [("mode", C.c_int32), ("max_worlds", C.c_int32), ("arena_bytes", C.c_uint64),
                ("world_budget", C.c_uint64), ("max_launch_rays", C.c_uint32), ("max_inflight", C.c_int32)]


class Request(C.Structure):
    _fields_ = [("world", C.c_int32), ("kind", C.c_int32), ("n", C.c_uint32), ("max_iter", C.c_int32),
                ("d_rays", C.c_void_p), ("sun_dir", C.c_float * 3), ("shadow_iter", C.c_int32),
                ("cam", C.POINTER(Camera)), ("gi", C.POINTER(GiParams)), ("primary_iter", C.c_int32),
                ("frame_shadow_iter", C.c_int32), ("arrival_ns", C.c_int64)]


class RequestResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("kind", C.c_int32), ("n", C.c_uint32), ("d_hits", C.c_void_p),
                ("d_vis", C.c_void_p), ("gbuffer", GBuffer), ("d_mask", C.c_void_p), ("d_rgb", C.c_void_p),
                ("t_arrival_ns", C.c_int64), ("t_admit_ns", C.c_int64), ("t_start_ns", C.c_int64),
                ("t_done_ns", C.c_int64)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("rays", C.c_uint64),
                ("launches", C.c_uint64), ("frames", C.c_uint64), ("arena_used", C.c_uint64),
                ("arena_peak", C.c_uint64), ("arena_bytes", C.c_uint64), ("world_bytes", C.c_uint64),
                ("pending", C.c_uint32), ("active", C.c_uint32)]


class GraphStats(C.Structure):
    _fields_ = [("launches", C.c_uint64), ("instantiations", C.c_uint64), ("updates", C.c_uint64),
                ("nodes", C.c_uint32), ("pad", C.c_uint32)]


_vp = C.c_void_p
_i = C.c_int
_f = C.c_float
_u32 = C.c_uint32
_PROTOS = {
    "vpt_last_error": (C.c_char_p, []),
    "vpt_version": (C.c_char_p, []),
    "vpt_check_log_division": (C.c_int, [C.POINTER(C.c_ulonglong), _vp]),
    "vpt_ctx_create": (_i, [_i, C.POINTER(_vp)]),
    "vpt_ctx_destroy": (_i, [_vp]),
    "vpt_ctx_device": (_i, [_vp, C.POINTER(_i)]),
    "vpt_world_create": (_i, [_vp, _i, _i, _i, C.POINTER(_vp)]),
    "vpt_world_destroy": (_i, [_vp]),
    "vpt_world_dims": (_i, [_vp, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i)]),
    "vpt_world_bytes": (_i, [_vp, C.POINTER(C.c_size_t)]),
    "vpt_world_upload": (_i, [_vp, _vp, _vp]),
    "vpt_world_upload_device": (_i, [_vp, _vp, _vp]),
    "vpt_world_set_blocks": (_i, [_vp, _vp, _vp, _i, _vp]),
    "vpt_df_build": (_i, [_vp, _vp]),
    "vpt_df_update": (_i, [_vp, _vp, _i, _vp, _vp]),
    "vpt_world_save": (_i, [_vp, C.c_char_p]),
    "vpt_world_load": (_i, [_vp, C.c_char_p, _vp]),
    "vpt_mc_import": (_i, [C.c_char_p, _i, _i, _i, C.POINTER(_f), _vp, _vp, _i, _vp, _vp]),
    "vpt_df_download": (_i, [_vp, _vp, _vp]),
    "vpt_grid_download": (_i, [_vp, _vp, _vp]),
    "vpt_trace_primary": (_i, [_vp, C.POINTER(Camera), _i, C.POINTER(GBuffer), _vp]),
    "vpt_trace_shadow": (_i, [_vp, C.POINTER(Camera), C.POINTER(GBuffer), C.POINTER(_f), _i, _vp, _vp]),
    "vpt_trace_shadow2": (_i, [_vp, C.POINTER(Camera), C.POINTER(GBuffer), C.POINTER(_f), _i, _vp, _vp, _vp]),
    "vpt_trace_reflect": (_i, [_vp, C.POINTER(Camera), C.POINTER(GBuffer), C.POINTER(ReflectParams), _vp, _vp, _vp]),
    "vpt_reflect_params_default": (_i, [C.POINTER(_f), C.c_uint32, C.POINTER(ReflectParams)]),
    "vpt_reflect_shading_from_gi": (_i, [C.POINTER(GiParams), C.POINTER(_f), _vp, _vp, C.POINTER(ReflectShading)]),
    "vpt_trace_reflect_shaded": (_i, [_vp, C.POINTER(Camera), C.POINTER(GBuffer), C.POINTER(ReflectParams),
                                      C.POINTER(ReflectShading), _vp, _vp, _vp, _vp]),
    "vpt_trace_reflect_deferred": (_i, [_vp, C.POINTER(Camera), C.POINTER(GBuffer), C.POINTER(ReflectParams),
                                        C.POINTER(_f), _vp, _vp, _vp, _vp]),
    "vpt_shade_reflect": (_i, [_vp, C.POINTER(Camera), C.POINTER(GBuffer), C.POINTER(ReflectParams),
                               C.POINTER(ReflectShading), _vp, _vp, _vp, _vp]),
    "vpt_composite_params_from_gi": (_i, [C.POINTER(GiParams), C.POINTER(CompositeParams)]),
    "vpt_composite": (_i, [_vp, C.POINTER(Camera), C.POINTER(GBuffer), C.POINTER(CompositeParams), _vp, _vp, _vp, _i,
                           _vp, _vp]),
    "vpt_trace_gi": (_i, [_vp, C.POINTER(Camera), C.POINTER(GBuffer), C.POINTER(GiParams), _vp, _vp, _vp]),
    "vpt_trace_gi_noalloc": (_i, [_vp, C.POINTER(Camera), C.POINTER(GBuffer), C.POINTER(GiParams), _vp, _vp,
                                  _vp]),
    "vpt_trace_frame": (_i, [_vp, C.POINTER(Camera), C.POINTER(GiParams), _i, _i, C.POINTER(GBuffer), _vp, _vp, _vp,
                             _vp, _vp]),
    "vpt_graph_create": (_i, [_vp, C.POINTER(_vp)]),
    "vpt_graph_destroy": (_i, [_vp]),
    "vpt_graph_begin": (_i, [_vp, _vp, C.POINTER(_vp)]),
    "vpt_graph_launch": (_i, [_vp, _vp]),
    "vpt_graph_abort": (_i, [_vp]),
    "vpt_graph_stats_get": (_i, [_vp, C.POINTER(GraphStats)]),
    "vpt_trace_rays": (_i, [_vp, _vp, _i, _i, _vp, _vp]),
    "vpt_trace_gi_rays": (_i, [_vp, _vp, _i, _i, C.POINTER(_f), _i, _vp, _vp, _vp]),
    "vpt_world_set_sample_counter": (_i, [_vp, _vp]),
    "vpt_batch_create": (_i, [_vp, _i, C.POINTER(_vp)]),
    "vpt_batch_destroy": (_i, [_vp]),
    "vpt_batch_set_world": (_i, [_vp, _i, _vp]),
    "vpt_batch_get_cfg": (_i, [_vp, C.POINTER(LaunchCfg)]),
    "vpt_batch_set_cfg": (_i, [_vp, C.POINTER(LaunchCfg)]),
    "vpt_batch_trace": (_i, [_vp, _vp, _i, _vp, _vp]),
    "vpt_batch_autotune": (_i, [_vp, _vp, _i, _vp, C.POINTER(_f), _vp]),
    "vpt_frame_create": (_i, [_vp, _i, _i, C.POINTER(_vp)]),
    "vpt_frame_destroy": (_i, [_vp]),
    "vpt_frame_set_cfg": (_i, [_vp, C.POINTER(LaunchCfg)]),
    "vpt_frame_get_cfg": (_i, [_vp, C.POINTER(LaunchCfg)]),
    "vpt_frame_trace": (_i, [_vp, _vp, C.POINTER(Camera), C.POINTER(GiParams), _i, _i, C.POINTER(GBuffer), _vp, _vp,
                             _vp, _i, _vp]),
    "vpt_frame_autotune": (_i, [_vp, _vp, C.POINTER(Camera), C.POINTER(GiParams), _i, _i, C.POINTER(GBuffer), _vp,
                                _vp, _vp, C.POINTER(_f), _vp]),
    "vpt_now_ns": (C.c_int64, []),
    "vpt_server_create": (_i, [_vp, C.POINTER(ServerCfg), C.POINTER(_vp)]),
    "vpt_server_destroy": (_i, [_vp]),
    "vpt_server_add_world": (_i, [_vp, _vp, C.POINTER(_i)]),
    "vpt_server_remove_world": (_i, [_vp, _i]),
    "vpt_server_set_launch_cfg": (_i, [_vp, C.POINTER(LaunchCfg)]),
    "vpt_server_submit": (_i, [_vp, C.POINTER(Request), C.POINTER(C.c_uint64)]),
    "vpt_server_wait": (_i, [_vp, C.c_uint64, _i, C.POINTER(RequestResult)]),
    "vpt_server_release": (_i, [_vp, C.c_uint64]),
    "vpt_server_stats_get": (_i, [_vp, C.POINTER(ServerStats)]),
    "vpt_unet_create": (_i, [_vp, _i, _i, _i, C.POINTER(_vp)]),
    "vpt_unet_destroy": (_i, [_vp]),
    "vpt_unet_param_count": (_i, [_vp, C.POINTER(C.c_size_t)]),
    "vpt_unet_init": (_i, [_vp, _u32]),
    "vpt_unet_set_params": (_i, [_vp, _vp, C.c_size_t]),
    "vpt_unet_get_params": (_i, [_vp, _vp, C.c_size_t]),
    "vpt_unet_get_grads": (_i, [_vp, _vp, C.c_size_t]),
    "vpt_unet_forward": (_i, [_vp, _vp, _i, _i, _i, _i, _vp, _vp]),
    "vpt_unet_backward": (_i, [_vp, _vp, _vp, _i, _i, _i, C.POINTER(_f), _vp]),
    "vpt_unet_adam_step": (_i, [_vp, _f, _vp]),
    "vpt_unet_assemble": (_i, [_vp, C.POINTER(GBuffer), _vp, _vp, _vp]),
    "vpt_unet_denoise": (_i, [_vp, C.POINTER(GBuffer), _vp, _vp, _vp]),
    "vpt_unet_read_activation": (_i, [_vp, _i, _i, _i, _i, _vp]),
    "vpt_texel_accumulate": (_i, [_vp, _i, _vp, _i, _vp]),
    "vpt_irradiance_decode": (_i, [C.POINTER(GBuffer), _vp, _vp, _f, _vp, _vp]),
    "vpt_svgf_create": (_i, [_vp, _i, _i, C.POINTER(SvgfParams), C.POINTER(_vp)]),
    "vpt_svgf_destroy": (_i, [_vp]),
    "vpt_svgf_reset": (_i, [_vp, _vp]),
    "vpt_svgf_frame": (_i, [_vp, C.POINTER(Camera), C.POINTER(_f), C.POINTER(GBuffer), C.POINTER(GBuffer), _vp, _f,
                            _vp, _vp, _vp]),
    "vpt_svgf_read_plane": (_i, [_vp, _i, _vp, _vp]),
    "vpt_svgf_final": (_i, [_vp, C.POINTER(_vp)]),
    "vpt_svgf_read_utility": (_i, [_vp, _vp, _vp]),
    "vpt_shadow_denoise_create": (_i, [_vp, _i, _i, _vp]),
    "vpt_shadow_denoise_destroy": (_i, [_vp]),
    "vpt_shadow_denoise_reset": (_i, [_vp, _vp]),
    "vpt_shadow_denoise_frame": (_i, [_vp, C.POINTER(Camera), C.POINTER(_f), C.POINTER(GBuffer), C.POINTER(GBuffer),
                                      _vp, _vp, _f, _vp, _vp]),
    "vpt_shadow_denoise_read": (_i, [_vp, _i, _vp, _vp]),
    "vpt_generate_terrain": (_i, [_i, _i, _i, _u32, _u32, _i, _i, _i, _vp]),
    "vpt_camera_make": (_i, [C.POINTER(_f), _f, _f, _f, _i, _i, _f, _f, C.POINTER(Camera), C.POINTER(_f),
                             C.POINTER(_f)]),
    "vpt_gi_params_default": (_i, [_f, _u32, _u32, C.POINTER(GiParams)]),
    "vpt_svgf_params_default": (_i, [C.POINTER(SvgfParams)]),
    "vpt_raycast_detect": (_i, [_vp, _i, _i, _i, C.POINTER(_f), C.POINTER(_f), C.POINTER(C.c_int32)]),
}

_lib = None


def lib() -> C.CDLL:
    """Loads libvpt.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} not found: build it with `python -m voxelpathtracer_amd.build`")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _PROTOS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols() -> list[str]:
    return list(_PROTOS)


def check(rc: int) -> None:
    if rc != VPT_OK:
        raise VptError(rc, lib().vpt_last_error().decode(errors="replace"))


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))
