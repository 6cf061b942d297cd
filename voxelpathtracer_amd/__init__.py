"""MI355X-native (gfx950) hot path of a distance-field voxel path tracer:
Manhattan distance-field build, DF-accelerated voxel traversal for primary,
sun-shadow and diffuse-GI rays, and the SVGF denoiser, behind a C ABI
(include/vpt/vpt.h) implemented by hand-written HIP kernels in libvpt.so.
"""
from ._lib import LIB_PATH, VptError, lib  # noqa: F401
from .world import (Context, DeviceGBuffer, World, composite, composite_params, generate_terrain, gi_params,  # noqa: F401
                    make_camera, reflect_shading)
from .pipeline import SVGF, Graph, Pipeline, ShadowDenoiser, decode_texels, svgf_params_default, viewproj  # noqa: F401
from .batch import HIT_DTYPE, QRAY_DTYPE, Batch, Frame, make_qrays  # noqa: F401
from .server import Server  # noqa: F401

__version__ = "0.1.0"
