"""Shared test helpers: workloads, storage rounding and the oracle SVGF chain."""
from __future__ import annotations

import numpy as np

import oracle
from voxelpathtracer_amd import decode_texels, generate_terrain, make_camera, viewproj

CAM_POS = (192.5, 100.5, 192.5)
CAM_YAW = 45.0
CAM_PITCH = -20.0


def f16(a) -> np.ndarray:
    """Round float32 values through binary16 (RNE) like a 16F render target."""
    return np.asarray(a, np.float32).astype(np.float16).astype(np.float32)


def u8n(a) -> np.ndarray:
    """Round through unorm8 (round(clamp(x,0,1)*255)/255)."""
    a = np.clip(np.asarray(a, np.float32), 0, 1)
    return (np.floor(a * np.float32(255.0) + np.float32(0.5)).astype(np.uint8).astype(np.float32) / np.float32(255.0))


def random_world(rng, nx, ny, nz, density):
    g = (rng.random((nz, ny, nx)) < density).astype(np.uint8)
    g *= rng.integers(1, 128, size=g.shape).astype(np.uint8)
    return g


def terrain(nx=384, ny=128, nz=384, **kw):
    return generate_terrain(nx, ny, nz, **kw)


def camera(W, H, pos=CAM_POS, yaw=CAM_YAW, pitch=CAM_PITCH):
    cam, view, proj = make_camera(pos, yaw, pitch, W, H)
    iv = np.array(cam.inv_view, np.float32)
    ip = np.array(cam.inv_proj, np.float32)
    return cam, view, proj, iv, ip


def svgf_camera_path(frame: int):
    """The SVGF workload's camera: +0.1 voxel/frame along x, +0.2 deg yaw/frame."""
    x, y, z = CAM_POS
    return (x + 0.1 * frame, y, z), CAM_YAW + 0.2 * frame, CAM_PITCH


def gi_planes_from_oracle(gi: dict) -> dict:
    """Raw GI FBO values as stored (RGBA16F / RG16F / R16F / RG8)."""
    return dict(sh=f16(gi["sh"][..., 0:4]), cocg=f16(gi["sh"][..., 4:6]), luma=f16(gi["luma"]), ao=u8n(gi["aosky"]))


class OracleSVGF:
    """The SVGF chain of Pipeline.cpp:2284-2611 on the oracle passes."""

    def __init__(self, W, H, steps=(16, 8, 4, 2, 1), scale=1.35, phi_bias=3.325, pre_spatial=True):
        self.W, self.H = W, H
        self.steps, self.scale, self.phi_bias, self.pre_spatial = steps, scale, phi_bias, pre_spatial
        z = np.zeros
        self.hist = dict(sh=z((H, W, 4), np.float32), cocg=z((H, W, 2), np.float32), util=z((H, W, 3), np.float32),
                         ao=z((H, W, 2), np.float32))

    def frame(self, iv, ip, prev_vp, cur_gb, prev_gb, raw: dict, time: float) -> dict:
        out = {}
        if self.pre_spatial:
            pre = oracle.svgf_prespatial(iv, ip, cur_gb["t"], cur_gb["face"], raw["sh"], raw["cocg"], raw["luma"],
                                         raw["ao"])
            cur = dict(sh=pre["sh"], cocg=pre["cocg"], luma=pre["util"], ao=pre["ao"])
            out["pre"] = pre
        else:
            cur = raw
        tmp = oracle.svgf_temporal(iv, ip, prev_vp, cur_gb, prev_gb, cur, self.hist)
        out["temporal"] = tmp
        self.hist = tmp
        var = oracle.svgf_variance(iv, ip, cur_gb["t"], cur_gb["face"], tmp["sh"], tmp["cocg"], tmp["util"])
        out["variance"] = var
        src = dict(sh=var["sh"], cocg=var["cocg"], var=var["var"], ao=tmp["ao"])
        for step in self.steps:
            src = oracle.svgf_atrous(cur_gb["t"], cur_gb["face"], src["sh"], src["cocg"], src["var"], src["ao"],
                                     tmp["util"], step, self.scale, self.phi_bias, time)
        out["final"] = src
        out["rgb"] = oracle.sh_decode(src["sh"], src["cocg"], cur_gb["face"])
        return out


def texels_to_planes(tex) -> dict:
    d = decode_texels(tex)
    return d


def rel_l1(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-30))


__all__ = ["f16", "u8n", "random_world", "terrain", "camera", "svgf_camera_path", "gi_planes_from_oracle",
           "OracleSVGF", "texels_to_planes", "rel_l1", "viewproj"]
