"""Writes tests/golden/small_world.npz: a small fixed workload and the
oracle's outputs for it (DF, G-buffer, shadow mask, GI, two SVGF frames).

The reference ships no golden vectors (it has no tests), so these fixtures
are produced by this repository's oracle; they pin the oracle and the GPU
path against drift.  Regenerate with
    python tests/golden/make_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from tests.helpers import OracleSVGF, camera, gi_planes_from_oracle, viewproj  # noqa: E402
from voxelpathtracer_amd import generate_terrain, gi_params  # noqa: E402

OUT = Path(__file__).resolve().parent / "small_world.npz"
NX, NY, NZ = 64, 48, 64
W, H = 64, 36
POS0 = (32.5, 40.5, 32.5)


def frame_camera(f):
    return camera(W, H, pos=(POS0[0] + 0.1 * f, POS0[1], POS0[2]), yaw=30.0 + 0.2 * f, pitch=-30.0)


def make() -> dict:
    grid = generate_terrain(NX, NY, NZ, seed=0x5EED0101, seed_random=0x5EED0102, h_min=8, h_max=30)
    df = oracle.df_build(grid)
    out = dict(grid=grid, df=df)
    orc = OracleSVGF(W, H)
    prev_gb = None
    prev_vm = None
    for f in range(2):
        cam, view, proj, iv, ip = frame_camera(f)
        p = oracle.trace_primary(grid, df, iv, ip, W, H, 350)
        prm = gi_params(50.0, f)
        s = oracle.trace_shadow(grid, df, iv, ip, p["t"], p["face"], np.array(prm.sun_dir, np.float32), 350)
        g = oracle.trace_gi(grid, df, iv, ip, p["t"], p["face"], oracle.gi_params_from(prm))
        raw = gi_planes_from_oracle(g)
        vp = viewproj(*(prev_vm if prev_vm else (view, proj)))
        sv = orc.frame(iv, ip, vp, p, prev_gb if prev_gb else p, raw, f / 60.0)
        for k in ("t", "face", "block", "voxel"):
            out[f"f{f}_{k}"] = p[k]
        out[f"f{f}_shadow"] = s["mask"]
        out[f"f{f}_gi_rgb"] = g["rgb"]
        out[f"f{f}_gi_sh"] = raw["sh"]
        out[f"f{f}_gi_cocg"] = raw["cocg"]
        out[f"f{f}_gi_luma"] = raw["luma"]
        out[f"f{f}_gi_ao"] = raw["ao"]
        out[f"f{f}_svgf_rgb"] = sv["rgb"]
        out[f"f{f}_svgf_sh"] = sv["final"]["sh"]
        out[f"f{f}_temporal_util"] = sv["temporal"]["util"]
        prev_gb = p
        prev_vm = (view, proj)
    return out


if __name__ == "__main__":
    data = make()
    np.savez_compressed(OUT, **data)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes, {len(data)} arrays)")
