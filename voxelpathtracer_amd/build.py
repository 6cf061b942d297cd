"""Builds libvpt.so (the gfx950 kernels + C ABI) in-tree with hipcc.

Run as ``python -m voxelpathtracer_amd.build`` or through
``__graft_entry__.build()``.  The library lands in
``voxelpathtracer_amd/_build/libvpt.so``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OUT = PKG / "_build"
LIB = OUT / "libvpt.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: no a*b+c fusion, so every expression rounds exactly as
# written (bit-exact hits against the CPU oracle).  Packed fp32 ops are
# disabled for correctness on this hardware.  Division and sqrt stay
# correctly rounded.
COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Wall",
    "-Wno-unused-function",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]
DEVICE = [
    "--offload-arch=gfx950",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]

SOURCES = ["vpt_host.cpp", "vpt_mca.cpp", "vpt_world.hip", "vpt_trace.hip", "vpt_batch.hip", "vpt_svgf.hip", "vpt_server.hip",
           "vpt_unet.hip", "vpt_graph.hip", "vpt_composite.hip"]


def _compile(src: str) -> Path:
    obj = OUT / (Path(src).stem + ".o")
    cmd = [HIPCC, *COMMON]
    if src.endswith(".hip"):
        cmd += [*DEVICE, "-x", "hip"]
    else:
        cmd += ["-x", "c++"]
    cmd += ["-c", str(CSRC / src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = True) -> Path:
    OUT.mkdir(parents=True, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(4, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT / "libvpt.so.tmp"
    cmd = [HIPCC, *DEVICE, "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-lz"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_demo(verbose)
    return LIB


DEMO = OUT / "vpt_frame_demo"
SERVE_BENCH = OUT / "vpt_serve_bench"


def _build_example(src: str, out: Path, verbose: bool) -> Path:
    cmd = [HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", "-pthread", f"-I{ROOT / 'include'}",
           str(CSRC / "examples" / src), f"-L{OUT}", "-lvpt", "-Wl,-rpath,$ORIGIN", "-o", str(out)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"example build failed:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    if verbose:
        print(f"built {out}")
    return out


def build_demo(verbose: bool = True) -> Path:
    """The C++ host examples: the frame loop over the World / Pipeline
    façades, and the serving benchmark over vpt_server."""
    _build_example("vpt_serve_bench.cpp", SERVE_BENCH, verbose)
    return _build_example("vpt_frame_demo.cpp", DEMO, verbose)


def build_oracle(verbose: bool = True) -> Path:
    """Builds the CPU oracle (test infrastructure) with its Makefile."""
    make = shutil.which("make") or "make"
    res = subprocess.run([make, "-s", "-C", str(ROOT / "oracle")], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stdout}\n{res.stderr}")
    lib = ROOT / "oracle" / "_build" / "libvpt_oracle.so"
    if verbose:
        print(f"built {lib}")
    # the reference's own Minecraft region reader, where the reference tree
    # is present (the importer's checker; the GPU box only runs GPU tests)
    if Path("/root/reference/Dependencies/enkiMI/enkimi.c").exists():
        res = subprocess.run([make, "-s", "-C", str(ROOT / "oracle"), "ref"], capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"oracle/_ref build failed:\n{res.stdout}\n{res.stderr}")
        if verbose:
            print(f"built {ROOT / 'oracle' / '_ref' / 'libmca_ref.so'}")
    return lib


if __name__ == "__main__":
    build()
    if "--oracle" in sys.argv:
        build_oracle()
