"""Writes the Minecraft-import fixtures tests/golden/mca_<case>.npz: the
synthetic region files of each case (mca_synth.build_case, seed 1) and the
block grid the reference's own reader and import loop produce from them
(oracle/_ref/libmca_ref.so, built by `make -C oracle ref` from the
reference's Dependencies/enkiMI), with the sections it imported.  It also
writes tests/golden/mca_medival.json: for the reference's own test world
(`Test MC Worlds/Medival`, seven region files, stated origin 136 0 -94) the
sections that reader walks and the SHA-256 and solid count of the grid it
imports at the stated origin and at origins whose 384x128x384 window covers
each region's saved chunks.  Run from the repository root where the
reference tree is present:

    python tests/golden/make_mca_golden.py
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import os
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
import mca_synth as M  # noqa: E402

REF = HERE.parent.parent / "oracle" / "_ref" / "libmca_ref.so"
MEDIVAL = Path("/root/reference/Test MC Worlds/Medival")
# the stated origin (Origin.txt), which lands on no saved chunk (the files
# hold a ring of chunks around the build), then one window per region file
MEDIVAL_ORIGINS = [(136, 0, -94), (-300, 40, -560), (-560, 40, -300), (-540, 40, 150), (192, 40, -560),
                   (550, 40, -300), (550, 40, -560), (550, 40, 150)]
MEDIVAL_LUTS = {"identity": lambda: np.arange(256, dtype=np.uint8), "lut5": lambda: M.block_lut(5)}


def reference_import(ref, files: dict, dims, origin, lut):
    """The reference's ImportWorld over `files` (name order): grid, sections."""
    nx, ny, nz = dims
    grid = np.zeros((nz, ny, nx), np.uint8)
    n = C.c_int32(0)
    org = (C.c_int32 * 3)(*[int(v) for v in origin])
    with tempfile.TemporaryDirectory() as d:
        for name in sorted(files):
            p = os.path.join(d, name)
            Path(p).write_bytes(files[name])
            rc = ref.mca_ref_import_file(p.encode(), nx, ny, nz, org, lut.ctypes.data_as(C.c_void_p),
                                         grid.ctypes.data_as(C.c_void_p), C.byref(n))
            assert rc == 0, p
    return grid, n.value


def reference_names(ref) -> dict:
    ref.mca_ref_table_name.restype = C.c_char_p
    table = {}
    for i in range(ref.mca_ref_table_size()):
        table.setdefault(ref.mca_ref_table_name(i).decode(), ref.mca_ref_table_value(i))
    return table


def medival_cases(ref) -> list:
    files = {p.name: p.read_bytes() for p in sorted(MEDIVAL.glob("*.mca"))}
    out = []
    for origin in MEDIVAL_ORIGINS:
        for lname, lut in MEDIVAL_LUTS.items():
            grid, sections = reference_import(ref, files, (384, 128, 384), origin, lut())
            out.append(dict(origin=list(origin), lut=lname, sections=sections, solid=int((grid > 0).sum()),
                            sha256=hashlib.sha256(grid.tobytes()).hexdigest()))
    return out


def main():
    ref = C.CDLL(str(REF))
    table = reference_names(ref)
    for name, (bid, dv) in M.NAMES.items():     # the cases' names agree with the reader's table
        assert table.get(name) == bid | (dv << 8), name
    assert "test:unknown" not in table
    lut = M.block_lut()
    for case in M.CASES:
        files = M.build_case(case, 1)
        grid, sections = reference_import(ref, files, M.DIMS, M.ORIGIN, lut)
        arrays = {f"file:{k}": np.frombuffer(v, np.uint8) for k, v in files.items()}
        np.savez_compressed(HERE / f"mca_{case}.npz", expected=grid, sections=np.int32(sections), **arrays)
        print(case, "sections", sections, "solid", int((grid > 0).sum()))
    cases = medival_cases(ref)
    (HERE / "mca_medival.json").write_text(json.dumps(dict(files=sorted(p.name for p in MEDIVAL.glob("*.mca")),
                                                           cases=cases), indent=1) + "\n")
    print("medival", [(c["origin"], c["lut"], c["solid"]) for c in cases])


if __name__ == "__main__":
    main()
