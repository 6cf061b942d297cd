"""Writes the Minecraft-import fixtures tests/golden/mca_<case>.npz: the
synthetic region files of each case (mca_synth.build_case, seed 1) and the
block grid the reference's own reader and import loop produce from them
(oracle/_ref/libmca_ref.so, built by `make -C oracle ref` from the
reference's Dependencies/enkiMI), with the sections it imported.  Run from
the repository root where the reference tree is present:

    python tests/golden/make_mca_golden.py
"""
from __future__ import annotations

import ctypes as C
import os
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
import mca_synth as M  # noqa: E402

REF = HERE.parent.parent / "oracle" / "_ref" / "libmca_ref.so"


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


if __name__ == "__main__":
    main()
