"""Minecraft world import (vpt_mc_import, host only) against the reference's
own reader: the committed fixtures hold the grids the reference's enkiMI
reader and import loop (Core/NBT/Importer.cpp:68-168) produce from
synthetic region files in every chunk layout; where the reference tree is
present, fresh cases are compared with that reader live.  A round trip of a
known grid through legacy region files pins the import independently."""
from __future__ import annotations

import ctypes as C
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from tests.golden import mca_synth as M
from voxelpathtracer_amd._lib import VptError
from voxelpathtracer_amd.world import mc_import

GOLDEN = Path(__file__).resolve().parent / "golden"
REF_LIB = Path(__file__).resolve().parent.parent / "oracle" / "_ref" / "libmca_ref.so"
MEDIVAL = Path("/root/reference/Test MC Worlds/Medival")


def _write(tmp_path: Path, files: dict) -> Path:
    d = tmp_path / "region"
    d.mkdir(exist_ok=True)
    for name, data in files.items():
        (d / name).write_bytes(bytes(data))
    return d


@pytest.mark.parametrize("case", M.CASES)
def test_fixture_matches_reference(tmp_path, case):
    z = np.load(GOLDEN / f"mca_{case}.npz")
    files = {k[5:]: z[k].tobytes() for k in z.files if k.startswith("file:")}
    grid, stats = mc_import(_write(tmp_path, files), M.DIMS, M.ORIGIN, M.block_lut(), M.NAMES)
    np.testing.assert_array_equal(grid, z["expected"], err_msg=case)
    assert stats[0] == len(files) and stats[2] == int(z["sections"])
    assert (grid > 0).sum() > 1000


@pytest.mark.skipif(not REF_LIB.exists(), reason="the reference's reader is built only where the reference is present")
@pytest.mark.parametrize("seed", [2, 3, 4, 5])
def test_live_against_reference(tmp_path, seed):
    from tests.golden.make_mca_golden import reference_import
    ref = C.CDLL(str(REF_LIB))
    rng = np.random.default_rng(seed)
    origin = tuple(rng.uniform(-20, 20, 3))
    lut = M.block_lut(seed)
    for case in M.CASES:
        files = M.build_case(case, 100 + seed)
        want, sections = reference_import(ref, files, M.DIMS, origin, lut)
        (tmp_path / case).mkdir()
        d = _write(tmp_path / case, files)
        got, stats = mc_import(d, M.DIMS, origin, lut, M.NAMES)
        np.testing.assert_array_equal(got, want, err_msg=f"{case} seed {seed}")
        assert stats[2] == sections


def test_legacy_round_trip(tmp_path):
    """A known grid written as pre-1.13 chunks (data values 0) imports back
    unchanged with the identity id map: chunk (cx, cz) section Y holds world
    cells x = 16 cx + i, y = 16 Y + j, z = 16 cz + k, and the world centres
    x and z, so origin (nx/2 - x0, ..) cancels the centring."""
    nx, ny, nz = 32, 32, 48
    rng = np.random.default_rng(7)
    want = np.where(rng.random((nz, ny, nx)) < 0.3, rng.integers(1, 256, (nz, ny, nx)), 0).astype(np.uint8)
    chunks = {}
    for cz in range(nz // 16):
        for cx in range(nx // 16):
            secs = []
            for sy in range(ny // 16):
                block = want[16 * cz:16 * cz + 16, 16 * sy:16 * sy + 16, 16 * cx:16 * cx + 16]   # (z, y, x)
                blocks = np.ascontiguousarray(block.transpose(1, 0, 2)).reshape(-1)          # y*256 + z*16 + x
                secs.append(M.compound(("Y", M.Tag(M.BYTE, sy)), ("Blocks", M.Tag(M.BYTE_ARRAY, blocks)),
                                       ("Data", M.Tag(M.BYTE_ARRAY, np.zeros(2048, np.uint8)))))
            level = M.compound(("xPos", M.Tag(M.INT, cx)), ("zPos", M.Tag(M.INT, cz)),
                               ("Sections", M.lst(M.COMPOUND, secs)))
            chunks[cz * 32 + cx] = M.encode_root(M.compound(("Level", level)))
    d = _write(tmp_path, {"r.0.0.mca": M.region(chunks)})
    got, stats = mc_import(d, (nx, ny, nz), (nx // 2, 0, nz // 2), None, None)
    np.testing.assert_array_equal(got, want)
    assert stats == [1, len(chunks), len(chunks) * (ny // 16)]
    # data values other than 0 are not imported; the id map applies
    lut = np.arange(256, dtype=np.uint8)[::-1].copy()
    lut[0] = 0
    got2, _ = mc_import(d, (nx, ny, nz), (nx // 2, 0, nz // 2), lut, None)
    np.testing.assert_array_equal(got2, np.where(want > 0, 255 - want, 0).astype(np.uint8))


def test_palette_semantics(tmp_path):
    """1.18 chunk: a name missing from the table imports as stone (1:0), a
    non-zero data value is skipped, indices past the palette are air, and a
    section whose palette has one entry (no data array) is not imported."""
    pal = [M.compound(("Name", M.Tag(M.STRING, n))) for n in ("minecraft:air", "test:unknown", "minecraft:granite",
                                                               "minecraft:dirt")]
    idx = np.tile(np.arange(6), 4096 // 6 + 1)[:4096]            # 4, 5: past the 4-entry palette
    # data before palette, and a sibling after block_states: the palette loop
    # ends on block_states' End and the block_states loop on `biomes`, so the
    # section's own End records it at its Y (the reader's order dependence)
    biomes = M.compound(("palette", M.lst(M.STRING, [M.Tag(M.STRING, "minecraft:plains")])))
    sec = M.compound(("Y", M.Tag(M.BYTE, 0)), ("block_states", M.compound(
        ("data", M.Tag(M.LONG_ARRAY, M.pack_indices(idx, 4, False))), ("palette", M.lst(M.COMPOUND, pal)))),
        ("biomes", biomes))
    single = M.compound(("Y", M.Tag(M.BYTE, 1)), ("block_states", M.compound(
        ("palette", M.lst(M.COMPOUND, [M.compound(("Name", M.Tag(M.STRING, "minecraft:stone")))])))),
        ("biomes", biomes))
    root = M.compound(("DataVersion", M.Tag(M.INT, 2975)), ("xPos", M.Tag(M.INT, 0)), ("zPos", M.Tag(M.INT, 0)),
                      ("sections", M.lst(M.COMPOUND, [sec, single])))
    d = _write(tmp_path, {"r.0.0.mca": M.region({0: M.encode_root(root)})})
    names = {"minecraft:granite": (1, 1), "minecraft:dirt": (3, 0), "minecraft:air": (0, 0)}
    grid, stats = mc_import(d, (16, 32, 16), (8, 0, 8), None, names)
    assert stats == [1, 1, 1]
    cells = grid[:, :16, :].transpose(1, 0, 2).reshape(-1)       # y*256 + z*16 + x
    want = np.array([0, 1, 0, 3, 0, 0], np.uint8)[idx]            # air, stone, granite has data 1, dirt, air, air
    np.testing.assert_array_equal(cells, want)
    assert (grid[:, 16:, :] == 0).all()


def test_import_errors(tmp_path):
    with pytest.raises(VptError):
        mc_import(tmp_path / "missing", (16, 16, 16))
    lut = np.arange(256, dtype=np.uint8)
    lut[0] = 3
    with pytest.raises(VptError):
        mc_import(tmp_path, (16, 16, 16), (0, 0, 0), lut)
    grid, stats = mc_import(tmp_path, (16, 16, 16))            # no region files: an empty world
    assert stats == [0, 0, 0] and not grid.any()


def _chunks_of(region: bytes) -> dict:
    """{index: decompressed NBT} of a region written by mca_synth.region."""
    import gzip
    import struct
    import zlib
    out = {}
    for idx in range(1024):
        loc = struct.unpack(">I", region[4 * idx:4 * idx + 4])[0]
        if loc:
            off = (loc >> 8) * 4096
            n, ctype = struct.unpack(">iB", region[off:off + 5])
            data = region[off + 5:off + 4 + n]
            if ctype == 2:
                try:
                    data = zlib.decompress(data)
                except zlib.error:
                    continue
            elif ctype == 1:
                data = gzip.decompress(data)
            out[idx] = data
    return out


@pytest.mark.parametrize("seed", range(4))
def test_mutated_files_import_without_error(tmp_path, seed):
    """Corrupted region files (flipped bytes, huge or negative lengths and
    counts, truncated NBT, header sectors past the end of the file) import
    without an error: a chunk the reader cannot walk is skipped, as the
    reference's reader skips it, and every read stays inside its buffer."""
    import struct
    rng = np.random.default_rng(seed)
    for it in range(25):
        case = M.CASES[it % len(M.CASES)]
        d = tmp_path / f"c{it}"
        d.mkdir()
        for name, reg in M.build_case(case, 1000 * seed + it).items():
            chunks = {}
            for i, nbt in _chunks_of(reg).items():
                b = bytearray(nbt)
                for _ in range(int(rng.integers(1, 5))):
                    p = int(rng.integers(0, len(b)))
                    k = int(rng.integers(0, 4))
                    if k == 0:
                        b[p] ^= 1 << int(rng.integers(0, 8))
                    elif k == 1 and p + 4 <= len(b):
                        b[p:p + 4] = struct.pack(">I", int(rng.choice([0x7fffffff, 0x80000000, 0xffffffff, 4096])))
                    elif k == 2:
                        del b[max(p, 1):]
                    else:
                        b[p] = int(rng.integers(0, 13))
                chunks[i] = bytes(b)
            reg = bytearray(M.region(chunks, {i: str(rng.choice(["zlib", "gzip", "raw"])) for i in chunks}))
            if rng.random() < 0.4:   # a location far past the end of the file, or a cut file
                j = int(rng.integers(0, 1024))
                reg[4 * j:4 * j + 4] = struct.pack(">I", int(rng.integers(0, 2**32)))
            if rng.random() < 0.2:
                del reg[int(rng.integers(0, len(reg))):]
            (d / name).write_bytes(bytes(reg))
        grid, stats = mc_import(d, M.DIMS, M.ORIGIN, M.block_lut(), M.NAMES)
        assert stats[0] == len(list(d.iterdir())) and stats[1] >= 0 and stats[2] >= 0


@pytest.mark.skipif(not MEDIVAL.exists(), reason="the reference's test world is read where the reference is present")
def test_reference_test_world():
    """The reference's own test world (Test MC Worlds/Medival, pre-1.13
    chunks): at its stated origin 136 0 -94 and at origins whose window
    covers each region file's saved chunks, the imported grid has the
    digest, solid count and section count the reference's reader gave
    (tests/golden/mca_medival.json), and, where that reader is built,
    equals it live.  At the stated origin the grid is empty: the files hold
    a ring of chunks around the build, none inside the 384 x 384 window."""
    from tests.golden.make_mca_golden import MEDIVAL_LUTS, reference_import, reference_names
    golden = json.loads((GOLDEN / "mca_medival.json").read_text())
    assert sorted(p.name for p in MEDIVAL.glob("*.mca")) == golden["files"]
    ref = C.CDLL(str(REF_LIB)) if REF_LIB.exists() else None
    files = {p.name: p.read_bytes() for p in sorted(MEDIVAL.glob("*.mca"))}
    for case in golden["cases"]:
        lut = MEDIVAL_LUTS[case["lut"]]()
        got, stats = mc_import(MEDIVAL, (384, 128, 384), tuple(map(float, case["origin"])), lut, None)
        assert stats[0] == 7 and stats[2] == case["sections"], (case, stats)
        assert int((got > 0).sum()) == case["solid"], case
        assert hashlib.sha256(got.tobytes()).hexdigest() == case["sha256"], case
        if ref is not None:
            want, sections = reference_import(ref, files, (384, 128, 384), case["origin"], lut)
            np.testing.assert_array_equal(got, want)
    if ref is not None:
        assert len(reference_names(ref)) > 700
    assert sum(c["solid"] > 0 for c in golden["cases"]) == len(golden["cases"]) - 2
