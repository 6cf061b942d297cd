"""Synthetic Minecraft region files for the import tests: an NBT encoder
that keeps tag order as given (the reference's reader depends on it) and a
region writer (1024 big-endian sector locations, zlib or gzip chunk
payloads).  Chunk layouts follow the public chunk format: before 1.13
(Level.Sections Blocks / Data / Add), 1.13-1.17 (Level.Sections Palette /
BlockStates) and 1.18+ (sections[].block_states)."""
from __future__ import annotations

import gzip
import struct
import zlib

import numpy as np

END, BYTE, SHORT, INT, LONG, FLOAT, DOUBLE, BYTE_ARRAY, STRING, LIST, COMPOUND, INT_ARRAY, LONG_ARRAY = range(13)


class Tag:
    """A typed NBT value; compounds are ordered lists of (name, Tag)."""

    def __init__(self, tid, value, item=None):
        self.tid, self.value, self.item = tid, value, item


def compound(*items):
    return Tag(COMPOUND, list(items))


def lst(item_tid, values):
    return Tag(LIST, list(values), item_tid)


def _payload(t: Tag) -> bytes:
    v = t.value
    if t.tid == BYTE:
        return struct.pack(">b", v)
    if t.tid == SHORT:
        return struct.pack(">h", v)
    if t.tid == INT:
        return struct.pack(">i", v)
    if t.tid == LONG:
        return struct.pack(">q", v)
    if t.tid == FLOAT:
        return struct.pack(">f", v)
    if t.tid == DOUBLE:
        return struct.pack(">d", v)
    if t.tid == BYTE_ARRAY:
        b = bytes(np.asarray(v, np.uint8))
        return struct.pack(">i", len(b)) + b
    if t.tid == STRING:
        b = v.encode()
        return struct.pack(">H", len(b)) + b
    if t.tid == LIST:
        return struct.pack(">bi", t.item, len(v)) + b"".join(_payload(x) for x in v)
    if t.tid == COMPOUND:
        return b"".join(named(n, x) for n, x in v) + bytes([END])
    if t.tid == INT_ARRAY:
        a = np.asarray(v, ">i4")
        return struct.pack(">i", len(a)) + a.tobytes()
    if t.tid == LONG_ARRAY:
        a = np.asarray(v, np.uint64).astype(">u8")
        return struct.pack(">i", len(a)) + a.tobytes()
    raise ValueError(t.tid)


def named(name: str, t: Tag) -> bytes:
    b = name.encode()
    return bytes([t.tid]) + struct.pack(">H", len(b)) + b + _payload(t)


def encode_root(t: Tag, name: str = "") -> bytes:
    return named(name, t)


def region(chunks: dict, compression: dict | None = None) -> bytes:
    """chunks: {index 0..1023: root NBT bytes}; compression: {index: 'zlib' |
    'gzip' | 'raw' | 'garbage'} (zlib by default)."""
    compression = compression or {}
    header = bytearray(8192)
    body = bytearray()
    sector = 2
    for idx in sorted(chunks):
        nbt = chunks[idx]
        kind = compression.get(idx, "zlib")
        if kind == "zlib":
            data, ctype = zlib.compress(nbt), 2
        elif kind == "gzip":
            data, ctype = gzip.compress(nbt, mtime=0), 1
        elif kind == "raw":
            data, ctype = nbt, 3
        else:
            data, ctype = b"\x78\x9c" + bytes(range(40)), 2
        blob = struct.pack(">iB", len(data) + 1, ctype) + data
        n_sec = (len(blob) + 4095) // 4096
        blob += bytes(n_sec * 4096 - len(blob))
        header[4 * idx:4 * idx + 4] = struct.pack(">I", (sector << 8) | min(n_sec, 255))
        body += blob
        sector += n_sec
    return bytes(header) + bytes(body)


def pack_indices(idx: np.ndarray, bits: int, spanning: bool) -> np.ndarray:
    """Palette indices (4096, y-z-x order) -> longs: spanning (before data
    version 2556) runs indices across longs, packed keeps 64 // bits per long."""
    idx = np.asarray(idx, np.uint64)
    if spanning:
        n = (4096 * bits + 63) // 64
        out = [0] * n
        for i, v in enumerate(idx.tolist()):
            pos = i * bits
            w, o = divmod(pos, 64)
            out[w] |= (v << o) & 0xFFFFFFFFFFFFFFFF
            if o + bits > 64:
                out[w + 1] |= v >> (64 - o)
        return np.array(out, np.uint64)
    per = 64 // bits
    n = (4096 + per - 1) // per
    out = [0] * n
    for i, v in enumerate(idx.tolist()):
        w, k = divmod(i, per)
        out[w] |= v << (k * bits)
    return np.array(out, np.uint64)


def palette_bits(size: int) -> int:
    return max(4, int(size - 1).bit_length())


def legacy_section(rng, y, ids, data_zero_frac=0.8, with_y=True, with_add=False):
    blocks = rng.choice(ids, 4096).astype(np.uint8)
    nib = np.where(rng.random(4096) < data_zero_frac, 0, rng.integers(1, 16, 4096)).astype(np.uint8)
    data = (nib[0::2] | (nib[1::2] << 4)).astype(np.uint8)
    items = []
    if with_y:
        items.append(("Y", Tag(BYTE, y)))
    items += [("Blocks", Tag(BYTE_ARRAY, blocks)), ("Data", Tag(BYTE_ARRAY, data)),
              ("SkyLight", Tag(BYTE_ARRAY, np.zeros(2048, np.uint8)))]
    if with_add:
        items.append(("Add", Tag(BYTE_ARRAY, rng.integers(0, 256, 2048).astype(np.uint8))))
    return compound(*items)


def palette_entries(rng, names, size):
    pal = []
    for k in range(size):
        name = names[k % len(names)] if k < len(names) else names[rng.integers(len(names))]
        props = [("axis", Tag(STRING, "y"))] if k % 3 == 1 else []
        items = [("Name", Tag(STRING, name))]
        if props:
            items = [("Properties", compound(*props))] + items if k % 2 else items + [("Properties", compound(*props))]
        pal.append(compound(*items))
    return pal


# Palette names used by the cases and their legacy (id, data value), as
# Minecraft's pre-flattening ids define them; "test:unknown" is in no table.
NAMES = {
    "minecraft:air": (0, 0), "minecraft:stone": (1, 0), "minecraft:granite": (1, 1),
    "minecraft:grass_block": (2, 0), "minecraft:dirt": (3, 0), "minecraft:cobblestone": (4, 0),
    "minecraft:oak_planks": (5, 0), "minecraft:bedrock": (7, 0), "minecraft:sand": (12, 0),
    "minecraft:gravel": (13, 0), "minecraft:gold_ore": (14, 0), "minecraft:oak_log": (17, 0),
    "minecraft:glass": (20, 0), "minecraft:sandstone": (24, 0), "minecraft:bricks": (45, 0),
    "minecraft:obsidian": (49, 0),
}
PALETTE_NAMES = list(NAMES) + ["test:unknown"]
LEGACY_IDS = np.array([0, 0, 0, 1, 2, 3, 4, 5, 7, 12, 13, 17, 20, 24, 45, 49, 98, 155, 200], np.uint8)


def _chunk_legacy(rng, cx, cz, ys, order="std", data_version=None, missing_y=()):
    secs = [legacy_section(rng, y, LEGACY_IDS, with_y=y not in missing_y, with_add=(k % 2 == 1))
            for k, y in enumerate(ys)]
    level = [("xPos", Tag(INT, cx)), ("zPos", Tag(INT, cz)), ("LastUpdate", Tag(LONG, 7)),
             ("Sections", lst(COMPOUND, secs)), ("Entities", lst(COMPOUND, []))]
    if order == "pos_last":
        level = level[2:] + level[:2]
    root = [("Level", compound(*level))]
    if data_version is not None:
        root = [("DataVersion", Tag(INT, data_version))] + root if order != "dv_last" else root + [
            ("DataVersion", Tag(INT, data_version))]
    return encode_root(compound(*root))


def _palette_section(rng, y, size, spanning, order):
    bits = palette_bits(size)
    idx = rng.integers(0, size, 4096)
    idx[rng.random(4096) < 0.01] = size + 3 if size + 3 < (1 << bits) else 0   # indices past the palette: air
    states = Tag(LONG_ARRAY, pack_indices(idx, bits, spanning))
    pal = lst(COMPOUND, palette_entries(rng, PALETTE_NAMES, size))
    parts = {"Y": ("Y", Tag(BYTE, y)), "S": ("BlockStates", states), "P": ("Palette", pal),
             "L": ("BlockLight", Tag(BYTE_ARRAY, np.zeros(2048, np.uint8)))}
    return compound(*[parts[c] for c in order])


def _chunk_palette(rng, cx, cz, ys, sizes, data_version, orders):
    spanning = data_version < 2556
    secs = [_palette_section(rng, y, s, spanning, o) for y, s, o in zip(ys, sizes, orders)]
    level = [("xPos", Tag(INT, cx)), ("zPos", Tag(INT, cz)), ("Sections", lst(COMPOUND, secs)),
             ("Status", Tag(STRING, "full"))]
    return encode_root(compound(("DataVersion", Tag(INT, data_version)), ("Level", compound(*level))))


def _section_118(rng, y, size, order, single=False):
    items = {"Y": ("Y", Tag(BYTE, y))}
    if single:
        bs = compound(("palette", lst(COMPOUND, palette_entries(rng, PALETTE_NAMES[3:], 1))))
    else:
        bits = palette_bits(size)
        idx = rng.integers(0, size, 4096)
        data = ("data", Tag(LONG_ARRAY, pack_indices(idx, bits, False)))
        pal = ("palette", lst(COMPOUND, palette_entries(rng, PALETTE_NAMES, size)))
        bs = compound(*([data, pal] if "d" in order else [pal, data]))
    items["B"] = ("block_states", bs)
    items["b"] = ("biomes", compound(("palette", lst(STRING, [Tag(STRING, "minecraft:plains")]))))
    return compound(*[items[c] for c in order if c in items])


def _chunk_118(rng, cx, cz, specs, order="std"):
    secs = [_section_118(rng, y, size, o, single) for y, size, o, single in specs]
    root = [("DataVersion", Tag(INT, 2975)), ("xPos", Tag(INT, cx)), ("yPos", Tag(INT, -4)), ("zPos", Tag(INT, cz)),
            ("sections", lst(COMPOUND, secs)), ("Status", Tag(STRING, "minecraft:full"))]
    if order == "sections_first":
        root = [root[4], root[0], root[1], root[3], root[5]]
    return encode_root(compound(*root))


def build_case(name: str, seed: int) -> dict:
    """{file name: region bytes} of one import case."""
    rng = np.random.default_rng(seed)
    if name == "legacy":
        chunks = {0: _chunk_legacy(rng, 0, 0, [0, 1]), 1: _chunk_legacy(rng, 1, 0, [0, 2], missing_y=(2,)),
                  32: _chunk_legacy(rng, 0, 1, [1], order="pos_last", data_version=1343),
                  33: _chunk_legacy(rng, -1, -1, [0, 1], data_version=1343, order="dv_last")}
        return {"r.0.0.mca": region(chunks),
                "r.-1.-1.mca": region({1023: _chunk_legacy(rng, -2, -2, [0, 3])})}
    if name == "palette_spanning":
        ys, sizes = [0, 1, 2, 3], [5, 17, 40, 16]
        chunks = {0: _chunk_palette(rng, 0, 0, ys, sizes, 1976, ["YLSP", "LPSY", "PYS", "SPY"]),
                  1: _chunk_palette(rng, 1, 1, [0, 1], [3, 70], 2230, ["YSP", "YSP"])}
        return {"r.0.0.mca": region(chunks)}
    if name == "palette_packed":
        chunks = {0: _chunk_palette(rng, 0, 0, [0, 1, 2], [5, 17, 33], 2586, ["YSP", "SPY", "YPS"]),
                  2: _chunk_palette(rng, -1, 0, [1], [9], 2586, ["YSP"])}
        return {"r.0.0.mca": region(chunks)}
    if name == "sections_118":
        specs = [(-1, 6, "YBb", False), (0, 20, "YdBb", False), (1, 1, "YB", True), (2, 9, "BYb", False),
                 (3, 12, "bBY", False)]
        chunks = {0: _chunk_118(rng, 0, 0, specs), 1: _chunk_118(rng, 1, 0, specs[:2], order="sections_first")}
        return {"r.0.0.mca": region(chunks)}
    if name == "payloads":
        good = _chunk_legacy(rng, 0, 0, [0])
        no_x = encode_root(compound(("Level", compound(("zPos", Tag(INT, 0)), ("Sections", lst(COMPOUND, [
            legacy_section(rng, 1, LEGACY_IDS)]))))))
        chunks = {0: good, 1: _chunk_legacy(rng, 1, 0, [1]), 2: no_x, 3: _chunk_legacy(rng, 0, 1, [2]),
                  4: _chunk_legacy(rng, 1, 1, [0])}
        return {"r.0.0.mca": region(chunks, {1: "gzip", 3: "garbage", 4: "raw"})}
    raise KeyError(name)


CASES = ("legacy", "palette_spanning", "palette_packed", "sections_118", "payloads")
# world, origin and block-id map of every case: 48 x 48 x 48 centred on the
# region's first chunks, origin truncated toward zero
DIMS = (48, 48, 48)
ORIGIN = (8.7, -3.9, 9.2)


def block_lut(seed: int = 5) -> np.ndarray:
    rng = np.random.default_rng(seed)
    lut = rng.integers(1, 128, 256).astype(np.uint8)
    lut[0] = 0
    lut[rng.integers(1, 256, 20)] = 0          # ids that import as air
    return lut
