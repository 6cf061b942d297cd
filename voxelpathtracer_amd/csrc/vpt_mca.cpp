// This file is an altered C++ port of parts of enkiMI (the region and NBT
// chunk reader the reference vendors as Dependencies/enkiMI/enkimi.c):
// rewritten with bounds checks, the system zlib, and tag names matched
// without rewriting the stream.  It is not the original software.  enkiMI's
// licence notice follows, as its terms require:
//
//   Copyright (c) 2017 Juliette Foucaut & Doug Binks
//
//   This software is provided 'as-is', without any express or implied
//   warranty. In no event will the authors be held liable for any damages
//   arising from the use of this software.
//
//   Permission is granted to anyone to use this software for any purpose,
//   including commercial applications, and to alter it and redistribute it
//   freely, subject to the following restrictions:
//
//   1. The origin of this software must not be misrepresented; you must not
//      claim that you wrote the original software. If you use this software
//      in a product, an acknowledgement in the product documentation would be
//      appreciated but is not required.
//   2. Altered source versions must be plainly marked as such, and must not be
//      misrepresented as being the original software.
//   3. This notice may not be removed or altered from any source distribution.
//
// Minecraft world import: MCWorldImporter::ImportWorld (Core/NBT/Importer.cpp:
// 86-168) over the enkiMI region reader it links (Dependencies/enkiMI/
// enkimi.c:1535-1638, 1743-1895, 1905-1966, 1998-2406).  Host code only:
// region files -> a host block grid, zlib inflating the chunks.
//
// The chunk reader walks the NBT stream tag by tag with enkiMI's level
// bookkeeping and matches tag names wherever the walk stands, so a chunk
// decodes exactly as the reference decodes it, quirks included: a loop
// over a compound or list ends by consuming the tag that follows it (a
// `Y` or `BlockStates` right after a `Palette`, or `data` right after a
// `palette`, is then not seen), the first `Blocks` / `Data` / `BlockStates`
// / `data` / `Palette` / `palette` of a section is kept and the last `Y`, a
// section without `Y` takes the previous one's plus one, `Add` nibbles are
// ignored, and a palette section without a data array is not imported.
#include <dirent.h>
#include <stdio.h>
#include <string.h>
#include <zlib.h>

#include <algorithm>
#include <memory>
#include <string>
#include <string_view>
#include <unordered_map>
#include <vector>

#include "vpt/vpt.h"
#include "vpt_internal.h"

using vpt::set_error;

namespace {

enum : uint8_t {
    kEnd = 0, kByte = 1, kShort = 2, kInt = 3, kLong = 4, kFloat = 5, kDouble = 6,
    kByteArray = 7, kString = 8, kList = 9, kCompound = 10, kIntArray = 11, kLongArray = 12,
};

constexpr size_t kSector = 4096;          // region sector (enkimi.c SECTOR_SIZE)
constexpr size_t kRegionHeader = 8192;    // 1024 locations + 1024 timestamps
constexpr int kSectionsYOffset = 128;     // section index = Y + 128 (enkimi.h:167)
constexpr int kSections = 256;
constexpr int kMaxLevel = 512;

uint32_t be32(const uint8_t* p) { return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3]; }

// A chunk's compressed payload -> NBT bytes (enkimi.c:1547-1638): a gzip
// member without header flags (deflate data after the 10-byte header,
// ISIZE from its last 4 bytes), anything else as a zlib stream.  false:
// the reference's reader leaves the chunk empty.
bool inflate_chunk(const uint8_t* src, size_t n, std::vector<uint8_t>& out)
{
    out.clear();
    z_stream zs;
    memset(&zs, 0, sizeof(zs));
    int wbits = 15;
    if (n > 10 && src[0] == 0x1f && src[1] == 0x8b) {
        if (src[3] != 0) return false;
        wbits = -15;
        src += 10;
        n -= 10;
    }
    if (inflateInit2(&zs, wbits) != Z_OK) return false;
    zs.next_in = const_cast<uint8_t*>(src);
    zs.avail_in = (uInt)n;
    out.resize(std::max<size_t>(n * 4 + 1024, 4096));
    int rc = Z_OK;
    while (rc == Z_OK) {
        if (zs.total_out == out.size()) out.resize(out.size() * 2);
        zs.next_out = out.data() + zs.total_out;
        zs.avail_out = (uInt)(out.size() - zs.total_out);
        rc = inflate(&zs, Z_FINISH);
        if (rc == Z_BUF_ERROR && zs.avail_out == 0) rc = Z_OK;   // output full: grow and go on
    }
    const bool ok = rc == Z_STREAM_END;
    out.resize(zs.total_out);
    inflateEnd(&zs);
    return ok;
}

// One tag as the walk reports it.  name is empty for list items and End.
struct Tag {
    uint8_t id = kEnd;
    std::string_view name;
    uint8_t item_id = kEnd;   // lists: item type, count, items handed out so far
    int32_t n_items = 0;
    int32_t cur_item = 0;
};

// Depth-first walk of an NBT byte stream with enkiMI's levels
// (enkiNBTReadNextTag, enkimi.c:1806-1895): the root is at level -1, the
// children of a compound or list one level below their parent, and an End
// tag is reported at its compound's level.  Payload reads start where the
// current tag's payload starts (pos) and do not move the walk.  Malformed
// data (a read or jump past the end) stops the walk.
class NbtWalk {
  public:
    NbtWalk(const uint8_t* p, size_t n) : beg_(p), end_(p + n), pos_(p), next_(p) {}

    int level() const { return level_; }
    const Tag& tag() const { return cur_; }
    const Tag& parent(int l) const { return parents_[l]; }
    bool ok() const { return ok_; }

    bool next()
    {
        if (!ok_) return false;
        if (cur_.id == kCompound || cur_.id == kList) {
            if (++level_ == kMaxLevel) return fail();
            parents_[level_] = cur_;
        }
        while (level_ >= 0 && parents_[level_].id == kList) {
            Tag& l = parents_[level_];
            if (l.cur_item == l.n_items) {
                level_--;
                continue;
            }
            cur_ = Tag();
            cur_.id = l.item_id;
            if (cur_.id == kList) {
                if (!has(pos_, 5)) return fail();
                cur_.item_id = *pos_++;
                cur_.n_items = (int32_t)be32(pos_);
                pos_ += 4;
                next_ = pos_;
            }
            if (!skip()) return fail();
            pos_ = next_;
            l.cur_item++;
            return true;
        }
        if (next_ >= end_) return false;
        pos_ = next_;
        cur_ = Tag();
        cur_.id = *pos_++;
        if (cur_.id > kLongArray) return fail();
        if (cur_.id != kEnd) {
            if (!has(pos_, 2)) return fail();
            const size_t len = ((size_t)pos_[0] << 8) | pos_[1];
            pos_ += 2;
            if (!has(pos_, len)) return fail();
            cur_.name = std::string_view(reinterpret_cast<const char*>(pos_), len);
            pos_ += len;
        }
        if (cur_.id == kList) {
            if (!has(pos_, 5)) return fail();
            cur_.item_id = *pos_++;
            cur_.n_items = (int32_t)be32(pos_);
            pos_ += 4;
        }
        next_ = pos_;
        if (!skip()) return fail();
        if (level_ >= 0 && cur_.id == kEnd) level_--;
        return true;
    }

    // Payload reads at the current tag's payload start, whatever its type
    // (the reference reads DataVersion, xPos and zPos as ints, Y as a byte).
    bool read_i8(int8_t& v) const
    {
        if (!has(pos_, 1)) return false;
        v = (int8_t)pos_[0];
        return true;
    }
    bool read_i32(int32_t& v) const
    {
        if (!has(pos_, 4)) return false;
        v = (int32_t)be32(pos_);
        return true;
    }
    // An array payload: its element count and first element.
    bool read_array(int32_t& count, const uint8_t*& data) const
    {
        if (!read_i32(count)) return false;
        data = pos_ + 4;
        return true;
    }
    bool read_string(std::string_view& s) const
    {
        if (!has(pos_, 2)) return false;
        const size_t len = ((size_t)pos_[0] << 8) | pos_[1];
        if (!has(pos_ + 2, len)) return false;
        s = std::string_view(reinterpret_cast<const char*>(pos_ + 2), len);
        return true;
    }
    const uint8_t* end() const { return end_; }

  private:
    bool has(const uint8_t* p, size_t k) const { return p >= beg_ && p <= end_ && (size_t)(end_ - p) >= k; }
    bool fail()
    {
        ok_ = false;
        return false;
    }
    // Moves next_ past the current tag's payload (SkipDataToNextTag); the
    // payload of a compound or list is walked, not skipped.
    bool skip()
    {
        size_t k = 0;
        switch (cur_.id) {
        case kEnd: case kList: case kCompound: return true;
        case kByte: k = 1; break;
        case kShort: k = 2; break;
        case kInt: case kFloat: k = 4; break;
        case kLong: case kDouble: k = 8; break;
        case kString: {
            if (!has(pos_, 2)) return false;
            k = 2 + (((size_t)pos_[0] << 8) | pos_[1]);
            next_ = pos_;
            break;
        }
        case kByteArray: case kIntArray: case kLongArray: {
            if (!has(pos_, 4)) return false;
            const int32_t cnt = (int32_t)be32(pos_);
            if (cnt < 0) return false;
            const size_t w = cur_.id == kByteArray ? 1 : (cur_.id == kIntArray ? 4 : 8);
            k = 4 + (size_t)cnt * w;
            next_ = pos_;
            break;
        }
        default: return false;
        }
        if (!has(next_, k)) return false;
        next_ += k;
        return true;
    }

    const uint8_t* beg_;
    const uint8_t* end_;
    const uint8_t* pos_;
    const uint8_t* next_;
    int level_ = -1;
    bool ok_ = true;
    Tag cur_;
    Tag parents_[kMaxLevel];
};

// Section palette (enkiChunkSectionPalette): table index per entry (-1:
// not in the table), bits per index, the BlockStates / data longs.
struct Palette {
    std::vector<int32_t> index;
    uint32_t bits = 0;
    int32_t n_longs = 0;
};

struct Section {
    const uint8_t* blocks = nullptr;   // legacy: 4096 ids, or palette indices (big-endian longs)
    const uint8_t* data = nullptr;     // legacy: 2048 bytes of 4-bit data values
    Palette palette;
};

struct Chunk {
    int32_t data_version = 0, x = 0, z = 0;
    Section sections[kSections];
};

using NameTable = std::unordered_map<std::string_view, int32_t>;

// LoadChunkPalette (enkimi.c:1998-2061) at a Palette / palette list tag.
void load_palette(NbtWalk& s, Palette& pal, const NameTable& names)
{
    const uint32_t size = (uint32_t)s.tag().n_items;
    // floor(1 + log2(max(size - 1, 15))): the bit length of size - 1, at least 4
    uint32_t bits = 4;
    while (bits < 32 && ((size - 1) >> bits) != 0) bits++;
    pal.bits = bits;
    pal.index.assign(size, -1);
    const int lp = s.level();
    while (s.next() && lp < s.level()) {
        const int entry = s.parent(lp + 1).cur_item - 1;
        if (s.tag().id == kString && s.tag().name == "Name" && entry >= 0 && (uint32_t)entry < size) {
            std::string_view v;
            if (s.read_string(v)) {
                const auto it = names.find(v);
                pal.index[entry] = it == names.end() ? -1 : it->second;
            }
        }
        if (s.tag().name == "Properties") {
            const int lq = s.level();
            while (s.next() && lq < s.level()) {
            }
        }
    }
}

// enkiNBTReadChunk (enkimi.c:2064-2271).  false: the chunk lacks xPos,
// zPos or its sections and contributes nothing.
bool read_chunk(NbtWalk& s, Chunk& c, const NameTable& names)
{
    bool fx = false, fz = false, fsec = false;
    while (s.next()) {
        const Tag& t = s.tag();
        int32_t v;
        if (t.name == "DataVersion") {
            if (s.read_i32(v)) c.data_version = v;
        } else if (!fx && t.name == "xPos") {
            fx = true;
            if (s.read_i32(v)) c.x = v;
        } else if (!fz && t.name == "zPos") {
            fz = true;
            if (s.read_i32(v)) c.z = v;
        } else if (t.name == "yPos") {
        } else if (!fsec && t.name == "sections") {
            // data version 2844+: sections[] { Y, block_states { palette, data } }
            fsec = true;
            int8_t sy = 0;
            const uint8_t* states = nullptr;
            Palette pal;
            const int ls = s.level();
            while (s.next() && s.level() > ls) {
                if (s.tag().name == "block_states") {
                    const int lb = s.level();
                    while (s.next() && s.level() > lb) {
                        if (!states && s.tag().name == "data") {
                            int32_t n;
                            const uint8_t* p;
                            if (s.read_array(n, p)) {
                                pal.n_longs = n;
                                states = p;
                            }
                        } else if (pal.index.empty() && s.tag().name == "palette") {
                            load_palette(s, pal, names);
                        }
                    }
                } else if (s.tag().name == "Y") {
                    s.read_i8(sy);
                } else if (s.tag().id == kEnd && s.level() == ls + 1) {
                    const int idx = (int)sy + kSectionsYOffset;
                    if (idx >= 0 && idx < kSections) {
                        c.sections[idx].palette = pal;
                        c.sections[idx].blocks = states;
                    }
                    states = nullptr;
                    pal = Palette();
                    ++sy;
                }
            }
        } else if (t.name == "Level") {
            // before data version 2844: Level { xPos, zPos, Sections[] }
            const int ll = s.level();
            while (s.next() && s.level() > ll) {
                const Tag& u = s.tag();
                if (!fx && u.name == "xPos") {
                    fx = true;
                    if (s.read_i32(v)) c.x = v;
                } else if (!fz && u.name == "zPos") {
                    fz = true;
                    if (s.read_i32(v)) c.z = v;
                } else if (!fsec && u.name == "Sections") {
                    fsec = true;
                    int8_t sy = 0;
                    const uint8_t *blocks = nullptr, *data = nullptr, *states = nullptr;
                    Palette pal;
                    const int ls = s.level();
                    while (s.next() && s.level() > ls) {
                        const Tag& q = s.tag();
                        int32_t n;
                        const uint8_t* p;
                        if (!blocks && q.name == "Blocks") {
                            if (s.read_array(n, p)) blocks = p;
                        } else if (q.name == "Add") {
                        } else if (!data && q.name == "Data") {
                            if (s.read_array(n, p)) data = p;
                        } else if (q.name == "Y") {
                            s.read_i8(sy);
                        } else if (!states && q.name == "BlockStates") {
                            if (s.read_array(n, p)) {
                                pal.n_longs = n;
                                states = p;
                            }
                        } else if (pal.index.empty() && q.name == "Palette") {
                            load_palette(s, pal, names);
                        } else if (q.id == kEnd && s.level() == ls + 1) {
                            const int idx = (int)sy + kSectionsYOffset;
                            if (idx >= 0 && idx < kSections) {
                                if (blocks) {
                                    c.sections[idx].blocks = blocks;
                                    c.sections[idx].data = data;
                                }
                                if (states && !pal.index.empty()) {
                                    c.sections[idx].palette = pal;
                                    c.sections[idx].blocks = states;
                                }
                            }
                            blocks = data = states = nullptr;
                            pal = Palette();
                            ++sy;
                        }
                    }
                }
            }
        }
        if (fx && fz && fsec && c.data_version) break;
    }
    return fx && fz && fsec;
}

struct Voxel {
    uint8_t id, data;
};

inline uint64_t be64(const uint8_t* p)
{
    return ((uint64_t)be32(p) << 32) | be32(p + 4);
}

// enkiGetChunkSectionVoxelData (enkimi.c:2291-2406) at offset pos = y*256 +
// z*16 + x.  Palette indices: data version 2556+ packs 64 / bits indices per
// long, older versions run them across longs.  An index past the palette is
// air, a palette name missing from the table stone (1:0).  As in the
// reference, arrays shorter than a section are read on into the bytes that
// follow them; only the end of the chunk's data bounds the reads (air past
// it).
Voxel section_voxel(const Chunk& c, const Section& s, uint32_t pos, const uint8_t* end, const uint8_t* ids,
                    const uint8_t* dvs)
{
    Voxel r = {0, 0};
    if (!s.palette.index.empty()) {
        const uint32_t bits = s.palette.bits;
        uint32_t value = 0;
        if (c.data_version >= 2556) {
            const uint32_t per = 64 / bits, w = pos / per, in = bits * (pos - w * per);
            if ((size_t)(end - s.blocks) < 8 * ((size_t)w + 1)) return r;
            value = (uint32_t)((be64(s.blocks + 8 * (size_t)w) >> in) & ((~0ull) >> (64 - bits)));
        } else {
            const uint32_t pb = bits * pos, w = pb / 64, in = pb - w * 64;
            if ((size_t)(end - s.blocks) < 8 * ((size_t)w + 1)) return r;
            const uint32_t lo_bits = std::min(64 - in, bits), hi_bits = bits - lo_bits;
            value = (uint32_t)((be64(s.blocks + 8 * (size_t)w) >> in) & ((~0ull) >> (64 - lo_bits)));
            if (hi_bits) {
                if ((size_t)(end - s.blocks) < 8 * ((size_t)w + 2)) return r;
                value |= (uint32_t)(be64(s.blocks + 8 * ((size_t)w + 1)) & ((~0ull) >> (64 - hi_bits))) << lo_bits;
            }
        }
        if (value < s.palette.index.size()) {
            const int32_t k = s.palette.index[value];
            r.id = 1;
            if (k >= 0) {
                r.id = ids[k];
                r.data = dvs[k];
            }
        }
        return r;
    }
    if ((size_t)(end - s.blocks) <= pos) return r;
    r.id = s.blocks[pos];
    if (s.data && (size_t)(end - s.data) > pos / 2) r.data = 0xf & (s.data[pos / 2] >> (4 * (pos & 1)));
    return r;
}

struct Target {
    uint8_t* grid;
    int nx, ny, nz;
    int ox, oy, oz;   // ivec3(origin)
    const uint8_t* lut;
};

// WriteVoxel (Importer.cpp:68-84): world position - origin, centred in x
// and z; air and positions outside the world are not written.
inline void write_voxel(const Target& t, uint8_t v, int x, int y, int z)
{
    x = x - t.ox + t.nx / 2;
    y = y - t.oy;
    z = z - t.oz + t.nz / 2;
    if (y >= t.ny || x >= t.nx || z >= t.nz || y < 0 || x < 0 || z < 0 || v == 0) return;
    t.grid[(size_t)x + (size_t)t.nx * ((size_t)y + (size_t)t.ny * (size_t)z)] = v;
}

// ImportRegionFile (Importer.cpp:86-148).  Returns the number of chunks
// that contributed, -1 if the file cannot be read.
int import_region(const std::string& path, const Target& t, const NameTable& names, const uint8_t* ids,
                  const uint8_t* dvs, int32_t* n_sections)
{
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) return -1;
    std::vector<uint8_t> file;
    uint8_t buf[65536];
    size_t k;
    while ((k = fread(buf, 1, sizeof(buf), f)) > 0) file.insert(file.end(), buf, buf + k);
    fclose(f);
    int chunks = 0;
    std::vector<uint8_t> nbt;
    std::unique_ptr<Chunk> c(new Chunk);
    for (int i = 0; i < 1024 && file.size() >= kRegionHeader; i++) {
        const size_t loc = (((size_t)file[4 * i] << 16) | ((size_t)file[4 * i + 1] << 8) | file[4 * i + 2]) * kSector;
        if (loc < kRegionHeader || loc + 6 > file.size()) continue;
        const size_t len = be32(&file[loc]);
        if (len < 1 || len - 1 + loc + 5 > file.size()) continue;
        if (!inflate_chunk(&file[loc + 5], len - 1, nbt) || nbt.empty()) continue;
        *c = Chunk();
        NbtWalk s(nbt.data(), nbt.size());
        if (!read_chunk(s, *c, names)) continue;
        chunks++;
        for (int sec = 0; sec < kSections; sec++) {
            const Section& q = c->sections[sec];
            if (!q.blocks) continue;
            if (n_sections) (*n_sections)++;
            const int bx = c->x * 16, by = (sec - kSectionsYOffset) * 16, bz = c->z * 16;
            for (int y = 0; y < 16; y++)
                for (int z = 0; z < 16; z++)
                    for (int x = 0; x < 16; x++) {
                        const Voxel v = section_voxel(*c, q, (uint32_t)(y * 256 + z * 16 + x), s.end(), ids, dvs);
                        if (v.data == 0) write_voxel(t, t.lut[v.id], bx + x, by + y, bz + z);
                    }
        }
    }
    return chunks;
}

}  // namespace

extern "C" VPT_API int vpt_mc_import(const char* dir, int nx, int ny, int nz, const float origin[3],
                                     const uint8_t mc_to_block[256], const vpt_mc_block* names, int n_names,
                                     uint8_t* host_grid, int32_t* stats)
{
    if (!dir || !origin || !mc_to_block || !host_grid || (n_names > 0 && !names) || n_names < 0)
        return set_error(VPT_EINVAL, "vpt_mc_import: null argument");
    if (nx < 1 || ny < 1 || nz < 1 || (size_t)nx * ny * nz > ((size_t)1 << 34))
        return set_error(VPT_EINVAL, "vpt_mc_import: bad dims %dx%dx%d", nx, ny, nz);
    if (mc_to_block[0] != 0) return set_error(VPT_EINVAL, "vpt_mc_import: Minecraft id 0 (air) must map to 0");
    DIR* d = opendir(dir);
    if (!d) return set_error(VPT_EINVAL, "vpt_mc_import: cannot open directory %s", dir);
    std::vector<std::string> files;
    while (const dirent* e = readdir(d)) {
        const std::string n = e->d_name;
        const size_t dot = n.find_last_of('.');
        if (dot != std::string::npos && n.substr(dot + 1) == "mca") files.push_back(std::string(dir) + "/" + n);
    }
    closedir(d);
    std::sort(files.begin(), files.end());

    // palette names -> (legacy id, data value); the first entry of a name wins
    NameTable table;
    std::vector<uint8_t> ids(std::max(n_names, 1)), dvs(std::max(n_names, 1));
    for (int i = 0; i < n_names; i++) {
        if (!names[i].name) return set_error(VPT_EINVAL, "vpt_mc_import: null name in entry %d", i);
        ids[i] = names[i].id;
        dvs[i] = names[i].data;
        table.emplace(std::string_view(names[i].name), i);
    }
    memset(host_grid, 0, (size_t)nx * ny * nz);
    // ImportOrigin -> glm::ivec3: truncation toward zero
    const Target t = {host_grid, nx, ny, nz, (int)origin[0], (int)origin[1], (int)origin[2], mc_to_block};
    int32_t n_files = 0, n_chunks = 0, n_sections = 0;
    for (const std::string& f : files) {
        const int c = import_region(f, t, table, ids.data(), dvs.data(), &n_sections);
        if (c < 0) return set_error(VPT_EINVAL, "vpt_mc_import: cannot read %s", f.c_str());
        n_files++;
        n_chunks += c;
    }
    if (stats) {
        stats[0] = n_files;
        stats[1] = n_chunks;
        stats[2] = n_sections;
    }
    return VPT_OK;
}
