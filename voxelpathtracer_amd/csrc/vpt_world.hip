// Context, world residency and the distance-field build.
//
// DF build (replaces ManhattanDistance{X,Y,Z}.comp dispatched by
// World::GenerateDistanceField, World.cpp:69-113).  Each reference pass is a
// forward relaxation d[i] = min(d[i], d[i-1] + 1) followed by a backward one
// along one axis (the X pass starts from solid ? 0 : 254); composed they give
// min(254, L1 distance to the nearest solid voxel).  The passes keep the
// reference's order and integer semantics (the unorm8 store/load round trip
// is exact for 0..254), so the result is identical byte for byte.
//
//   df_xy_slab_kernel : one workgroup per z-slab when the slab (nx * ny
//                       bytes) fits in 128 KiB of LDS: the slab is staged
//                       with 16-byte loads, one wavefront per row runs the X
//                       pass from the row's solid bitmap, four columns per
//                       lane the Y pass in segments, and the slab is written
//                       back with 16-byte stores.
//   df_x_kernel / df_y_kernel : the same passes in global memory for larger
//                       slabs (lanes along x for the Y pass, so each step is
//                       a coalesced row access).
//   df_z_tile_kernel  : the Z pass over a column tile staged in LDS; it
//                       writes whole bricks of the bricked DF.
//   df_z_kernel       : one lane per (x, y) column, lanes along x; the
//                       backward sweep writes the final byte straight into
//                       the bricked DF (tiles too large for LDS).
#include <hip/hip_runtime.h>
#include <string.h>

#include <stdio.h>

#include <algorithm>
#include <climits>
#include <unordered_set>
#include <vector>

#include "vpt_device.h"
#include "vpt_internal.h"

using vpt::set_error;

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return set_error(VPT_EDEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));      \
    } while (0)

namespace {

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard()
    {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// ---- layout conversion ----------------------------------------------------

// Linear -> bricked copy: one thread per 16-byte group along x.
__global__ void brick_kernel(const uint8_t* __restrict__ lin, uint8_t* __restrict__ brk, int nx, int ny, int nz,
                             uint32_t row_stride, uint32_t slab_stride)
{
    size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t groups = (size_t)(nx / 16) * ny * nz;
    if (gid >= groups) return;
    int gx = (int)(gid % (nx / 16));
    size_t rest = gid / (nx / 16);
    int y = (int)(rest % ny);
    int z = (int)(rest / ny);
    const uint4 v = *reinterpret_cast<const uint4*>(lin + (size_t)z * nx * ny + (size_t)y * nx + (size_t)gx * 16);
    uint32_t words[4] = {v.x, v.y, v.z, v.w};
    for (int k = 0; k < 16; k++) {
        uint8_t b = (uint8_t)(words[k >> 2] >> ((k & 3) * 8));
        brk[vpt::brick_offset(gx * 16 + k, y, z, row_stride, slab_stride)] = b;
    }
}

// Bricked -> linear copy (downloads).
__global__ void unbrick_kernel(const uint8_t* __restrict__ brk, uint8_t* __restrict__ lin, int nx, int ny, int nz,
                               uint32_t row_stride, uint32_t slab_stride)
{
    size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t groups = (size_t)(nx / 16) * ny * nz;
    if (gid >= groups) return;
    int gx = (int)(gid % (nx / 16));
    size_t rest = gid / (nx / 16);
    int y = (int)(rest % ny);
    int z = (int)(rest / ny);
    uint32_t words[4] = {0, 0, 0, 0};
    for (int k = 0; k < 16; k++) {
        uint32_t b = brk[vpt::brick_offset(gx * 16 + k, y, z, row_stride, slab_stride)];
        words[k >> 2] |= b << ((k & 3) * 8);
    }
    *reinterpret_cast<uint4*>(lin + (size_t)z * nx * ny + (size_t)y * nx + (size_t)gx * 16) =
        make_uint4(words[0], words[1], words[2], words[3]);
}

__global__ void set_blocks_kernel(uint8_t* __restrict__ lin, uint8_t* __restrict__ brk, const uint32_t* __restrict__ edits,
                                  const uint8_t* __restrict__ ids, int n, int nx, int ny, int nz, uint32_t row_stride,
                                  uint32_t slab_stride)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t e = edits[i];
    int x = (int)(e & 1023u), y = (int)((e >> 10) & 1023u), z = (int)((e >> 20) & 1023u);
    if (x >= nx || y >= ny || z >= nz) return;
    lin[(size_t)z * nx * ny + (size_t)y * nx + x] = ids[i];
    brk[vpt::brick_offset(x, y, z, row_stride, slab_stride)] = ids[i];
}

// Short edit lists travel in the kernel arguments: no staging buffer and
// no host-to-device copy (an interactive edit is a handful of blocks).
constexpr int kArgEdits = 128;
struct ArgEdits {
    uint32_t edit[kArgEdits];
    uint8_t id[kArgEdits];
};

__global__ void set_blocks_args_kernel(uint8_t* __restrict__ lin, uint8_t* __restrict__ brk, ArgEdits ed, int n,
                                       int nx, int ny, uint32_t row_stride, uint32_t slab_stride)
{
    const int i = threadIdx.x;
    if (i >= n) return;
    const uint32_t e = ed.edit[i];
    const int x = (int)(e & 1023u), y = (int)((e >> 10) & 1023u), z = (int)((e >> 20) & 1023u);
    lin[(size_t)z * nx * ny + (size_t)y * nx + x] = ed.id[i];
    brk[vpt::brick_offset(x, y, z, row_stride, slab_stride)] = ed.id[i];
}

// ---- DF passes ---------------------------------------------------------------

// Packed relaxation of four byte columns at once: a column chunk's bytes
// 0 / 2 and 1 / 3 as two pairs of 16-bit lanes (values stay at most 255, so
// prev + 1 never carries into the next lane).
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

struct Cols4 {
    u16x2 e, o;   // bytes 0, 2 and 1, 3
};

__device__ inline Cols4 cols4_of(uint32_t d)
{
    return Cols4{__builtin_bit_cast(u16x2, d & 0x00ff00ffu), __builtin_bit_cast(u16x2, (d >> 8) & 0x00ff00ffu)};
}

__device__ inline uint32_t cols4_pack(Cols4 c)
{
    return __builtin_bit_cast(uint32_t, c.e) | (__builtin_bit_cast(uint32_t, c.o) << 8);
}

// min(d, c + k) per column
__device__ inline Cols4 cols4_relax(Cols4 d, Cols4 c, unsigned short k)
{
    const u16x2 kk = {k, k};
    return Cols4{__builtin_elementwise_min(d.e, c.e + kk), __builtin_elementwise_min(d.o, c.o + kk)};
}

// The value entering segment `seg` of a forward (dir = +1) or backward
// (dir = -1) relaxation, folded from the published ends of the segments
// behind it in sweep order: F = min(end, cin + len - 1), cin = F + 1.  The
// ends of all kMaxSeg slots are read together (slots past the segments read
// a valid entry that is not used), then folded.
template <int kMaxSeg>
__device__ __attribute__((always_inline)) inline Cols4 fold_carry(const uint32_t (*ends)[2], int seg, int nseg, int dir, int len, int q, int qn)
{
    const int behind = dir > 0 ? seg : nseg - 1 - seg;
    uint32_t e0[kMaxSeg], e1[kMaxSeg];
#pragma unroll
    for (int k = 0; k < kMaxSeg; k++) {
        const int s2 = max(0, dir > 0 ? k : nseg - 1 - k);
        e0[k] = ends[s2 * qn + q][0];
        e1[k] = ends[s2 * qn + q][1];
    }
    const unsigned short big = 0x4000;
    const u16x2 one = {1, 1};
    Cols4 cin{u16x2{big, big}, u16x2{big, big}};
#pragma unroll
    for (int k = 0; k < kMaxSeg; k++) {
        Cols4 f = cols4_relax(Cols4{__builtin_bit_cast(u16x2, e0[k]), __builtin_bit_cast(u16x2, e1[k])}, cin,
                              (unsigned short)(len - 1));
        f.e = f.e + one;
        f.o = f.o + one;
        if (k < behind) cin = f;
    }
    return cin;
}

// One segment's share of a column relaxation d[i] = min(d[i], d[i -+ 1] + 1)
// over n cells of LDS dwords p[0], p[stride], ... (4 columns per dword),
// forward (dir = +1) or backward (dir = -1): first the segment's own chain,
// its end published in ends[seg]; then, after the block barrier, the value
// entering the segment from the segments before it (fold_carry) and
// min(cell, cin + offset) over the segment -- what the sequential sweep
// gives, since its result is min over the cells behind of (d + distance).
// Every thread of the block calls it (it holds barriers); nseg <= kMaxSeg.
template <int kMaxSeg>
__device__ __attribute__((always_inline)) inline void relax_segments(uint32_t* p, int stride, int len, int seg, int nseg, int dir, bool active,
                                      uint32_t (*ends)[2], int q, int qn)
{
    const int first = dir > 0 ? seg * len : (seg + 1) * len - 1;
    if (active) {
        Cols4 prev = cols4_of(p[first * stride]);
        for (int k = 1; k < len; k++) {
            const int i = first + dir * k;
            prev = cols4_relax(cols4_of(p[i * stride]), prev, 1);
            p[i * stride] = cols4_pack(prev);
        }
        ends[seg * qn + q][0] = __builtin_bit_cast(uint32_t, prev.e);
        ends[seg * qn + q][1] = __builtin_bit_cast(uint32_t, prev.o);
    }
    __syncthreads();
    if (active && (dir > 0 ? seg > 0 : seg < nseg - 1)) {
        const Cols4 cin = fold_carry<kMaxSeg>(ends, seg, nseg, dir, len, q, qn);
        for (int k = 0; k < len; k++) {
            const int i = first + dir * k;
            p[i * stride] = cols4_pack(cols4_relax(cols4_of(p[i * stride]), cin, (unsigned short)k));
        }
    }
    __syncthreads();
}

// Both relaxations (forward, then backward) of one segment of exactly kLen
// cells held in registers, with the same carries between segments as
// relax_segments: the segment's cells are read once, relaxed in registers,
// and written once.  The backward chain runs before the forward carry is
// applied: over the cells at or after k, the carry's term cin + j + (j - k)
// is least at j = k, so min(cell, cin + k) may be taken after the backward
// chain as well as before it; the segment's backward end includes it.
// Every thread of the block calls it (it holds barriers); nseg <= kMaxSeg.
template <int kLen, int kMaxSeg>
__device__ __attribute__((always_inline)) inline void relax_segments_reg(uint32_t* p, int stride, int seg, int nseg, bool active,
                                          uint32_t (*ends)[2], int q, int qn)
{
    Cols4 r[kLen];
    const int first = seg * kLen;
    const unsigned short big = 0x4000;
    if (active) {
#pragma unroll
        for (int k = 0; k < kLen; k++) r[k] = cols4_of(p[(first + k) * stride]);
#pragma unroll
        for (int k = 1; k < kLen; k++) r[k] = cols4_relax(r[k], r[k - 1], 1);
        ends[seg * qn + q][0] = __builtin_bit_cast(uint32_t, r[kLen - 1].e);
        ends[seg * qn + q][1] = __builtin_bit_cast(uint32_t, r[kLen - 1].o);
    }
    __syncthreads();
    Cols4 cin_f{u16x2{big, big}, u16x2{big, big}};
    if (active) {
        cin_f = fold_carry<kMaxSeg>(ends, seg, nseg, +1, kLen, q, qn);
#pragma unroll
        for (int k = kLen - 2; k >= 0; k--) r[k] = cols4_relax(r[k], r[k + 1], 1);
    }
    __syncthreads();   // every forward end read
    if (active) {
        const Cols4 e0 = cols4_relax(r[0], cin_f, 0);
        ends[seg * qn + q][0] = __builtin_bit_cast(uint32_t, e0.e);
        ends[seg * qn + q][1] = __builtin_bit_cast(uint32_t, e0.o);
    }
    __syncthreads();
    if (active) {
        const Cols4 cin_b = fold_carry<kMaxSeg>(ends, seg, nseg, -1, kLen, q, qn);
#pragma unroll
        for (int k = 0; k < kLen; k++)
            p[(first + k) * stride] = cols4_pack(
                cols4_relax(cols4_relax(r[k], cin_f, (unsigned short)k), cin_b, (unsigned short)(kLen - 1 - k)));
    }
    __syncthreads();
}


constexpr int kDfCap = 254;

// X then Y pass over one z-slab held in LDS (slabs of at most kSlabMax
// bytes; larger slabs use df_x_kernel + df_y_kernel).
//
// X pass: the reference's forward/backward sweep along a row equals
// min(cap, distance to the nearest solid cell of the row), which one
// wavefront computes per row in registers.  A lane holds the dword of cells
// 256 c + 4 l .. + 3 of each 256-cell chunk c; the ballots of "byte k is
// solid" (k = 0..3) give, per chunk, which lanes hold a solid cell and the
// highest / lowest solid byte of each.  The nearest solid cell before the
// lane's dword is then the highest solid lane below it (its highest solid
// byte), or else the last solid cell of the chunks before; the nearest after
// it likewise; the four bytes take those or their own solid bytes.  The
// dwords of several rows are read together, so a wavefront waits on LDS
// once per batch of rows.  A row with no solid cell is cap everywhere, a row
// of solid cells 0 everywhere.
// Y pass: four adjacent columns per lane, the rows split into segments
// (relax_segments).
constexpr int kSlabMax = 128 * 1024;
constexpr int kSlabThreads = 1024;   // threads of a slab workgroup (the launch uses this)
constexpr int kNone = 1 << 20;

// Ballots of "byte k of the lane's dword is solid" over one 256-cell chunk.
struct ChunkBits {
    uint64_t b[4];
    __device__ uint64_t any() const { return b[0] | b[1] | b[2] | b[3]; }
    // per lane: its highest solid byte is odd / at least 2
    __device__ uint64_t hi_odd() const { return b[3] | (b[1] & ~b[2]); }
    __device__ uint64_t hi_ge2() const { return b[2] | b[3]; }
    // per lane: its lowest solid byte is odd / at least 2
    __device__ uint64_t lo_odd() const { return (b[1] & ~b[0]) | (b[3] & ~(b[0] | b[1] | b[2])); }
    __device__ uint64_t lo_ge2() const { return ~(b[0] | b[1]); }
};

__device__ inline ChunkBits chunk_bits(uint32_t v, bool valid)
{
    ChunkBits cb;
#pragma unroll
    for (int k = 0; k < 4; k++) cb.b[k] = __ballot(valid && ((v >> (8 * k)) & 0xffu) != 0u);
    return cb;
}

// Cell of the highest / lowest solid byte of lane `l` of chunk c (l holds one).
__device__ inline int chunk_last_cell(const ChunkBits& cb, int c, uint32_t l)
{
    return 256 * c + 4 * (int)l + 2 * (int)((cb.hi_ge2() >> l) & 1ull) + (int)((cb.hi_odd() >> l) & 1ull);
}
__device__ inline int chunk_first_cell(const ChunkBits& cb, int c, uint32_t l)
{
    return 256 * c + 4 * (int)l + 2 * (int)((cb.lo_ge2() >> l) & 1ull) + (int)((cb.lo_odd() >> l) & 1ull);
}

// kChunks: 256-cell chunks of a row (nx <= 256 * kChunks).  Two workgroups
// per CU: at most 64 VGPRs.
template <int kChunks>
__global__ void __launch_bounds__(kSlabThreads) __attribute__((amdgpu_waves_per_eu(8))) df_xy_slab_kernel(
    const uint8_t* __restrict__ grid, uint8_t* __restrict__ out, int nx, int ny, int nz, const int32_t* __restrict__ gate)
{
    if (gate && !*gate) return;
    extern __shared__ uint8_t slab[];
    const int z = blockIdx.x;
    const size_t base = (size_t)z * nx * ny;
    const int bytes = nx * ny;
    // unrolled so that a wavefront has several loads in flight per wait
#pragma unroll 4
    for (int i = threadIdx.x * 16; i < bytes; i += kSlabThreads * 16)
        *reinterpret_cast<uint4*>(slab + i) = *reinterpret_cast<const uint4*>(grid + base + i);
    __syncthreads();
    const int cap = min(kDfCap, nx + ny + nz);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, waves = kSlabThreads / 64;
    // ---- X pass, one wavefront per row, rows in batches -----------------
    constexpr int kBatch = 16 / kChunks;                        // rows whose dwords are read together
    const uint64_t below = (1ull << lane) - 1ull;               // lanes under this one
    const uint64_t above = lane == 63 ? 0ull : ~0ull << (lane + 1);
    for (int y0 = wave; y0 < ny; y0 += waves * kBatch) {
        uint32_t v[kBatch][kChunks];
#pragma unroll
        for (int r = 0; r < kBatch; r++)
#pragma unroll
            for (int c = 0; c < kChunks; c++) {
                const int y = y0 + r * waves, x0 = 256 * c + 4 * lane;
                v[r][c] = y < ny && x0 < nx ? *reinterpret_cast<const uint32_t*>(slab + y * nx + x0) : 0u;
            }
#pragma unroll
        for (int r = 0; r < kBatch; r++) {
            const int y = y0 + r * waves;
            if (y >= ny) break;
            uint8_t* row = slab + y * nx;
            bool any_solid = false, any_air = false;
#pragma unroll
            for (int c = 0; c < kChunks; c++) {
                if (256 * c + 4 * lane < nx) {
                    const uint32_t w = v[r][c];
                    any_solid = any_solid || w != 0u;
                    any_air = any_air || (((w - 0x01010101u) & ~w & 0x80808080u) != 0u);   // a zero byte
                }
            }
            const bool solid_row = __ballot(any_solid) != 0ull;
            if (!solid_row || __ballot(any_air) == 0ull) {
                const uint32_t fill = solid_row ? 0u : (uint32_t)cap * 0x01010101u;
#pragma unroll
                for (int c = 0; c < kChunks; c++)
                    if (256 * c + 4 * lane < nx) *reinterpret_cast<uint32_t*>(row + 256 * c + 4 * lane) = fill;
                continue;
            }
            // last / first solid cell of each chunk (scalars)
            int last_of[kChunks], first_of[kChunks];
#pragma unroll
            for (int c = 0; c < kChunks; c++) {
                const ChunkBits cb = chunk_bits(v[r][c], 256 * c + 4 * lane < nx);
                const uint64_t a = cb.any();
                last_of[c] = a ? chunk_last_cell(cb, c, 63u - (uint32_t)__clzll((long long)a)) : -kNone;
                first_of[c] = a ? chunk_first_cell(cb, c, (uint32_t)__ffsll((long long)a) - 1u) : kNone;
            }
#pragma unroll
            for (int c = 0; c < kChunks; c++) {
                const int x0 = 256 * c + 4 * lane;
                const uint32_t w = v[r][c];
                const ChunkBits cb = chunk_bits(w, x0 < nx);
                int prev_cell = -kNone, next_cell = kNone;   // nearest solid cells outside the lane's dword
#pragma unroll
                for (int c2 = 0; c2 < kChunks; c2++) {
                    if (c2 < c) prev_cell = max(prev_cell, last_of[c2]);
                    if (c2 > c) next_cell = min(next_cell, first_of[c2]);
                }
                const uint64_t mb = cb.any() & below, ma = cb.any() & above;
                if (mb) prev_cell = chunk_last_cell(cb, c, 63u - (uint32_t)__clzll((long long)mb));
                if (ma) next_cell = chunk_first_cell(cb, c, (uint32_t)__ffsll((long long)ma) - 1u);
                int before[4], after[4];
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    prev_cell = ((w >> (8 * k)) & 0xffu) ? x0 + k : prev_cell;
                    before[k] = prev_cell;
                }
#pragma unroll
                for (int k = 3; k >= 0; k--) {
                    next_cell = ((w >> (8 * k)) & 0xffu) ? x0 + k : next_cell;
                    after[k] = next_cell;
                }
                uint32_t o = 0u;
#pragma unroll
                for (int k = 0; k < 4; k++)
                    o |= (uint32_t)min(cap, min(x0 + k - before[k], after[k] - x0 - k)) << (8 * k);
                if (x0 < nx) *reinterpret_cast<uint32_t*>(row + x0) = o;
            }
        }
    }
    __syncthreads();
    // ---- Y pass (ManhattanDistanceY.comp:103-118): four adjacent columns per
    // lane (one LDS dword per row), the rows split into segments
    // (relax_segments); every thread takes part in its barriers
    {
        __shared__ uint32_t s_end_y[kSlabThreads][2];
        const int qn = nx / 4;
        int nseg = 8;
        if (ny % 16 == 0 && ny / 16 <= 8 && qn * (ny / 16) <= kSlabThreads) {
            nseg = ny / 16;   // 16-row segments, relaxed in registers
        } else {
            while (nseg > 1 && (qn * nseg > kSlabThreads || ny % nseg != 0)) nseg >>= 1;
        }
        // uniform values in scalar registers, so the segment loops branch on them
        nseg = __builtin_amdgcn_readfirstlane(nseg);
        const int len = __builtin_amdgcn_readfirstlane(ny / nseg);
        const bool active = (int)threadIdx.x < qn * nseg;
        const int q = active ? (int)threadIdx.x % qn : 0, sg = active ? (int)threadIdx.x / qn : 0;
        uint32_t* col = reinterpret_cast<uint32_t*>(slab) + q;
        if (len == 16) {
            relax_segments_reg<16, 8>(col, qn, sg, nseg, active, s_end_y, q, qn);
        } else {
            relax_segments<8>(col, qn, len, sg, nseg, +1, active, s_end_y, q, qn);
            relax_segments<8>(col, qn, len, sg, nseg, -1, active, s_end_y, q, qn);
        }
    }
    for (int i = threadIdx.x * 16; i < bytes; i += kSlabThreads * 16)
        *reinterpret_cast<uint4*>(out + base + i) = *reinterpret_cast<const uint4*>(slab + i);
}

// Z pass over a TW x 8 x nz column tile held in LDS (TW = 16 or 32; tiles
// of at most kZTileMax bytes; larger ones use df_z_kernel): the forward and
// backward relaxations along z (ManhattanDistanceZ.comp:151-165) of the
// tile's 8 * TW columns, four adjacent columns per lane (one LDS dword per
// z) and each column's z range split into 512 / (2 * TW) segments
// (relax_segments); then the tile -- TW / 8 whole 8x8 bricks wide and one
// brick high -- is written to the bricked DF with 16-byte stores.
constexpr int kZTileMax = 128 * 1024;
constexpr int kZThreads = 512;   // threads of a 16-wide tile

template <int TW, int NT>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(6))) df_z_tile_kernel(const uint8_t* __restrict__ in, uint8_t* __restrict__ df_brick,
                                                        int nx, int ny, int nz, uint32_t row_stride,
                                                        uint32_t slab_stride, const int32_t* __restrict__ gate)
{
    static_assert(TW == 16 || TW == 32, "tile width: two or four bricks");
    constexpr int kRow = 8 * TW;               // bytes of one z layer of the tile
    constexpr int kQ = kRow / 4;               // dwords of a layer: the block's columns, four per dword
    constexpr int kSeg = NT / kQ;              // z segments per column
    constexpr int kVec = TW / 16;              // 16-byte groups per row
    if (gate && !*gate) return;
    extern __shared__ uint8_t tile[];   // [z][y 0..7][x 0..TW-1]
    const int tiles_x = nx / TW;
    // consecutive tiles on one XCD: the tiles that share a 128-byte line of
    // a row meet in the same L2
    const uint32_t t = vpt::xcd_remap(blockIdx.x, gridDim.x);
    const int x0 = (int)(t % (uint32_t)tiles_x) * TW, y0 = (int)(t / (uint32_t)tiles_x) * 8;
    const size_t plane = (size_t)nx * ny;
    // stage: 8 rows of TW bytes per z, 16-byte loads; unrolled so that a
    // wavefront has several loads in flight per wait
#pragma unroll 8
    for (int i = threadIdx.x; i < nz * 8 * kVec; i += NT) {
        const int z = i / (8 * kVec), r = (i / kVec) & 7, h = i % kVec;
        *reinterpret_cast<uint4*>(tile + z * kRow + r * TW + h * 16) =
            *reinterpret_cast<const uint4*>(in + (size_t)z * plane + (size_t)(y0 + r) * nx + x0 + h * 16);
    }
    __syncthreads();
    {
        __shared__ uint32_t s_end[NT][2];
        const int q = (int)(threadIdx.x % kQ), sg = (int)(threadIdx.x / kQ);   // dword of a z layer, segment
        uint32_t* col = reinterpret_cast<uint32_t*>(tile) + q;
        // segments of 8, 16 or 24 layers are relaxed in registers
        if (nz / kSeg == 24) {
            relax_segments_reg<24, kSeg>(col, kQ, sg, kSeg, true, s_end, q, kQ);
        } else if (nz / kSeg == 16) {
            relax_segments_reg<16, kSeg>(col, kQ, sg, kSeg, true, s_end, q, kQ);
        } else if (nz / kSeg == 8) {
            relax_segments_reg<8, kSeg>(col, kQ, sg, kSeg, true, s_end, q, kQ);
        } else {
            relax_segments<kSeg>(col, kQ, nz / kSeg, sg, kSeg, +1, true, s_end, q, kQ);
            relax_segments<kSeg>(col, kQ, nz / kSeg, sg, kSeg, -1, true, s_end, q, kQ);
        }
    }
    // write whole bricks: brick (bx, bz) of this tile, 512 bytes each, one
    // 16-byte group per lane: bytes m..m+15 of the Morton order
    const int nbz = nz / 8;
    for (int i = threadIdx.x; i < (TW / 8) * nbz * 32; i += NT) {
        const int g = i & 31, brick = i >> 5;       // 32 groups of 16 bytes per brick
        const int bxl = brick % (TW / 8), bz = brick / (TW / 8);
        uint32_t words[4] = {0u, 0u, 0u, 0u};
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const uint32_t m = (uint32_t)(g * 16 + k);   // Morton index in the brick
            const uint32_t lx = (m & 1u) | ((m >> 2) & 2u) | ((m >> 4) & 4u);
            const uint32_t ly = ((m >> 1) & 1u) | ((m >> 3) & 2u) | ((m >> 5) & 4u);
            const uint32_t lz = ((m >> 2) & 1u) | ((m >> 4) & 2u) | ((m >> 6) & 4u);
            const uint32_t b = tile[(bz * 8 + lz) * kRow + ly * TW + bxl * 8 + lx];
            words[k >> 2] |= b << (8 * (k & 3));
        }
        const size_t dst = (size_t)((x0 >> 3) + bxl) * vpt::kBrickBytes + (size_t)(y0 >> 3) * row_stride +
                           (size_t)bz * slab_stride + (size_t)g * 16;
        *reinterpret_cast<uint4*>(df_brick + dst) = make_uint4(words[0], words[1], words[2], words[3]);
    }
}

// X pass in global memory (slabs too large for LDS): one lane per row.
__global__ void df_x_kernel(const uint8_t* __restrict__ grid, uint8_t* __restrict__ out, int nx, int ny, int nz,
                            const int32_t* __restrict__ gate)
{
    if (gate && !*gate) return;
    size_t row_id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row_id >= (size_t)ny * nz) return;
    const int cap = min(kDfCap, nx + ny + nz);
    const uint8_t* g = grid + row_id * nx;
    uint8_t* o = out + row_id * nx;
    int prev = g[0] > 0 ? 0 : cap;
    o[0] = (uint8_t)prev;
    for (int x = 1; x < nx; x++) {
        int v = g[x] > 0 ? 0 : min(cap, 1 + prev);
        o[x] = (uint8_t)v;
        prev = v;
    }
    int next = prev;
    for (int x = nx - 2; x >= 0; x--) {
        int v = o[x];
        if (next < v) v = 1 + next;
        o[x] = (uint8_t)v;
        next = v;
    }
}

// Y pass in global memory: one lane per (x, z) column, lanes along x.
__global__ void df_y_kernel(uint8_t* __restrict__ buf, int nx, int ny, int nz, const int32_t* __restrict__ gate)
{
    if (gate && !*gate) return;
    size_t col = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= (size_t)nx * nz) return;
    int x = (int)(col % nx), z = (int)(col / nx);
    uint8_t* p = buf + (size_t)z * nx * ny + x;
    int prev = p[0];
    for (int y = 1; y < ny; y++) {
        int v = p[(size_t)y * nx];
        if (prev < v) v = 1 + prev;
        p[(size_t)y * nx] = (uint8_t)v;
        prev = v;
    }
    int next = prev;
    for (int y = ny - 2; y >= 0; y--) {
        int v = p[(size_t)y * nx];
        if (next < v) v = 1 + next;
        p[(size_t)y * nx] = (uint8_t)v;
        next = v;
    }
}

// Z pass (ManhattanDistanceZ.comp:151-165): one lane per (x, y) column.
// The final value of every cell is known during the backward sweep and is
// written straight into the bricked DF.
__global__ void df_z_kernel(uint8_t* __restrict__ buf, uint8_t* __restrict__ df_brick, int nx, int ny, int nz,
                            uint32_t row_stride, uint32_t slab_stride, const int32_t* __restrict__ gate)
{
    if (gate && !*gate) return;
    size_t col = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= (size_t)nx * ny) return;
    int x = (int)(col % nx), y = (int)(col / nx);
    const size_t plane = (size_t)nx * ny;
    uint8_t* p = buf + (size_t)y * nx + x;
    int prev = p[0];
    for (int z = 1; z < nz; z++) {
        int v = p[(size_t)z * plane];
        if (prev < v) v = 1 + prev;
        p[(size_t)z * plane] = (uint8_t)v;
        prev = v;
    }
    int next = prev;
    df_brick[vpt::brick_offset(x, y, nz - 1, row_stride, slab_stride)] = (uint8_t)next;
    for (int z = nz - 2; z >= 0; z--) {
        int v = p[(size_t)z * plane];
        if (next < v) v = 1 + next;
        df_brick[vpt::brick_offset(x, y, z, row_stride, slab_stride)] = (uint8_t)v;
        next = v;
    }
}

int check_world(const vpt_world* w, const char* fn)
{
    if (!w) return set_error(VPT_EINVAL, "%s: null world", fn);
    return VPT_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" VPT_API int vpt_ctx_create(int device, vpt_ctx** out)
{
    if (!out) return set_error(VPT_EINVAL, "vpt_ctx_create: null output");
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count));
    if (device < 0 || device >= count)
        return set_error(VPT_EINVAL, "vpt_ctx_create: device %d out of range (%d devices)", device, count);
    DeviceGuard g(device);
    HIP_TRY(hipFree(nullptr));  // initialise the device context
    if (int rc = vpt::gi_noalloc_init(device)) return rc;
    vpt_ctx* c = new vpt_ctx();
    c->device = device;
    *out = c;
    return VPT_OK;
}

extern "C" VPT_API int vpt_ctx_destroy(vpt_ctx* ctx)
{
    if (!ctx) return VPT_OK;
    DeviceGuard g(ctx->device);
    for (vpt::GiScratch& sc : ctx->gi_scratch) {
        if (sc.done) (void)hipEventDestroy((hipEvent_t)sc.done);
        (void)hipFree(sc.d_buf);
    }
    for (void* p : ctx->gi_retired) (void)hipFree(p);
    if (ctx->side_stream) (void)hipStreamDestroy((hipStream_t)ctx->side_stream);
    if (ctx->ev_fork) (void)hipEventDestroy((hipEvent_t)ctx->ev_fork);
    if (ctx->ev_join) (void)hipEventDestroy((hipEvent_t)ctx->ev_join);
    delete ctx;
    return VPT_OK;
}

extern "C" VPT_API int vpt_ctx_device(const vpt_ctx* ctx, int* device)
{
    if (!ctx || !device) return set_error(VPT_EINVAL, "vpt_ctx_device: null argument");
    *device = ctx->device;
    return VPT_OK;
}

extern "C" VPT_API int vpt_world_create(vpt_ctx* ctx, int nx, int ny, int nz, vpt_world** out)
{
    if (!ctx || !out) return set_error(VPT_EINVAL, "vpt_world_create: null argument");
    if (nx < 16 || ny < 8 || nz < 8 || nx > 1024 || ny > 1024 || nz > 1024 || nx % 16 || ny % 8 || nz % 8)
        return set_error(VPT_EINVAL,
                         "vpt_world_create: dims %dx%dx%d invalid (nx multiple of 16, ny/nz multiples of 8, <= 1024)",
                         nx, ny, nz);
    DeviceGuard g(ctx->device);
    vpt_world* w = new vpt_world();
    w->ctx = ctx;
    w->device = ctx->device;
    w->nx = nx;
    w->ny = ny;
    w->nz = nz;
    w->nbx = nx / 8;
    w->nby = ny / 8;
    w->nbz = nz / 8;
    size_t v = w->voxels();
    hipError_t e = hipMalloc(&w->d_grid_lin, v);
    if (e == hipSuccess) e = hipMalloc(&w->d_grid, v);
    if (e == hipSuccess) e = hipMalloc(&w->d_df, v);
    if (e == hipSuccess) e = hipMalloc(&w->d_tmp, v);
    if (e == hipSuccess) e = hipMalloc(&w->d_dfbox, 8 * sizeof(int32_t));
    if (e != hipSuccess) {
        (void)hipFree(w->d_grid_lin);
        (void)hipFree(w->d_grid);
        (void)hipFree(w->d_df);
        (void)hipFree(w->d_tmp);
        (void)hipFree(w->d_dfbox);
        delete w;
        return set_error(VPT_ENOMEM, "vpt_world_create: hipMalloc of %zu bytes failed", 4 * v);
    }
    w->bytes = 4 * v;
    HIP_TRY(hipMemset(w->d_grid_lin, 0, v));
    HIP_TRY(hipMemset(w->d_grid, 0, v));
    HIP_TRY(hipMemset(w->d_df, 0, v));
    *out = w;
    return VPT_OK;
}

extern "C" VPT_API int vpt_world_destroy(vpt_world* w)
{
    if (!w) return VPT_OK;
    DeviceGuard g(w->device);
    (void)hipFree(w->d_grid_lin);
    (void)hipFree(w->d_grid);
    (void)hipFree(w->d_df);
    (void)hipFree(w->d_tmp);
    (void)hipFree(w->d_dfbox);
    delete w;
    return VPT_OK;
}

extern "C" VPT_API int vpt_world_dims(const vpt_world* w, int* nx, int* ny, int* nz)
{
    if (int rc = check_world(w, "vpt_world_dims")) return rc;
    if (nx) *nx = w->nx;
    if (ny) *ny = w->ny;
    if (nz) *nz = w->nz;
    return VPT_OK;
}

extern "C" VPT_API int vpt_world_bytes(const vpt_world* w, size_t* bytes)
{
    if (int rc = check_world(w, "vpt_world_bytes")) return rc;
    if (!bytes) return set_error(VPT_EINVAL, "vpt_world_bytes: null output");
    *bytes = w->bytes;
    return VPT_OK;
}

static int rebrick_grid(vpt_world* w, hipStream_t s)
{
    size_t groups = (size_t)(w->nx / 16) * w->ny * w->nz;
    int threads = 256;
    int blocks = (int)((groups + threads - 1) / threads);
    uint32_t rs = (uint32_t)w->nbx * vpt::kBrickBytes, ss = rs * (uint32_t)w->nby;
    hipLaunchKernelGGL(brick_kernel, dim3(blocks), dim3(threads), 0, s, w->d_grid_lin, w->d_grid, w->nx, w->ny,
                       w->nz, rs, ss);
    HIP_TRY(hipGetLastError());
    return VPT_OK;
}

extern "C" VPT_API int vpt_world_upload(vpt_world* w, const uint8_t* host_grid, void* stream)
{
    if (int rc = check_world(w, "vpt_world_upload")) return rc;
    if (!host_grid) return set_error(VPT_EINVAL, "vpt_world_upload: null grid");
    DeviceGuard g(w->device);
    hipStream_t s = as_stream(stream);
    HIP_TRY(hipMemcpyAsync(w->d_grid_lin, host_grid, w->voxels(), hipMemcpyHostToDevice, s));
    return rebrick_grid(w, s);
}

extern "C" VPT_API int vpt_world_upload_device(vpt_world* w, const uint8_t* d_grid, void* stream)
{
    if (int rc = check_world(w, "vpt_world_upload_device")) return rc;
    if (!d_grid) return set_error(VPT_EINVAL, "vpt_world_upload_device: null grid");
    DeviceGuard g(w->device);
    hipStream_t s = as_stream(stream);
    HIP_TRY(hipMemcpyAsync(w->d_grid_lin, d_grid, w->voxels(), hipMemcpyDeviceToDevice, s));
    return rebrick_grid(w, s);
}

extern "C" VPT_API int vpt_world_set_blocks(vpt_world* w, const uint32_t* edits, const uint8_t* ids, int n,
                                            void* stream)
{
    if (int rc = check_world(w, "vpt_world_set_blocks")) return rc;
    if (n < 0 || (n > 0 && (!edits || !ids))) return set_error(VPT_EINVAL, "vpt_world_set_blocks: bad edit list");
    if (n == 0) return VPT_OK;
    for (int i = 0; i < n; i++) {
        uint32_t e = edits[i];
        int x = (int)(e & 1023u), y = (int)((e >> 10) & 1023u), z = (int)((e >> 20) & 1023u);
        if (x >= w->nx || y >= w->ny || z >= w->nz || (e >> 30))
            return set_error(VPT_EINVAL, "vpt_world_set_blocks: edit %d at (%d,%d,%d) outside the world", i, x, y, z);
    }
    // The kernels write all edits at once; a voxel named twice keeps the
    // last id, as sequential SetBlock calls do.
    std::vector<uint32_t> uniq_e;
    std::vector<uint8_t> uniq_i;
    {
        std::unordered_set<uint32_t> seen;
        seen.reserve((size_t)n * 2);
        for (int i = n - 1; i >= 0; i--)
            if (seen.insert(edits[i]).second) {
                uniq_e.push_back(edits[i]);
                uniq_i.push_back(ids[i]);
            }
        if ((int)uniq_e.size() != n) {
            edits = uniq_e.data();
            ids = uniq_i.data();
            n = (int)uniq_e.size();
        }
    }
    DeviceGuard g(w->device);
    hipStream_t s = as_stream(stream);
    if (n <= kArgEdits) {
        ArgEdits ed;
        std::copy(edits, edits + n, ed.edit);
        std::copy(ids, ids + n, ed.id);
        const uint32_t rs = (uint32_t)w->nbx * vpt::kBrickBytes, ss = rs * (uint32_t)w->nby;
        hipLaunchKernelGGL(set_blocks_args_kernel, dim3(1), dim3(kArgEdits), 0, s, w->d_grid_lin, w->d_grid, ed, n,
                           w->nx, w->ny, rs, ss);
        HIP_TRY(hipGetLastError());
        return VPT_OK;
    }
    uint32_t* d_e = nullptr;
    uint8_t* d_i = nullptr;
    HIP_TRY(hipMallocAsync(reinterpret_cast<void**>(&d_e), sizeof(uint32_t) * n, s));
    HIP_TRY(hipMallocAsync(reinterpret_cast<void**>(&d_i), n, s));
    HIP_TRY(hipMemcpyAsync(d_e, edits, sizeof(uint32_t) * n, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_i, ids, n, hipMemcpyHostToDevice, s));
    uint32_t rs = (uint32_t)w->nbx * vpt::kBrickBytes, ss = rs * (uint32_t)w->nby;
    hipLaunchKernelGGL(set_blocks_kernel, dim3((n + 255) / 256), dim3(256), 0, s, w->d_grid_lin, w->d_grid, d_e, d_i,
                       n, w->nx, w->ny, w->nz, rs, ss);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipFreeAsync(d_e, s));
    HIP_TRY(hipFreeAsync(d_i, s));
    return VPT_OK;
}

namespace {

// GenerateDistanceField's three passes (World.cpp:69-113).  With a gate
// pointer every kernel returns at once unless *gate is nonzero, so a caller
// can enqueue the rebuild behind a decision taken on the device.
int df_build_launch(vpt_world* w, hipStream_t s, const int32_t* gate)
{
    const int nx = w->nx, ny = w->ny, nz = w->nz;
    const size_t slab = (size_t)nx * ny;
    if (slab <= (size_t)kSlabMax) {
        const int chunks = (nx + 255) / 256;
        auto kern = chunks == 1   ? df_xy_slab_kernel<1>
                    : chunks == 2 ? df_xy_slab_kernel<2>
                    : chunks == 3 ? df_xy_slab_kernel<3>
                                  : df_xy_slab_kernel<4>;
        HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                    kSlabMax));
        hipLaunchKernelGGL(kern, dim3(nz), dim3(kSlabThreads), slab, s, w->d_grid_lin, w->d_tmp, nx, ny, nz, gate);
        HIP_TRY(hipGetLastError());
    } else {
        size_t rows = (size_t)ny * nz;
        hipLaunchKernelGGL(df_x_kernel, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, s, w->d_grid_lin,
                           w->d_tmp, nx, ny, nz, gate);
        HIP_TRY(hipGetLastError());
        size_t cols = (size_t)nx * nz;
        hipLaunchKernelGGL(df_y_kernel, dim3((unsigned)((cols + 255) / 256)), dim3(256), 0, s, w->d_tmp, nx, ny, nz,
                           gate);
        HIP_TRY(hipGetLastError());
    }
    uint32_t rs = (uint32_t)w->nbx * vpt::kBrickBytes, ss = rs * (uint32_t)w->nby;
    // 16-byte-wide tiles (two bricks) give 3 blocks per CU in LDS at nz =
    // 384 and twice the blocks of 32-wide ones; both need nz to split into
    // their segments (16 / 8) and the tile to fit in LDS
    if (nx % 16 == 0 && nz % (kZThreads / 32) == 0 && (size_t)nz * 128 <= (size_t)kZTileMax) {
        HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(&df_z_tile_kernel<16, kZThreads>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, kZTileMax));
        hipLaunchKernelGGL((df_z_tile_kernel<16, kZThreads>), dim3((unsigned)((nx / 16) * (ny / 8))), dim3(kZThreads),
                           (size_t)nz * 128, s, w->d_tmp, w->d_df, nx, ny, nz, rs, ss, gate);
    } else if (nx % 32 == 0 && nz % 8 == 0 && (size_t)nz * 256 <= (size_t)kZTileMax) {
        HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(&df_z_tile_kernel<32, 512>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, kZTileMax));
        hipLaunchKernelGGL((df_z_tile_kernel<32, 512>), dim3((unsigned)((nx / 32) * (ny / 8))), dim3(512),
                           (size_t)nz * 256, s, w->d_tmp, w->d_df, nx, ny, nz, rs, ss, gate);
    } else {
        size_t cols = (size_t)nx * ny;
        hipLaunchKernelGGL(df_z_kernel, dim3((unsigned)((cols + 255) / 256)), dim3(256), 0, s, w->d_tmp, w->d_df, nx,
                           ny, nz, rs, ss, gate);
    }
    HIP_TRY(hipGetLastError());
    return VPT_OK;
}

}  // namespace

extern "C" VPT_API int vpt_df_build(vpt_world* w, void* stream)
{
    if (int rc = check_world(w, "vpt_df_build")) return rc;
    DeviceGuard g(w->device);
    return df_build_launch(w, as_stream(stream), nullptr);
}

static int download_bricked(vpt_world* w, const uint8_t* d_src, uint8_t* host, hipStream_t s)
{
    size_t groups = (size_t)(w->nx / 16) * w->ny * w->nz;
    uint32_t rs = (uint32_t)w->nbx * vpt::kBrickBytes, ss = rs * (uint32_t)w->nby;
    hipLaunchKernelGGL(unbrick_kernel, dim3((unsigned)((groups + 255) / 256)), dim3(256), 0, s, d_src, w->d_tmp,
                       w->nx, w->ny, w->nz, rs, ss);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(host, w->d_tmp, w->voxels(), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return VPT_OK;
}

extern "C" VPT_API int vpt_df_download(vpt_world* w, uint8_t* host_df, void* stream)
{
    if (int rc = check_world(w, "vpt_df_download")) return rc;
    if (!host_df) return set_error(VPT_EINVAL, "vpt_df_download: null output");
    DeviceGuard g(w->device);
    return download_bricked(w, w->d_df, host_df, as_stream(stream));
}

extern "C" VPT_API int vpt_grid_download(vpt_world* w, uint8_t* host_grid, void* stream)
{
    if (int rc = check_world(w, "vpt_grid_download")) return rc;
    if (!host_grid) return set_error(VPT_EINVAL, "vpt_grid_download: null output");
    DeviceGuard g(w->device);
    // Read the bricked copy back, so the test also covers the layout.
    return download_bricked(w, w->d_grid, host_grid, as_stream(stream));
}

extern "C" VPT_API int vpt_world_set_sample_counter(vpt_world* w, unsigned long long* d_counter)
{
    if (int rc = check_world(w, "vpt_world_set_sample_counter")) return rc;
    w->d_counter = d_counter;
    return VPT_OK;
}

extern "C" VPT_API int vpt_world_save(vpt_world* w, const char* path)
{
    if (int rc = check_world(w, "vpt_world_save")) return rc;
    if (!path) return set_error(VPT_EINVAL, "vpt_world_save: null path");
    std::vector<uint8_t> host(w->voxels());
    DeviceGuard g(w->device);
    HIP_TRY(hipMemcpy(host.data(), w->d_grid_lin, host.size(), hipMemcpyDeviceToHost));
    FILE* f = fopen(path, "wb");
    if (!f) return set_error(VPT_EINVAL, "vpt_world_save: cannot open %s for writing", path);
    const size_t n = fwrite(host.data(), 1, host.size(), f);
    const int closed = fclose(f);
    if (n != host.size() || closed != 0) return set_error(VPT_EINVAL, "vpt_world_save: short write to %s", path);
    return VPT_OK;
}

extern "C" VPT_API int vpt_world_load(vpt_world* w, const char* path, void* stream)
{
    if (int rc = check_world(w, "vpt_world_load")) return rc;
    if (!path) return set_error(VPT_EINVAL, "vpt_world_load: null path");
    FILE* f = fopen(path, "rb");
    if (!f) return set_error(VPT_EINVAL, "vpt_world_load: cannot open %s", path);
    std::vector<uint8_t> host(w->voxels());
    const size_t n = fread(host.data(), 1, host.size(), f);
    const bool extra = fgetc(f) != EOF;
    fclose(f);
    if (n != host.size() || extra)
        return set_error(VPT_EINVAL, "vpt_world_load: %s does not hold exactly %zu bytes (%dx%dx%d)", path,
                         host.size(), w->nx, w->ny, w->nz);
    if (int rc = vpt_world_upload(w, host.data(), stream)) return rc;
    // the host buffer is pageable and goes out of scope: finish the copy first
    DeviceGuard g(w->device);
    HIP_TRY(hipStreamSynchronize(as_stream(stream)));
    return VPT_OK;
}

// ---- incremental DF update ---------------------------------------------------
//
// The DF holds min(cap, L1 distance to the nearest block).  After block edits
// a cell p can change only if |p - e|_1 <= D_old(p) for some edit e (an added
// block shortens p's distance only when it is nearer than D_old(p); a removed
// one lengthens it only for cells whose nearest block it was, with
// D_old(p) = |p - e| < cap).  Let B be a box around the edits' bounding box
// K such that every cell q on a face of B that is not on the world's border
// has D_old(q) < dist_1(q, K).  A changed cell p outside B would, on a
// monotone lattice path back to its edit e, enter B at such a face cell q
// with D_old(q) >= D_old(p) - |p - q| >= |p - e| - |p - q| = |q - e| >=
// dist_1(q, K) (D_old is 1-Lipschitz): impossible.  So every cell outside B
// keeps its value, and inside B
//   D_new(p) = min(cap, min over blocks s in B of |p - s|,
//                       min over shell cells q of D_old(q) + |p - q|)
// (the shell is the one-cell layer around B: a shortest path to a block
// outside B leaves B through it).  Three separable min-plus sweeps over the
// lines through B compute this: B's cells start at 0 (block) or cap, the
// shell cells keep D_old (already their swept value, D_old being
// 1-Lipschitz), and every intermediate value stays an upper bound of D_new.
//
// One workgroup does it all.  It grows B from K in steps: each step tests,
// for every face, its current position and kCands - 1 positions further
// out at geometrically growing margins ((m + 1) 2^c - 1), all at once (one
// wavefront-aligned slice of face cells per face and candidate), and moves
// the face to the nearest position that passes; a step in which every face
// passes where it stands ends the search.  B plus its shell is then staged
// in LDS, swept, and B written back.  A box whose shell-extended volume
// passes the LDS budget, or a search past kSearchSteps steps, sets a gate
// instead, and the full rebuild enqueued behind the kernel runs.

namespace {

constexpr int kSlackBias = 4096;          // > the largest L1 distance in a world
constexpr int kSearchSteps = 64;
constexpr int kCands = 8;                 // positions tested per face and step
constexpr int kSlices = 6 * kCands;
constexpr int kStepCells = 256 * 1024;    // face cells per step before only the current positions are tested
constexpr int kLocalLds = 128 * 1024;     // cells of B plus shell held in LDS, one byte each
constexpr int kShortLine = 16;            // lines up to this length: one lane each; longer: one wavefront
constexpr int kBatch = 4;                 // global loads in flight per thread
constexpr int kRow = 16;                  // face-cell slices are padded to DPP rows of 16 lanes
constexpr int kZRun = 4;                  // cells along z per thread when staging / writing back

// The world dims and the edits' bounding box, inclusive.
struct EditBox {
    int k0[3], k1[3];
    int dims[3];
};

// Inclusive wavefront scans with DPP: row shifts by 1, 2, 4, 8 within each
// row of 16 lanes (lanes shifted in from outside the row read the identity),
// then row_bcast:15 and row_bcast:31 carry rows 0 and 0-1 into the rows
// above.  No LDS traffic, unlike shuffles.
template <bool kMax>
__device__ inline int wave_scan_dpp(int v)
{
    constexpr int kId = kMax ? INT_MIN : INT_MAX;
    auto op = [](int a, int b) { return kMax ? max(a, b) : min(a, b); };
    v = op(v, __builtin_amdgcn_update_dpp(kId, v, 0x111, 0xf, 0xf, false));   // row_shr:1
    v = op(v, __builtin_amdgcn_update_dpp(kId, v, 0x112, 0xf, 0xf, false));   // row_shr:2
    v = op(v, __builtin_amdgcn_update_dpp(kId, v, 0x114, 0xf, 0xf, false));   // row_shr:4
    v = op(v, __builtin_amdgcn_update_dpp(kId, v, 0x118, 0xf, 0xf, false));   // row_shr:8
    v = op(v, __builtin_amdgcn_update_dpp(kId, v, 0x142, 0xa, 0xf, false));   // row_bcast:15 into rows 1, 3
    v = op(v, __builtin_amdgcn_update_dpp(kId, v, 0x143, 0xc, 0xf, false));   // row_bcast:31 into rows 2, 3
    return v;
}

// Exact j / n and j % n for 0 <= j < 2^20, 0 < n <= 2^11, from a float
// reciprocal: the estimate is within one of the quotient, then corrected.
struct FastDiv {
    int n;
    float inv;
    __device__ explicit FastDiv(int n_) : n(n_), inv(1.0f / (float)n_) {}
    __device__ inline int divmod(int j, int& r) const
    {
        int q = (int)((float)j * inv);
        r = j - q * n;
        if (r < 0) {
            q--;
            r += n;
        } else if (r >= n) {
            q++;
            r -= n;
        }
        return q;
    }
};

__device__ inline int cand_margin(int m, int c) { return c == 0 ? m : (m + 1) * (1 << c) - 1; }

// Min-plus sweep of one line of L cells at stride st, forward then
// backward: cell k ends at min over j of v[j] + |k - j|.
__device__ inline void sweep_line_serial(uint8_t* p, int L, int st)
{
    int run = 1 << 20;
    for (int k = 0; k < L; k++) {
        run = min((int)p[k * st], run + 1);
        p[k * st] = (uint8_t)run;
    }
    run = 1 << 20;
    for (int k = L - 1; k >= 0; k--) {
        run = min((int)p[k * st], run + 1);
        p[k * st] = (uint8_t)run;
    }
}

// The same with one wavefront per line: forward f[k] = k + min_{j<=k}(v[j] - j)
// and backward b[k] = min_{j>=k}(f[j] + j) - k, both as prefix scans over
// lanes (the backward pass maps lane l of chunk c to k = L - 1 - 64 c - l).
__device__ inline void sweep_line_wave(uint8_t* p, int L, int st)
{
    const int lane = threadIdx.x & 63;
    constexpr int kBig = 1 << 20;
    int carry = kBig;
    for (int t = 0; t < L; t += 64) {
        const int k = t + lane;
        int a = k < L ? (int)p[k * st] - k : kBig;
        a = min(wave_scan_dpp<false>(a), carry);
        if (k < L) p[k * st] = (uint8_t)(a + k);
        carry = __builtin_amdgcn_readlane(a, 63);
    }
    carry = kBig;
    for (int t = 0; t < L; t += 64) {
        const int k = L - 1 - t - lane;
        int a = k >= 0 ? (int)p[k * st] + k : kBig;
        a = min(wave_scan_dpp<false>(a), carry);
        if (k >= 0) p[k * st] = (uint8_t)(a - k);
        carry = __builtin_amdgcn_readlane(a, 63);
    }
}

// out: [0..2] box begin, [3..5] box end (half-open), [6] 1 = updated here,
// [7] 1 = full rebuild needed (the box then reads as the whole world).
__global__ void __launch_bounds__(1024) df_update_kernel(const uint8_t* __restrict__ grid_lin,
                                                        uint8_t* __restrict__ df, EditBox eb, int cap,
                                                        uint32_t row_stride, uint32_t slab_stride,
                                                        int32_t* __restrict__ out)
{
    extern __shared__ uint8_t cells[];   // during the search: slice of each row of face cells
    __shared__ int s_margin[6];
    __shared__ int s_grow[6];            // face still looking for its position: test candidates
    __shared__ unsigned int s_slack[kSlices];
    __shared__ int s_begin[kSlices + 1];
    __shared__ int s_fixed[kSlices], s_area[kSlices], s_nu[kSlices];
    __shared__ int s_state;   // 0 search on, 1 box found, 2 full rebuild
    const int tid = threadIdx.x;
    if (tid < 6) {
        s_margin[tid] = 0;
        s_grow[tid] = 1;
    }
    int b0[3], b1[3];
    for (int step = 0;; step++) {
        __syncthreads();
        unsigned long long vol = 1;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            b0[a] = max(0, eb.k0[a] - s_margin[2 * a]);
            b1[a] = min(eb.dims[a], eb.k1[a] + s_margin[2 * a + 1] + 1);
            vol *= (unsigned long long)(b1[a] - b0[a]);
        }
        const int wx = b1[0] - b0[0], wy = b1[1] - b0[1], wz = b1[2] - b0[2];
        const int face_cells = 2 * (wy * wz + wx * wz + wx * wy);
        const bool full = vol > (unsigned long long)kLocalLds || step >= kSearchSteps;
        const bool wide = face_cells * kCands <= kStepCells;   // candidates affordable this step
        if (!full) {
            // slice f * kCands + c: face f at candidate c, padded to whole DPP rows
            if (tid < 64) {
                int count = 0;
                if (tid < kSlices) {
                    const int f = tid / kCands, c = tid % kCands, a = f >> 1, side = f & 1;
                    const int k0a = a == 0 ? eb.k0[0] : (a == 1 ? eb.k0[1] : eb.k0[2]);
                    const int k1a = a == 0 ? eb.k1[0] : (a == 1 ? eb.k1[1] : eb.k1[2]);
                    const int dima = a == 0 ? eb.dims[0] : (a == 1 ? eb.dims[1] : eb.dims[2]);
                    const int nu = a == 0 ? wy : wx, nv = a == 2 ? wy : wz;
                    const int mc = cand_margin(s_margin[f], c);
                    const int pos = side ? k1a + mc : k0a - mc;
                    const bool border = side ? pos >= dima - 1 : pos <= 0;
                    const bool tested = c == 0 || (wide && s_grow[f]);
                    s_fixed[tid] = pos;
                    s_area[tid] = nu * nv;
                    s_nu[tid] = nu;
                    s_slack[tid] = 0;
                    count = border || !tested ? 0 : (nu * nv + kRow - 1) & ~(kRow - 1);
                }
                int incl = count;   // exclusive prefix of the counts
                for (int d = 1; d < 64; d <<= 1) {
                    const int t = __shfl_up(incl, d);
                    if (tid >= d) incl += t;
                }
                if (tid <= kSlices) s_begin[tid] = incl - count;
            }
            __syncthreads();
            const int total = s_begin[kSlices], rows = total / kRow;
            for (int r = tid; r < rows; r += blockDim.x) {
                const int i = r * kRow;
                int lo = 0;   // last slice with s_begin <= i
#pragma unroll
                for (int w = 32; w > 0; w >>= 1)
                    if (lo + w <= kSlices - 1 && s_begin[lo + w] <= i) lo += w;
                cells[r] = (uint8_t)lo;
            }
            __syncthreads();
            // kBatch face cells per thread in flight; rows of 16 lanes share a
            // slice; every lane of a wavefront reaches the row reductions
            const int total64 = (total + 63) & ~63;
            for (int i0 = tid; i0 < total64; i0 += kBatch * 1024) {
                int sl[kBatch], d[kBatch], dist[kBatch];
#pragma unroll
                for (int b = 0; b < kBatch; b++) {
                    const int i = i0 + b * 1024;
                    sl[b] = i < total ? (int)cells[i / kRow] : -1;
                }
#pragma unroll
                for (int b = 0; b < kBatch; b++) {
                    d[b] = 0;
                    dist[b] = 1 << 20;   // padding lanes: slack far below zero
                    const int lo = sl[b];
                    if (lo >= 0) {
                        const int j = i0 + b * 1024 - s_begin[lo];
                        if (j < s_area[lo]) {
                            const int f = lo / kCands, fix = s_fixed[lo];
                            int ju;
                            const int jv = FastDiv(s_nu[lo]).divmod(j, ju);
                            const int x = f < 2 ? fix : b0[0] + ju;
                            const int y = f < 2 ? b0[1] + ju : (f < 4 ? fix : b0[1] + jv);
                            const int z = f < 4 ? b0[2] + jv : fix;
                            dist[b] = max(0, max(eb.k0[0] - x, x - eb.k1[0])) +
                                      max(0, max(eb.k0[1] - y, y - eb.k1[1])) +
                                      max(0, max(eb.k0[2] - z, z - eb.k1[2]));
                            d[b] = df[vpt::brick_offset(x, y, z, row_stride, slab_stride)];
                        }
                    }
                }
#pragma unroll
                for (int b = 0; b < kBatch; b++) {
                    if (i0 - (tid & 63) + b * 1024 >= total64) continue;   // uniform per wavefront
                    int m = max(0, d[b] - dist[b] + kSlackBias);
                    m = max(m, __builtin_amdgcn_update_dpp(0, m, 0x111, 0xf, 0xf, false));   // row_shr:1
                    m = max(m, __builtin_amdgcn_update_dpp(0, m, 0x112, 0xf, 0xf, false));   // row_shr:2
                    m = max(m, __builtin_amdgcn_update_dpp(0, m, 0x114, 0xf, 0xf, false));   // row_shr:4
                    m = max(m, __builtin_amdgcn_update_dpp(0, m, 0x118, 0xf, 0xf, false));   // row_shr:8
                    if ((tid & 15) == 15 && sl[b] >= 0 && m) atomicMax(&s_slack[sl[b]], (unsigned int)m);
                }
            }
        }
        __syncthreads();
        if (tid == 0) s_state = full ? 2 : 1;
        __syncthreads();
        if (!full && tid < 6) {
            // a slice passes when it has no cells (world border) or every slack is negative
            const int f = tid, m0 = s_margin[f];
            const bool tested = wide && s_grow[f];
            auto passes = [&](int c) { return s_slack[f * kCands + c] < (unsigned int)kSlackBias; };
            if (passes(0)) {
                s_grow[f] = 0;
            } else {
                int next = -1;
                for (int c = 1; tested && c < kCands && next < 0; c++)
                    if (passes(c)) next = cand_margin(m0, c);
                s_grow[f] = next < 0;
                if (next < 0) {
                    const int v = (int)s_slack[f * kCands] - kSlackBias;
                    next = max(max(m0 + v + 1, 2 * m0), tested ? cand_margin(m0, kCands - 1) + 1 : 0);
                }
                s_margin[f] = next;
                s_state = 0;   // every writer stores the same value
            }
        }
        __syncthreads();
        if (s_state) break;
    }
    // B plus its shell, clipped to the world
    int e0[3], e1[3];
    for (int a = 0; a < 3; a++) {
        e0[a] = max(0, b0[a] - 1);
        e1[a] = min(eb.dims[a], b1[a] + 1);
    }
    const int ex = e1[0] - e0[0], ey = e1[1] - e0[1], ez = e1[2] - e0[2];
    const bool local = s_state == 1 && (long long)ex * ey * ez <= kLocalLds;
    if (tid == 0) {
        for (int a = 0; a < 3; a++) {
            out[a] = local ? b0[a] : 0;
            out[3 + a] = local ? b1[a] : eb.dims[a];
        }
        out[6] = local;
        out[7] = !local;
    }
    if (!local) return;   // uniform: every thread computed the same box
    const int nx = eb.dims[0], ny = eb.dims[1];
    // seeds: B's cells 0 (block) or cap, shell cells their old distance.
    // One item = kZRun cells along z at one (x, y): the x-y part of both
    // addresses is computed once, and all loads of the item are in flight.
    const int zruns = (ez + kZRun - 1) / kZRun;
    const int items = ex * ey * zruns;
    const FastDiv div_ex(ex), div_ey(ey);
    const size_t plane = (size_t)nx * ny;
    for (int it = tid; it < items; it += blockDim.x) {
        int xl, yl;
        const int zr = div_ey.divmod(div_ex.divmod(it, xl), yl);
        const int x = e0[0] + xl, y = e0[1] + yl;
        const bool in_xy = x >= b0[0] && x < b1[0] && y >= b0[1] && y < b1[1];
        const uint8_t* g = grid_lin + (size_t)y * nx + x;
        const uint8_t* o = df + vpt::brick_offset(x, y, 0, row_stride, slab_stride);
        uint8_t blk[kZRun], old[kZRun];
#pragma unroll
        for (int k = 0; k < kZRun; k++) {
            const int z = e0[2] + min(zr * kZRun + k, ez - 1);
            blk[k] = g[(size_t)z * plane];
            old[k] = o[vpt::brick_offset(0, 0, z, row_stride, slab_stride)];
        }
#pragma unroll
        for (int k = 0; k < kZRun; k++) {
            const int zl = zr * kZRun + k, z = e0[2] + zl;
            if (zl < ez)
                cells[(zl * ey + yl) * ex + xl] =
                    in_xy && z >= b0[2] && z < b1[2] ? (blk[k] > 0 ? 0 : (uint8_t)cap) : old[k];
        }
    }
    __syncthreads();
    // sweeps along x, y, z over the lines through B (E-local coordinates)
    const int lo[3] = {b0[0] - e0[0], b0[1] - e0[1], b0[2] - e0[2]};
    const int len[3] = {ex, ey, ez};
    const int stride[3] = {1, ex, ex * ey};
    const int wave = tid >> 6, waves = blockDim.x >> 6;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const int u = a == 0 ? 1 : 0, v = a == 2 ? 1 : 2;
        const int nu = b1[u] - b0[u], lines = nu * (b1[v] - b0[v]);
        if (len[a] <= kShortLine) {
            for (int i = tid; i < lines; i += blockDim.x)
                sweep_line_serial(cells + (lo[u] + i % nu) * stride[u] + (lo[v] + i / nu) * stride[v], len[a],
                                  stride[a]);
        } else {
            for (int i = wave; i < lines; i += waves)
                sweep_line_wave(cells + (lo[u] + i % nu) * stride[u] + (lo[v] + i / nu) * stride[v], len[a],
                                stride[a]);
        }
        __syncthreads();
    }
    const int bx = b1[0] - b0[0], by = b1[1] - b0[1], bz = b1[2] - b0[2];
    const int bruns = (bz + kZRun - 1) / kZRun;
    const FastDiv div_bx(bx), div_by(by);
    for (int it = tid; it < bx * by * bruns; it += blockDim.x) {
        int xl, yl;
        const int zr = div_by.divmod(div_bx.divmod(it, xl), yl);
        const int x = b0[0] + xl, y = b0[1] + yl;
        uint8_t* o = df + vpt::brick_offset(x, y, 0, row_stride, slab_stride);
        const uint8_t* c = cells + ((b0[2] - e0[2]) * ey + (y - e0[1])) * ex + (x - e0[0]);
#pragma unroll
        for (int k = 0; k < kZRun; k++) {
            const int zl = zr * kZRun + k;
            if (zl < bz) o[vpt::brick_offset(0, 0, b0[2] + zl, row_stride, slab_stride)] = c[zl * ex * ey];
        }
    }
}

}  // namespace

extern "C" VPT_API int vpt_df_update(vpt_world* w, const uint32_t* edits, int n, int* region, void* stream)
{
    if (int rc = check_world(w, "vpt_df_update")) return rc;
    if (n < 0 || (n > 0 && !edits)) return set_error(VPT_EINVAL, "vpt_df_update: bad edit list");
    if (n == 0) {
        if (region) std::fill(region, region + 6, 0);
        return VPT_OK;
    }
    const int nx = w->nx, ny = w->ny, nz = w->nz;
    EditBox eb;
    eb.dims[0] = nx;
    eb.dims[1] = ny;
    eb.dims[2] = nz;
    for (int a = 0; a < 3; a++) {
        eb.k0[a] = eb.dims[a];
        eb.k1[a] = -1;
    }
    for (int i = 0; i < n; i++) {
        const int c[3] = {(int)(edits[i] & 1023u), (int)((edits[i] >> 10) & 1023u), (int)((edits[i] >> 20) & 1023u)};
        if (c[0] >= nx || c[1] >= ny || c[2] >= nz || (edits[i] >> 30))
            return set_error(VPT_EINVAL, "vpt_df_update: edit %d at (%d,%d,%d) outside the world", i, c[0], c[1], c[2]);
        for (int a = 0; a < 3; a++) {
            eb.k0[a] = std::min(eb.k0[a], c[a]);
            eb.k1[a] = std::max(eb.k1[a], c[a]);
        }
    }
    DeviceGuard g(w->device);
    hipStream_t s = as_stream(stream);
    const uint32_t rs = (uint32_t)w->nbx * vpt::kBrickBytes, ss = rs * (uint32_t)w->nby;
    const int cap = std::min(kDfCap, nx + ny + nz);
    HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(&df_update_kernel),
                                hipFuncAttributeMaxDynamicSharedMemorySize, kLocalLds));
    hipLaunchKernelGGL(df_update_kernel, dim3(1), dim3(1024), kLocalLds, s, w->d_grid_lin, w->d_df, eb, cap, rs, ss,
                       w->d_dfbox);
    HIP_TRY(hipGetLastError());
    if (int rc = df_build_launch(w, s, w->d_dfbox + 7)) return rc;
    if (region) {
        int32_t box[8];
        HIP_TRY(hipMemcpyAsync(box, w->d_dfbox, sizeof(box), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        std::copy(box, box + 6, region);
    }
    return VPT_OK;
}
