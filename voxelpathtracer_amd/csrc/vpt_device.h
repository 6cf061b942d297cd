// Device-side building blocks shared by the gfx950 kernels: brick
// addressing, the DF traversal, ray generation.  Every floating-point
// expression here is the same IEEE operation sequence as the reference
// shader it cites (compiled with -ffp-contract=off), which is what makes hit
// cells, faces and distances reproducible bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vpt/vpt_numerics.h"

namespace vpt {

constexpr int kBrick = 8;
constexpr int kBrickBytes = 512;

// Spread the three bits of v (0..7) to positions 0, 3, 6.
__host__ __device__ inline uint32_t spread3(uint32_t v) { return (v & 1u) | ((v & 2u) << 2) | ((v & 4u) << 4); }

// Read-only view of a world's bricked planes, passed to kernels by value.
struct WorldView {
    const uint8_t* grid;
    const uint8_t* df;
    int nx, ny, nz;
    uint32_t row_stride;    // nbx * 512
    uint32_t slab_stride;   // nbx * nby * 512
    float xmax, ymax, zmax; // dims - 1 as float (IsInVolume bounds)
};

// Byte `off` of a world plane (grid or DF).  The planes are device
// allocations, so the read is a global one also where the view was loaded
// from memory (a served world) rather than passed as a kernel argument,
// which would otherwise make it a flat access.
__device__ inline uint8_t plane_byte(const uint8_t* p, uint32_t off)
{
    typedef const __attribute__((address_space(1))) uint8_t GlobalByte;
    return ((GlobalByte*)p)[off];   // address-space cast
}

// Byte offset of cell (x, y, z) in a bricked plane: the three axes
// contribute independent terms.
__host__ __device__ inline uint32_t brick_offset(uint32_t x, uint32_t y, uint32_t z, uint32_t row_stride,
                                                 uint32_t slab_stride)
{
    return ((x >> 3) * kBrickBytes + spread3(x & 7u)) + ((y >> 3) * row_stride + (spread3(y & 7u) << 1)) +
           ((z >> 3) * slab_stride + (spread3(z & 7u) << 2));
}

// brick_offset for the traversal's hot loop, same value: the Morton spread
// of each axis's low three bits is one v_perm_b32 byte lookup in an 8-entry
// table held in its two source operands (byte k of {hi, lo} = spread of k),
// and the brick products are 24-bit multiplies (brick index < 2^7, strides
// < 2^24).  Full-rate instructions only.
__device__ inline uint32_t brick_offset_fast(uint32_t x, uint32_t y, uint32_t z, uint32_t row_stride,
                                             uint32_t slab_stride)
{
    const uint32_t mx = __builtin_amdgcn_perm(0x49484140u, 0x09080100u, (x & 7u) | 0x0c0c0c00u);
    const uint32_t my = __builtin_amdgcn_perm(0x92908280u, 0x12100200u, (y & 7u) | 0x0c0c0c00u);
    const uint32_t mz = __builtin_amdgcn_perm(0x49484140u, 0x09080100u, (z & 7u) | 0x0c0c0c00u);
    return ((x >> 3) << 9) + mx + __umul24(y >> 3, row_stride) + my + __umul24(z >> 3, slab_stride) + (mz << 2);
}

struct Vec3 {
    float x, y, z;
};

__device__ inline Vec3 mk3(float x, float y, float z)
{
    Vec3 r;
    r.x = x;
    r.y = y;
    r.z = z;
    return r;
}

__device__ inline float dot3(Vec3 a, Vec3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }

__device__ inline Vec3 normalize3(Vec3 v)
{
    float len = sqrtf((v.x * v.x + v.y * v.y) + v.z * v.z);
    return mk3(v.x / len, v.y / len, v.z / len);
}

__device__ inline Vec3 cross3(Vec3 a, Vec3 b)
{
    return mk3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}

// mat4 * vec4 rows 0..n-1, GLSL column-major, summed in column order.
__device__ inline float mat_row(const float* m, int r, float v0, float v1, float v2, float v3)
{
    float acc = m[r] * v0;
    acc = acc + m[4 + r] * v1;
    acc = acc + m[8 + r] * v2;
    acc = acc + m[12 + r] * v3;
    return acc;
}

// GetRayStuff / GetRayDirectionAt (InitialRayTraceFrag.glsl:398-414,
// ShadowRayTraceFrag.glsl:292-297), normalised as in main() (:427).
__device__ inline Vec3 ray_dir_at(const float* inv_view, const float* inv_proj, float sx, float sy)
{
    float cx = sx * 2.0f - 1.0f;
    float cy = sy * 2.0f - 1.0f;
    float ex = mat_row(inv_proj, 0, cx, cy, -1.0f, 1.0f);
    float ey = mat_row(inv_proj, 1, cx, cy, -1.0f, 1.0f);
    Vec3 rd;
    rd.x = mat_row(inv_view, 0, ex, ey, -1.0f, 0.0f);
    rd.y = mat_row(inv_view, 1, ex, ey, -1.0f, 0.0f);
    rd.z = mat_row(inv_view, 2, ex, ey, -1.0f, 0.0f);
    return rd;
}

__device__ inline float texcoord(int p, int n) { return ((float)p + 0.5f) / (float)n; }

// GetNormalFromID for a face byte; `miss` is the value of the reference's
// out-of-range branch (1.0 in most passes, 0.5 in the diffuse pass).
__device__ inline Vec3 face_normal(int face, float miss)
{
    switch (face) {
    case 0: return mk3(0.0f, 0.0f, 1.0f);
    case 1: return mk3(0.0f, 0.0f, -1.0f);
    case 2: return mk3(0.0f, 1.0f, 0.0f);
    case 3: return mk3(0.0f, -1.0f, 0.0f);
    case 4: return mk3(-1.0f, 0.0f, 0.0f);
    case 5: return mk3(1.0f, 0.0f, 0.0f);
    default: return mk3(miss, miss, miss);
    }
}

// GetNormalID of the normal with component -s on `axis`
// (InitialRayTraceFrag.glsl:143-185).
__device__ inline int face_from_axis(int axis, int s)
{
    if (s == 0) return 0;
    int n = -s;
    if (axis == 2) return n > 0 ? 0 : 1;
    if (axis == 1) return n > 0 ? 2 : 3;
    return n < 0 ? 4 : 5;
}

// The DF step E of the reference for a DF byte: floor(M * 0.57735026918f),
// except M == 1 -> 1.  The unorm8 round trip (M/255)*255 is exact for every
// byte, so M is the byte itself (the host tests check this against
// vpt_df_step for all 256 values).  E is 0 only for M = 0 and 1 only for
// M = 1, 2, 3 (3 * 0.57735026918f = 1.73, 4 * 0.57735026918f = 2.31), which
// is how trav_advance branches on the byte; the jump length E - 1 is then
// df_jump.
__device__ inline int df_step(int m)
{
    return m == 1 ? 1 : (int)((float)m * 0.57735026918f);
}

// (float)(E - 1) for a byte M >= 4: truncf is floor for the positive
// product, and E - 1 < 2^24 subtracts exactly.
__device__ inline float df_jump(int m)
{
    return truncf((float)m * 0.57735026918f) - 1.0f;
}

struct Hit {
    float t;
    int face;
    int block;
    uint32_t voxel;
    int samples;
};

__device__ inline uint32_t pack_cell(int x, int y, int z)
{
    return (uint32_t)x | ((uint32_t)y << 10) | ((uint32_t)z << 20);
}

// VoxelTraversalDF (InitialRayTraceFrag.glsl:307-374; identical copies in
// ShadowRayTraceFrag.glsl:222-289 and DiffuseRayTraceFrag.glsl:1043-1110),
// split into resumable pieces so that a persistent kernel can interleave
// several rays per lane:
//   trav_init     the loop's set-up,
//   trav_prepare  the head of one iteration (iteration cap, IsInVolume, the
//                 DF address); false = the loop has ended,
//   trav_advance  the rest of the iteration once the DF byte has arrived;
//                 false = the loop has ended,
//   trav_finish   the code after the loop.
// The per-ray state is the reference's: origin, direction, the reciprocal
// direction, initial origin, MinIdx, Intersection.  `samples` counts the
// loop's turns (one DF sample each); the loop ends after max_iter of them.
struct Trav {
    float ox, oy, oz;
    float dx, dy, dz;
    float ix, iy, iz;
    float o0x, o0y, o0z;
    int max_iter;
    int min_idx;
    int inter;
    int samples;
};

// The cell of the position, found by trav_prepare and used by trav_advance
// in the same loop turn.
struct Cell {
    uint32_t x, y, z;
};

__device__ inline int sgn(float v) { return v > 0.0f ? 1 : (v < 0.0f ? -1 : 0); }

// c ? x : y with both values computed in every lane first (the empty asm
// only pins them to registers and emits nothing), so that the choice stays
// one v_cndmask_b32 instead of two exec-masked arms.
__device__ __attribute__((always_inline)) inline float pick(bool c, float x, float y)
{
    asm volatile("" : "+v"(x), "+v"(y));
    return c ? x : y;
}

__device__ __attribute__((always_inline)) inline void trav_init(Trav& r, Vec3 o, Vec3 d, int max_iter)
{
    // o + 0 turns a -0 component into +0 and leaves every other value as it
    // is.  The two behave alike in all the loop's arithmetic (both floor to a
    // zero inside the volume), and no later step can produce -0 (a sum is -0
    // only when both terms are), so trav_prepare's bounds test never sees it.
    r.ox = o.x + 0.0f;
    r.oy = o.y + 0.0f;
    r.oz = o.z + 0.0f;
    r.dx = d.x;
    r.dy = d.y;
    r.dz = d.z;
    r.ix = 1.0f / d.x;
    r.iy = 1.0f / d.y;
    r.iz = 1.0f / d.z;
    r.o0x = o.x;
    r.o0y = o.y;
    r.o0z = o.z;
    r.max_iter = max_iter;
    r.min_idx = 0;
    r.inter = 0;
    r.samples = 0;
}

// Bit pattern of (float)dim, dim = max + 1 (0 for the empty view, max = -1).
__device__ inline uint32_t trav_bound(float max) { return __float_as_uint(max + 1.0f); }

// trav_prepare without the iteration cap, for loops that test the cap
// where trav_advance returns.
__device__ __attribute__((always_inline)) inline bool trav_locate(const WorldView& w, Trav& r, uint32_t& off,
                                                                  Cell& c)
{
    // IsInVolume(floor(o)), one unsigned compare per axis: floor(o) lies in
    // [0, dim - 1] exactly when o lies in [+0, dim); non-negative floats
    // order as their bit patterns, and negative values and NaN have patterns
    // above every bound.
    if (!(__float_as_uint(r.ox) < trav_bound(w.xmax) && __float_as_uint(r.oy) < trav_bound(w.ymax) &&
          __float_as_uint(r.oz) < trav_bound(w.zmax))) {
        r.inter = 0;
        return false;
    }
    // truncation is floor for o >= 0
    c.x = (uint32_t)r.ox;
    c.y = (uint32_t)r.oy;
    c.z = (uint32_t)r.oz;
    off = brick_offset_fast(c.x, c.y, c.z, w.row_stride, w.slab_stride);
    return true;
}

__device__ __attribute__((always_inline)) inline bool trav_prepare(const WorldView& w, Trav& r, uint32_t& off,
                                                                   Cell& c)
{
    return r.samples < r.max_iter && trav_locate(w, r, off, c);
}

__device__ __attribute__((always_inline)) inline bool trav_advance(Trav& r, int m, const Cell& c)
{
    r.samples++;
    if (m == 0) return false;   // E == 0
    if (m < 4) {
        // E == 1: the one-cell step, with selects on the step axis (no
        // divergent branches: fewer exec-mask moves in the loop)
        const int sx = sgn(r.dx), sy = sgn(r.dy), sz = sgn(r.dz);
        const int hx = (1 + sx) >> 1, hy = (1 + sy) >> 1, hz = (1 + sz) >> 1;
        // float(int(o)) of the reference is the cell's corner
        const float gx = (float)c.x, gy = (float)c.y, gz = (float)c.z;
        const float wx = r.ox - gx, wy = r.oy - gy, wz = r.oz - gz;
        const float fxx = ((float)hx - wx) * r.ix;
        const float fyy = ((float)hy - wy) * r.iy;
        const float fzz = ((float)hz - wz) * r.iz;
        const int mi = (fxx < fyy && sx != 0) ? ((fxx < fzz || sz == 0) ? 0 : 2) : ((fyy < fzz || sz == 0) ? 1 : 2);
        const bool m0 = mi == 0, m1 = mi == 1, m2 = mi == 2;
        const float fm = m0 ? fxx : (m1 ? fyy : fzz);
        // The reference forms float(int(o) + s) + w, then adds s * 0.0001 on
        // the step axis.  On that axis w is 1 - h, and the integers add
        // exactly: (g + s) + (1 - h) = g + (s + 1 - h), also for the s = 0
        // that the y axis can be picked with.  On the others s is 0 and w
        // moves by d * fm.
        const float ax = pick(m0, (gx + (float)(sx + 1 - hx)) + (float)sx * 0.0001f, gx + (wx + r.dx * fm));
        const float ay = pick(m1, (gy + (float)(sy + 1 - hy)) + (float)sy * 0.0001f, gy + (wy + r.dy * fm));
        const float az = pick(m2, (gz + (float)(sz + 1 - hz)) + (float)sz * 0.0001f, gz + (wz + r.dz * fm));
        r.ox = ax;
        r.oy = ay;
        r.oz = az;
        r.min_idx = mi;
        r.inter = 1;
    } else {
        const float k = df_jump(m);
        r.ox = r.ox + k * r.dx;
        r.oy = r.oy + k * r.dy;
        r.oz = r.oz + k * r.dz;
    }
    return r.samples < r.max_iter;
}

__device__ __attribute__((always_inline)) inline Hit trav_finish(const WorldView& w, const Trav& r)
{
    Hit h;
    h.t = -1.0f;
    h.face = 255;
    h.block = 0;
    h.voxel = 0xffffffffu;
    h.samples = r.samples;
    if (r.inter) {
        float fx = floorf(r.ox), fy = floorf(r.oy), fz = floorf(r.oz);
        if (fx >= 0.0f && fy >= 0.0f && fz >= 0.0f && fx <= w.xmax && fy <= w.ymax && fz <= w.zmax) {
            int cx = (int)fx, cy = (int)fy, cz = (int)fz;
            int b = plane_byte(w.grid, brick_offset_fast(cx, cy, cz, w.row_stride, w.slab_stride));
            if (b > 0) {
                float dx = r.ox - r.o0x, dy = r.oy - r.o0y, dz = r.oz - r.o0z;
                h.t = sqrtf((dx * dx + dy * dy) + dz * dz);
                // the three signs first, then a select: an indexed pick of the
                // direction components can end up as a scratch-memory load
                const int sx = sgn(r.dx), sy = sgn(r.dy), sz = sgn(r.dz);
                int s = sx;
                s = r.min_idx == 1 ? sy : s;
                s = r.min_idx == 2 ? sz : s;
                h.face = face_from_axis(r.min_idx, s);
                h.block = b;
                h.voxel = pack_cell(cx, cy, cz);
            }
        }
    }
    return h;
}

__device__ inline Hit traverse(const WorldView& w, Vec3 o, Vec3 d, int max_iter)
{
    Trav r;
    trav_init(r, o, d, max_iter);
    uint32_t off;
    Cell c;
    // the cap is tested once here and then where trav_advance returns
    if (max_iter > 0) {
        while (trav_locate(w, r, off, c)) {
            if (!trav_advance(r, plane_byte(w.df, off), c)) break;
        }
    }
    return trav_finish(w, r);
}

// GetVoxel(ivec3(floor(p))) (InitialRayTraceFrag.glsl:79-87).
__device__ inline int voxel_at(const WorldView& w, Vec3 p)
{
    float fx = floorf(p.x), fy = floorf(p.y), fz = floorf(p.z);
    if (!(fx >= 0.0f && fy >= 0.0f && fz >= 0.0f && fx <= w.xmax && fy <= w.ymax && fz <= w.zmax)) return 0;
    return plane_byte(w.grid, brick_offset_fast((uint32_t)(int)fx, (uint32_t)(int)fy, (uint32_t)(int)fz,
                                                w.row_stride, w.slab_stride));
}

// Hardware binary16 conversions (v_cvt_f32_f16 is exact; v_cvt_f16_f32
// rounds to nearest even with fp16 denormals preserved), i.e. the same
// results as vpt_f16_to_f32 / vpt_f32_to_f16 of the shared numerics.
__device__ inline float hw_h2f(uint32_t h) { return (float)__builtin_bit_cast(_Float16, (uint16_t)h); }
__device__ inline uint32_t hw_f2h(float x) { return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)x); }

// Blocks are dealt to the 8 XCDs round-robin by hardware; remap the block
// index so that XCD k receives a contiguous 1/8 of the tiles, keeping
// neighbouring tiles (which touch the same bricks) in one L2.
__device__ inline uint32_t xcd_remap(uint32_t bid, uint32_t nblocks)
{
    uint32_t per = nblocks / 8u;
    if (per == 0u || bid >= per * 8u) return bid;
    return (bid & 7u) * per + (bid >> 3);
}

}  // namespace vpt
