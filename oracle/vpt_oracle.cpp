/*
 * vpt_oracle.cpp — CPU oracle for the voxel path tracer hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a line-for-line C++17 restatement
 * of the reference's GLSL passes, used by tests/, by smoke() and by the
 * CPU-baseline leg of bench.py as the checker.  It is never linked into,
 * loaded by, or called from the product library (voxelpathtracer_amd/).
 *
 * Build: oracle/Makefile (g++ -O3 -fopenmp -ffp-contract=off, no fast-math).
 *
 * Reference files followed (paths relative to the reference repository):
 *   DF build        Core/Shaders/ManhattanDistance{X,Y,Z}.comp, Core/World.cpp:69-113
 *   traversal       Core/Shaders/InitialRayTraceFrag.glsl:68-102, :307-374
 *                   (identical copies: ShadowRayTraceFrag.glsl:222-289,
 *                    DiffuseRayTraceFrag.glsl:1043-1110)
 *   primary rays    InitialRayTraceFrag.glsl:143-185, :398-468; FBOVert.glsl:13-22
 *   shadow rays     ShadowRayTraceFrag.glsl:303-331, :414-513
 *   diffuse GI      DiffuseRayTraceFrag.glsl:535-664, :766-784, :822-935,
 *                   :945-967, :984-988, :1202-1222
 *   SVGF            Core/Shaders/Spatial3x3Initial.glsl, SVGF/TemporalFilter.glsl,
 *                   SVGF/VarianceEstimate.glsl, SVGF/SpatialFilter.glsl,
 *                   wiring Core/Pipeline.cpp:2284-2611
 *   SH decode       Core/Shaders/ColorPassFrag.glsl:356-368
 *   shadow denoise  Core/Shaders/ShadowTemporalFilter.glsl:98-267,
 *                   ShadowFilter.glsl:69-165, wiring Core/Pipeline.cpp:2854-2948
 *   reflections     Core/Shaders/ReflectionTraceFrag.glsl:764-1035 (traversal
 *                   core), wiring Core/Pipeline.cpp:3000-3060
 *   CPU picking     Core/World.cpp:496-546 (RaycastDetect)
 *
 * Parity status per function (what pins it besides this transcription):
 *   oracle_df_build        closed form min(254, L1) + brute force   -> pinned
 *   oracle_trace_*         textbook DDA agreement, RaycastDetect agreement,
 *                          axis-aligned known answers, invariants  -> pinned
 *                          (the reference's quirk cases are pinned by the
 *                          transcription only; see the traversal tests)
 *   oracle_gi              closed forms (SH round trip, luminance,
 *                          cosine-lobe statistics)                 -> pinned
 *   oracle_svgf_*          constant-image identity, weight
 *                          normalisation, running-mean closed form,
 *                          disocclusion passthrough; per-tap weights
 *                          and constants by closed forms on
 *                          structured inputs (impulses, edges, linear
 *                          and quadratic ramps, the distance ladder),
 *                          each of 43 planted mistakes failing one
 *                          (tests/test_svgf_pins_cpu.py)           -> pinned
 *   oracle_camera          closed forms of glm lookAt/perspective  -> pinned
 *   oracle_trace_shadow2   transversal: a known slab distance, the
 *                          miss / N.L / sky constants; mask equal to
 *                          oracle_trace_shadow                     -> pinned
 *   oracle_shadow_temporal static-camera frame-count ladder, closed-
 *                          form blend of uniform images, sky
 *                          passthrough                             -> pinned
 *   oracle_shadow_filter   uniform-image identity, sharp and sky
 *                          passthrough; per-tap weights, kernel-size,
 *                          spacing and luminance switches by edge
 *                          responses, 10 planted mistakes each
 *                          failing one (tests/test_svgf_pins_cpu.py)
 *                                                                  -> pinned
 *                          (the shadow temporal filter's per-tap
 *                          weights follow the transcription only)
 *   vpt_logf / vpt_powf    error bounds against double libm        -> pinned
 *   oracle_trace_reflect   corridor closed forms (first hit, face,
 *                          block, distance incl. the 0.001 offset,
 *                          shadow at the hit, the two-sample
 *                          average), an independent GL LINEAR /
 *                          NEAREST resampler, the 200 clamp, the
 *                          GGX half-vector polar law of the best of
 *                          three (KS) and its azimuth sectors, the
 *                          0.05 roughness floor                    -> pinned
 */
#include <stdint.h>
#include <string.h>
#include <math.h>
#include <algorithm>
#include <chrono>
#include <vector>
#include <omp.h>

#include "../include/vpt/vpt_numerics.h"

#define OAPI extern "C" __attribute__((visibility("default")))

namespace {

struct V3 { float x, y, z; };
static inline V3 v3(float x, float y, float z) { V3 r = {x, y, z}; return r; }
static inline float comp(const V3& v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
static inline void set_comp(V3& v, int i, float f) { if (i == 0) v.x = f; else if (i == 1) v.y = f; else v.z = f; }

struct Dims { int nx, ny, nz; };

/* IsInVolume(vec3(ivec3(floor(p)))) on the float value of floor(p)
 * (InitialRayTraceFrag.glsl:68-77).  A NaN coordinate counts as outside. */
static inline bool in_volume_f(float fx, float fy, float fz, const Dims& d)
{
    return fx >= 0.0f && fy >= 0.0f && fz >= 0.0f &&
           fx <= (float)(d.nx - 1) && fy <= (float)(d.ny - 1) && fz <= (float)(d.nz - 1);
}

static inline size_t lin(int x, int y, int z, const Dims& d)
{
    return (size_t)x + (size_t)y * d.nx + (size_t)z * (size_t)d.nx * d.ny;
}

/* GetVoxel(ivec3(floor(p))) (InitialRayTraceFrag.glsl:79-87) as a byte id. */
static inline int get_voxel_at(const uint8_t* grid, const Dims& d, V3 p)
{
    float fx = floorf(p.x), fy = floorf(p.y), fz = floorf(p.z);
    if (!in_volume_f(fx, fy, fz, d)) return 0;
    return grid[lin((int)fx, (int)fy, (int)fz, d)];
}

struct TraceResult {
    float t;          /* distance(origin, initial_origin) or -1 */
    int face;         /* 0..5, 255 on miss */
    int block;        /* hit block id, 0 on miss */
    int cx, cy, cz;   /* hit cell, -1 on miss */
    int samples;      /* loop iterations that read the DF */
};

/* Face id of the normal with component -s on axis `axis`
 * (GetNormalID, InitialRayTraceFrag.glsl:143-185: a zero normal matches no
 * entry and yields id 0). */
static inline int face_from_axis(int axis, int s)
{
    if (s == 0) return 0;
    int n = -s;
    if (axis == 2) return n > 0 ? 0 : 1;
    if (axis == 1) return n > 0 ? 2 : 3;
    return n < 0 ? 4 : 5;
}

/* VoxelTraversalDF (InitialRayTraceFrag.glsl:307-374). */
static TraceResult voxel_traversal_df(const uint8_t* grid, const uint8_t* df, const Dims& d,
                                      V3 origin, V3 direction, int max_iter)
{
    V3 initial_origin = origin;
    bool intersection = false;
    int min_idx = 0;
    int rs[3];
    for (int i = 0; i < 3; i++) {
        float c = comp(direction, i);
        rs[i] = c > 0.0f ? 1 : (c < 0.0f ? -1 : 0);      /* ivec3(sign(direction)) */
    }
    V3 inv = v3(1.0f / direction.x, 1.0f / direction.y, 1.0f / direction.z);
    int samples = 0;

    for (int itr = 0; itr < max_iter; itr++) {
        float fx = floorf(origin.x), fy = floorf(origin.y), fz = floorf(origin.z);
        if (!in_volume_f(fx, fy, fz, d)) {
            intersection = false;
            break;
        }
        int dist = df[lin((int)fx, (int)fy, (int)fz, d)];
        samples++;
        int euclidean = vpt_df_step(dist);
        if (euclidean == 0) break;
        if (euclidean == 1) {
            /* one DDA step (lines 339-356) */
            int g[3] = {(int)origin.x, (int)origin.y, (int)origin.z};
            float w[3] = {origin.x - (float)g[0], origin.y - (float)g[1], origin.z - (float)g[2]};
            int h[3] = {(1 + rs[0]) >> 1, (1 + rs[1]) >> 1, (1 + rs[2]) >> 1};
            float f[3] = {((float)h[0] - w[0]) * inv.x, ((float)h[1] - w[1]) * inv.y,
                          ((float)h[2] - w[2]) * inv.z};
            min_idx = (f[0] < f[1] && rs[0] != 0)
                          ? ((f[0] < f[2] || rs[2] == 0) ? 0 : 2)
                          : ((f[1] < f[2] || rs[2] == 0) ? 1 : 2);
            g[min_idx] += rs[min_idx];
            float fm = f[min_idx];
            w[0] = w[0] + direction.x * fm;
            w[1] = w[1] + direction.y * fm;
            w[2] = w[2] + direction.z * fm;
            w[min_idx] = (float)(1 - h[min_idx]);
            origin = v3((float)g[0] + w[0], (float)g[1] + w[1], (float)g[2] + w[2]);
            set_comp(origin, min_idx, comp(origin, min_idx) + (float)rs[min_idx] * 0.0001f);
            intersection = true;
        } else {
            float k = (float)(euclidean - 1);
            origin = v3(origin.x + k * direction.x, origin.y + k * direction.y, origin.z + k * direction.z);
        }
    }

    TraceResult r;
    r.t = -1.0f; r.face = 255; r.block = 0; r.cx = r.cy = r.cz = -1; r.samples = samples;
    if (intersection) {
        int b = get_voxel_at(grid, d, origin);
        if (b > 0) {
            float dx = origin.x - initial_origin.x;
            float dy = origin.y - initial_origin.y;
            float dz = origin.z - initial_origin.z;
            r.t = sqrtf((dx * dx + dy * dy) + dz * dz);
            r.face = face_from_axis(min_idx, rs[min_idx]);
            r.block = b;
            r.cx = (int)floorf(origin.x);
            r.cy = (int)floorf(origin.y);
            r.cz = (int)floorf(origin.z);
        }
    }
    return r;
}

/* mat4 * vec4 with GLSL column-major storage m[col*4 + row]; the sum is taken
 * in column order ((c0 + c1) + c2) + c3. */
static inline void mat_vec4(const float* m, const float v[4], float out[4])
{
    for (int r = 0; r < 4; r++) {
        float acc = m[0 * 4 + r] * v[0];
        acc = acc + m[1 * 4 + r] * v[1];
        acc = acc + m[2 * 4 + r] * v[2];
        acc = acc + m[3 * 4 + r] * v[3];
        out[r] = acc;
    }
}

static inline V3 normalize3(V3 v)
{
    float len = sqrtf((v.x * v.x + v.y * v.y) + v.z * v.z);
    return v3(v.x / len, v.y / len, v.z / len);
}

/* GetRayDirectionAt (ShadowRayTraceFrag.glsl:292-297; GetRayStuff
 * InitialRayTraceFrag.glsl:398-414 without TAA jitter), not normalised. */
static inline V3 ray_direction_at(const float* inv_view, const float* inv_proj, float sx, float sy)
{
    float clip[4] = {sx * 2.0f - 1.0f, sy * 2.0f - 1.0f, -1.0f, 1.0f};
    float e[4];
    mat_vec4(inv_proj, clip, e);
    float eye[4] = {e[0], e[1], -1.0f, 0.0f};
    float rd[4];
    mat_vec4(inv_view, eye, rd);
    return v3(rd[0], rd[1], rd[2]);
}

/* v_TexCoords at the centre of pixel (px, py) of a W x H target. */
static inline float texcoord(int p, int n) { return ((float)p + 0.5f) / (float)n; }

static inline V3 face_normal(int face, float miss_value)
{
    switch (face) {
    case 0: return v3(0.0f, 0.0f, 1.0f);
    case 1: return v3(0.0f, 0.0f, -1.0f);
    case 2: return v3(0.0f, 1.0f, 0.0f);
    case 3: return v3(0.0f, -1.0f, 0.0f);
    case 4: return v3(-1.0f, 0.0f, 0.0f);
    case 5: return v3(1.0f, 0.0f, 0.0f);
    default: return v3(miss_value, miss_value, miss_value);
    }
}

static inline float dot3(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline V3 cross3(V3 a, V3 b)
{
    return v3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
static inline float fclamp(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }

} /* namespace */

/* ======================================================================= */
/* Distance field                                                           */
/* ======================================================================= */

/* GenerateDistanceField (World.cpp:69-113): X, Y, Z passes in that order.
 * The DF image is R8 unorm: every store is unorm8(val / 255), every load is
 * floor(unorm * 255) (ManhattanDistanceX.comp:25-38). */
OAPI void oracle_df_build(const uint8_t* grid, int nx, int ny, int nz, uint8_t* df)
{
    Dims d = {nx, ny, nz};
    const float max_distance = (float)std::min(254, nx + ny + nz);
    auto store = [&](size_t i, float val) { df[i] = vpt_unorm8_encode(val / 255.0f); };
    auto load = [&](size_t i) { return floorf(vpt_unorm8_decode(df[i]) * 255.0f); };

    /* X pass (ManhattanDistanceX.comp:45-68) */
#pragma omp parallel for schedule(static)
    for (int zy = 0; zy < ny * nz; zy++) {
        int y = zy % ny, z = zy / ny;
        size_t i0 = lin(0, y, z, d);
        store(i0, grid[i0] > 0 ? 0.0f : max_distance);
        for (int x = 1; x < nx; x++) {
            size_t i = lin(x, y, z, d);
            store(i, grid[i] > 0 ? 0.0f : fminf(max_distance, 1.0f + load(i - 1)));
        }
        for (int x = nx - 2; x >= 0; x--) {
            size_t i = lin(x, y, z, d);
            if (load(i + 1) < load(i)) store(i, 1.0f + load(i + 1));
        }
    }
    /* Y pass (ManhattanDistanceY.comp:94-118) */
#pragma omp parallel for schedule(static)
    for (int zx = 0; zx < nx * nz; zx++) {
        int x = zx % nx, z = zx / nx;
        for (int y = 1; y < ny; y++) {
            size_t i = lin(x, y, z, d), p = lin(x, y - 1, z, d);
            if (load(p) < load(i)) store(i, 1.0f + load(p));
        }
        for (int y = ny - 2; y >= 0; y--) {
            size_t i = lin(x, y, z, d), p = lin(x, y + 1, z, d);
            if (load(p) < load(i)) store(i, 1.0f + load(p));
        }
    }
    /* Z pass (ManhattanDistanceZ.comp:142-165) */
#pragma omp parallel for schedule(static)
    for (int yx = 0; yx < nx * ny; yx++) {
        int x = yx % nx, y = yx / nx;
        for (int z = 1; z < nz; z++) {
            size_t i = lin(x, y, z, d), p = lin(x, y, z - 1, d);
            if (load(p) < load(i)) store(i, 1.0f + load(p));
        }
        for (int z = nz - 2; z >= 0; z--) {
            size_t i = lin(x, y, z, d), p = lin(x, y, z + 1, d);
            if (load(p) < load(i)) store(i, 1.0f + load(p));
        }
    }
}

/* ======================================================================= */
/* Rays                                                                     */
/* ======================================================================= */

/* Arbitrary ray list: rays[6*i] = origin xyz, direction xyz (already
 * normalised by the caller).  cell[3*i] = hit cell or -1. */
OAPI void oracle_trace_rays(const uint8_t* grid, const uint8_t* df, int nx, int ny, int nz,
                            const float* rays, int n, int max_iter,
                            float* t, int32_t* cell, uint8_t* face, uint8_t* block, int32_t* samples)
{
    Dims d = {nx, ny, nz};
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n; i++) {
        const float* r = rays + 6 * (size_t)i;
        TraceResult tr = voxel_traversal_df(grid, df, d, v3(r[0], r[1], r[2]), v3(r[3], r[4], r[5]), max_iter);
        t[i] = tr.t;
        if (cell) { cell[3 * i] = tr.cx; cell[3 * i + 1] = tr.cy; cell[3 * i + 2] = tr.cz; }
        if (face) face[i] = (uint8_t)tr.face;
        if (block) block[i] = (uint8_t)tr.block;
        if (samples) samples[i] = tr.samples;
    }
}

static inline uint32_t pack_cell(int x, int y, int z)
{
    if (x < 0) return 0xffffffffu;
    return (uint32_t)x | ((uint32_t)y << 10) | ((uint32_t)z << 20);
}

/* Primary trace (InitialRayTraceFrag.glsl:417-468).  Outputs per pixel
 * (row-major, row 0 = v_TexCoords.y smallest): t (o_HitDistance), face
 * (o_Normal id; 255 = miss, the unorm 1.0 of the reference), block
 * (o_BlockID), voxel (packed hit cell x | y<<10 | z<<20, or ~0). */
OAPI void oracle_trace_primary(const uint8_t* grid, const uint8_t* df, int nx, int ny, int nz,
                               const float* inv_view, const float* inv_proj, int W, int H, int max_iter,
                               float* t, uint8_t* face, uint8_t* block, uint32_t* voxel, int32_t* samples)
{
    Dims d = {nx, ny, nz};
    V3 org = v3(inv_view[12], inv_view[13], inv_view[14]);
#pragma omp parallel for schedule(dynamic, 16)
    for (int py = 0; py < H; py++) {
        for (int px = 0; px < W; px++) {
            size_t i = (size_t)py * W + px;
            V3 rd = ray_direction_at(inv_view, inv_proj, texcoord(px, W), texcoord(py, H));
            V3 dir = normalize3(rd);
            TraceResult tr = voxel_traversal_df(grid, df, d, org, dir, max_iter);
            bool intersect = tr.t > 0.0f && tr.block > 0;
            t[i] = tr.t;
            face[i] = (uint8_t)(intersect ? tr.face : 255);
            block[i] = (uint8_t)(intersect ? tr.block : 0);
            if (voxel) voxel[i] = intersect ? pack_cell(tr.cx, tr.cy, tr.cz) : 0xffffffffu;
            if (samples) samples[i] = tr.samples;
        }
    }
}

/* World-space position of a G-buffer pixel: GetPositionAt
 * (ShadowRayTraceFrag.glsl:299-303) at the pixel centre. */
static inline V3 gbuffer_position(const float* inv_view, const float* inv_proj, int px, int py, int W, int H, float dist)
{
    V3 org = v3(inv_view[12], inv_view[13], inv_view[14]);
    V3 dir = normalize3(ray_direction_at(inv_view, inv_proj, texcoord(px, W), texcoord(py, H)));
    return v3(org.x + dir.x * dist, org.y + dir.y * dist, org.z + dir.z * dist);
}

/* Sun shadow (ShadowRayTraceFrag.glsl:414-513, soft-shadow cone off,
 * alpha test off).  mask: 1 = in shadow, 0 = lit or sky. */
/* SHADOW_TRACE main (ShadowRayTraceFrag.glsl:414-513), both outputs.  The
 * second, o_IntersectionTransversal (R16F, :8), is clamp(T / 100, 1e-5,
 * 196), 4.25 / 100 when the ray misses, 1 / 100 when N.L <= 0.01 and 64 for
 * sky (:433-434, 486-487, 505-511); the ray is traced also when its origin
 * cell is solid. */
static void shadow_impl(const uint8_t* grid, const uint8_t* df, int nx, int ny, int nz,
                        const float* inv_view, const float* inv_proj, int W, int H,
                        const float* gb_t, const uint8_t* gb_face, const float* sun_dir, int max_iter,
                        uint8_t* mask, uint16_t* transversal, int32_t* samples, int32_t* rays)
{
    Dims d = {nx, ny, nz};
    V3 L = v3(sun_dir[0], sun_dir[1], sun_dir[2]);
#pragma omp parallel for schedule(dynamic, 16)
    for (int py = 0; py < H; py++) {
        for (int px = 0; px < W; px++) {
            size_t i = (size_t)py * W + px;
            int smp = 0, nr = 0;
            float tv;
            float dist = gb_t[i];
            if (dist < 0.0f) {
                mask[i] = 0;
                tv = 64.0f;
            } else {
                V3 P = gbuffer_position(inv_view, inv_proj, px, py, W, H, dist);
                V3 N = face_normal(gb_face[i], 1.0f);
                float ndotl = dot3(N, L);
                if (ndotl <= 0.01f) {
                    mask[i] = 1;
                    tv = 1.0f / 100.0f;
                } else {
                    V3 bias = v3(N.x * 0.06f, N.y * 0.06f, N.z * 0.06f);
                    V3 o = v3(P.x + bias.x, P.y + bias.y, P.z + bias.z);
                    float T = -1.0f;
                    int block_at = get_voxel_at(grid, d, o);
                    if (dist > 0.0f) {
                        TraceResult tr = voxel_traversal_df(grid, df, d, o, L, max_iter);
                        T = tr.t;
                        smp = tr.samples;
                        nr = 1;
                    }
                    mask[i] = (uint8_t)((T > 0.0f || block_at > 0) ? 1 : 0);
                    tv = fclamp(T / 100.0f, 0.00001f, 196.0f);
                    if (T < 0.0f) tv = 4.25f / 100.0f;
                }
            }
            if (transversal) transversal[i] = vpt_f32_to_f16(tv);
            if (samples) samples[i] = smp;
            if (rays) rays[i] = nr;
        }
    }
}

OAPI void oracle_trace_shadow(const uint8_t* grid, const uint8_t* df, int nx, int ny, int nz,
                              const float* inv_view, const float* inv_proj, int W, int H,
                              const float* gb_t, const uint8_t* gb_face, const float* sun_dir, int max_iter,
                              uint8_t* mask, int32_t* samples, int32_t* rays)
{
    shadow_impl(grid, df, nx, ny, nz, inv_view, inv_proj, W, H, gb_t, gb_face, sun_dir, max_iter, mask, nullptr,
                samples, rays);
}

OAPI void oracle_trace_shadow2(const uint8_t* grid, const uint8_t* df, int nx, int ny, int nz,
                               const float* inv_view, const float* inv_proj, int W, int H,
                               const float* gb_t, const uint8_t* gb_face, const float* sun_dir, int max_iter,
                               uint8_t* mask, uint16_t* transversal, int32_t* samples, int32_t* rays)
{
    shadow_impl(grid, df, nx, ny, nz, inv_view, inv_proj, W, H, gb_t, gb_face, sun_dir, max_iter, mask, transversal,
                samples, rays);
}

/* Parameters of the diffuse GI pass.  Host-side constants of the frame
 * (DiffuseRayTraceFrag.glsl:822-840) are precomputed by the caller. */
struct OGiParams {
    float sun_dir[3];          /* StrongerLightDirection */
    float light_color[3];      /* LIGHT_COLOR (already * 0.4 * GISunStrength) */
    float sky_radiance[3];     /* GetSkyColorAt(dir) * x for a constant sky */
    float sky_pixel_radiance[3]; /* texture(u_Skymap, v_RayDirection) * 2.66 */
    float albedo[128 * 3];     /* Lambert albedo per block id */
    uint32_t frame;
    int gi_iter;               /* u_DiffuseTraceLength (48) */
    int shadow_iter;           /* GetShadowAt cap (128) */
};

/* IrridianceToSH (DiffuseRayTraceFrag.glsl:766-784). */
static inline void irradiance_to_sh(V3 radiance, V3 dir, float sh[6])
{
    float co = radiance.x - radiance.z;
    float t = radiance.z + co * 0.5f;
    float cg = radiance.y - t;
    float y = fmaxf(t + cg * 0.5f, 0.0f);
    float l00 = 0.282095f;
    float l1_1 = 0.488603f * dir.y;
    float l10 = 0.488603f * dir.z;
    float l11 = 0.488603f * dir.x;
    sh[0] = fmaxf(l11 * y, -100.0f);
    sh[1] = fmaxf(l1_1 * y, -100.0f);
    sh[2] = fmaxf(l10 * y, -100.0f);
    sh[3] = fmaxf(l00 * y, -100.0f);
    sh[4] = co;
    sh[5] = cg;
}

static inline float luminance(V3 c) { return (c.x * 0.299f + c.y * 0.587f) + c.z * 0.114f; }

/* cosWeightedRandomHemisphereDirection (DiffuseRayTraceFrag.glsl:945-967)
 * with the two uniforms r drawn from PCG instead of hash2(). */
static inline V3 cos_weighted_dir(V3 n, float r0, float r1)
{
    V3 uu = normalize3(cross3(n, v3(0.0f, 1.0f, 1.0f)));
    V3 vv = cross3(uu, n);
    float ra = sqrtf(r1);
    float s, c;
    vpt_sincos_2pi(r0, &s, &c);
    float rx = ra * c;
    float ry = ra * s;
    float rz = sqrtf(1.0f - r1);
    V3 rr = v3((rx * uu.x + ry * vv.x) + rz * n.x, (rx * uu.y + ry * vv.y) + rz * n.y,
               (rx * uu.z + ry * vv.z) + rz * n.z);
    return normalize3(rr);
}

/* The GI pixel seed: pcg_hash(px + W*py + W*H*frame), chained for r.y. */
static inline void gi_random(int px, int py, int W, int H, uint32_t frame, float* r0, float* r1)
{
    uint32_t seed = (uint32_t)px + (uint32_t)W * (uint32_t)py + (uint32_t)W * (uint32_t)H * frame;
    uint32_t h0 = vpt_pcg_hash(seed);
    uint32_t h1 = vpt_pcg_hash(h0);
    *r0 = vpt_u32_to_unit(h0);
    *r1 = vpt_u32_to_unit(h1);
}

/* Diffuse GI, 1 spp, 1 bounce (DiffuseRayTraceFrag.glsl:822-935 with
 * CalculateDiffuse :535-664, MAX_BOUNCE_LIMIT = 1).  Substitutions fixed by
 * the workload: PCG randoms, Lambert sun BRDF (DiffuseHammon -> N.L / pi),
 * per-id albedo, constant sky.  Outputs:
 *   rgb[3*i]      clamp(radiance, 0, 8)   (sky pixels: sky_pixel_radiance)
 *   sh[6*i]       IrridianceToSH(radiance, first bounce dir), clamped +-100
 *   luma[i]       o_Utility, aosky[2*i] o_AOAndSkyLighting (clamped to [0,1])
 * The SVGF input planes are stored in binary16 / unorm8 by the caller. */
OAPI void oracle_trace_gi(const uint8_t* grid, const uint8_t* df, int nx, int ny, int nz,
                          const float* inv_view, const float* inv_proj, int W, int H,
                          const float* gb_t, const uint8_t* gb_face, const OGiParams* prm,
                          float* rgb, float* sh_out, float* luma, float* aosky, int32_t* samples,
                          int32_t* rays)
{
    Dims d = {nx, ny, nz};
    V3 L = v3(prm->sun_dir[0], prm->sun_dir[1], prm->sun_dir[2]);
    V3 light = v3(prm->light_color[0], prm->light_color[1], prm->light_color[2]);
    const float PI = 3.14159265359f;
#pragma omp parallel for schedule(dynamic, 16)
    for (int py = 0; py < H; py++) {
        for (int px = 0; px < W; px++) {
            size_t i = (size_t)py * W + px;
            float dist = gb_t[i];
            V3 N = face_normal(gb_face[i], 0.5f);
            int smp = 0;
            if (dist < 0.0f) {
                V3 sky = v3(prm->sky_pixel_radiance[0], prm->sky_pixel_radiance[1], prm->sky_pixel_radiance[2]);
                float sh[6];
                irradiance_to_sh(sky, N, sh);
                for (int k = 0; k < 6; k++) sh_out[6 * i + k] = sh[k];
                rgb[3 * i] = sky.x; rgb[3 * i + 1] = sky.y; rgb[3 * i + 2] = sky.z;
                luma[i] = 0.0f;                  /* GetLuminance(vec3(0)) */
                aosky[2 * i] = 1.0f; aosky[2 * i + 1] = 0.0f;
                if (samples) samples[i] = 0;
                if (rays) rays[i] = 0;
                continue;
            }
            V3 P = gbuffer_position(inv_view, inv_proj, px, py, W, H, dist);
            float r0, r1;
            gi_random(px, py, W, H, prm->frame, &r0, &r1);

            /* CalculateDiffuse, bounce 0 */
            const float bias = 0.06f;
            V3 o = v3(P.x + N.x * bias, P.y + N.y * bias, P.z + N.z * bias);
            V3 dir = cos_weighted_dir(N, r0, r1);
            float ao = 1.0f;
            bool skyhit = false;
            V3 contrib = v3(0.0f, 0.0f, 0.0f);
            TraceResult tr = voxel_traversal_df(grid, df, d, o, dir, prm->gi_iter);
            smp += tr.samples;
            int nr = 1;
            float T = tr.t;
            if (T > 0.0f && tr.block > 0) {
                V3 ip = v3(o.x + dir.x * T, o.y + dir.y * T, o.z + dir.z * T);
                V3 hn = face_normal(tr.face, 0.0f);
                int tex_ref = std::min(std::max(tr.block, 0), 127);
                V3 albedo = v3(prm->albedo[3 * tex_ref], prm->albedo[3 * tex_ref + 1], prm->albedo[3 * tex_ref + 2]);
                float ndotl = fmaxf(dot3(hn, L), 0.0f);
                float shadow_at;
                if (ndotl < 0.001f) {
                    shadow_at = 0.0f;
                } else {
                    V3 so = v3(ip.x + hn.x * 0.045f, ip.y + hn.y * 0.045f, ip.z + hn.z * 0.045f);
                    TraceResult st = voxel_traversal_df(grid, df, d, so, L, prm->shadow_iter);
                    smp += st.samples;
                    nr += 1;
                    shadow_at = st.t > 0.0f ? 1.0f : 0.0f;
                }
                /* SunBRDF (:499-508) with DiffuseHammon replaced by Lambert N.L/pi */
                float brdf = ndotl * (1.0f / PI);
                V3 rad35 = v3(light.x * 3.5f, light.y * 3.5f, light.z * 3.5f);
                float vis = 1.0f - shadow_at;
                V3 sun = v3(((albedo.x * brdf) * rad35.x) * vis * PI,
                            ((albedo.y * brdf) * rad35.y) * vis * PI,
                            ((albedo.z * brdf) * rad35.z) * vis * PI);
                contrib = sun;
            } else {
                V3 sky = v3(prm->sky_radiance[0], prm->sky_radiance[1], prm->sky_radiance[2]);
                contrib = sky;
                skyhit = true;
            }
            if (T < 2.0f && T > 0.0f) ao = fmaxf(T / 2.0f, 0.0f);

            V3 x = v3(fclamp(contrib.x, 0.0f, 8.0f), fclamp(contrib.y, 0.0f, 8.0f), fclamp(contrib.z, 0.0f, 8.0f));
            float sh[6];
            irradiance_to_sh(x, dir, sh);
            for (int k = 0; k < 4; k++) sh_out[6 * i + k] = fclamp(sh[k], -100.0f, 100.0f);
            sh_out[6 * i + 4] = fclamp(sh[4], -100.0f, 100.0f);
            sh_out[6 * i + 5] = fclamp(sh[5], -100.0f, 100.0f);
            rgb[3 * i] = x.x; rgb[3 * i + 1] = x.y; rgb[3 * i + 2] = x.z;
            float u = fmaxf(luminance(x), 0.01f);
            luma[i] = fclamp(u, 0.001f, 64.0f);
            aosky[2 * i] = fclamp(ao, 0.0f, 1.0f);
            aosky[2 * i + 1] = fclamp(skyhit ? 1.0f : 0.0f, 0.0f, 1.0f);
            if (samples) samples[i] = smp;
            if (rays) rays[i] = nr;
        }
    }
}

/* ======================================================================= */
/* SVGF                                                                     */
/* ======================================================================= */
/*
 * Textures are float arrays holding values already rounded to their GL
 * storage format; channels are interleaved (SH: 4, CoCg: 2, ...).  Sampling
 * follows GL: LINEAR = bilinear with REPEAT wrap at fractional coordinates,
 * NEAREST for the face/block planes (Framebuffer.cpp:55-73,
 * Pipeline.cpp:1094-1108).  A tap at an integer pixel offset reads the
 * texel directly (LINEAR at a texel centre returns the texel).
 */

namespace {

struct Tex {
    const float* p;
    int W, H, C;
    float at(int x, int y, int c) const { return p[((size_t)y * W + x) * C + c]; }
};

static inline int wrap(int i, int n) { int r = i % n; return r < 0 ? r + n : r; }

/* texture(sampler2D LINEAR REPEAT, (s, t)), channel c. */
static inline float bilinear(const Tex& T, float s, float t, int c)
{
    float u = s * (float)T.W - 0.5f;
    float v = t * (float)T.H - 0.5f;
    float fu = floorf(u), fv = floorf(v);
    float a = u - fu, b = v - fv;
    int i0 = wrap((int)fu, T.W), i1 = wrap((int)fu + 1, T.W);
    int j0 = wrap((int)fv, T.H), j1 = wrap((int)fv + 1, T.H);
    float top = T.at(i0, j0, c) * (1.0f - a) + T.at(i1, j0, c) * a;
    float bot = T.at(i0, j1, c) * (1.0f - a) + T.at(i1, j1, c) * a;
    return top * (1.0f - b) + bot * b;
}

/* texture(sampler2D NEAREST REPEAT) of a byte plane. */
static inline int nearest_u8(const uint8_t* p, int W, int H, float s, float t)
{
    int i = wrap((int)floorf(s * (float)W), W);
    int j = wrap((int)floorf(t * (float)H), H);
    return p[(size_t)j * W + i];
}

static inline bool in_screen(float x, float y) { return x > 0.0f && x < 1.0f && y > 0.0f && y < 1.0f; }

static inline float sh_to_y(float w) { return fmaxf(0.0f, 3.544905f * w); }

/* Planted mistakes, for the oracle's own pinning tests only: each id swaps
 * one constant or operation of the SVGF and shadow-filter passes for a
 * plausible wrong one, so that a test can show that its closed-form checks
 * fail on it.  0 (the default) is the transcription as written. */
enum Mutation {
    MUT_NONE = 0,
    MUT_PRE_K1, MUT_PRE_PHI, MUT_PRE_NW_SIGN, MUT_PRE_NW_DROP, MUT_PRE_FLOOR, MUT_PRE_ACCEPT,
    MUT_TMP_CENTRE, MUT_TMP_FLOOR, MUT_TMP_TOL_8_16, MUT_TMP_TOL_48_64, MUT_TMP_LAST_BAND, MUT_TMP_BLOCK_32,
    MUT_TMP_NORMAL_64, MUT_TMP_EDGE, MUT_TMP_MOMENT, MUT_TMP_CLAMP_COCG, MUT_TMP_SPP_CLAMP, MUT_TMP_SKY,
    MUT_VAR_THRESH, MUT_VAR_FACTOR3, MUT_VAR_PHI, MUT_VAR_DEPTH, MUT_VAR_WINDOW, MUT_VAR_SCALE, MUT_VAR_CLAMP,
    MUT_VAR_NW,
    MUT_AT_K1, MUT_AT_SCALE, MUT_AT_JIT_HALF, MUT_AT_JIT_AMP, MUT_AT_STRONG_FRAMES, MUT_AT_STRONG_STEP,
    MUT_AT_PHI_BIAS, MUT_AT_TWEAK, MUT_AT_GAUSS, MUT_AT_DEPTH, MUT_AT_DEPTH_CLAMP, MUT_AT_NW_CLAMP, MUT_AT_W_CLAMP,
    MUT_AT_VAR_PROP, MUT_AT_AO_STEP, MUT_AT_SKY, MUT_AT_LUMA_PHI,
    MUT_SF_SIZE, MUT_SF_SPACING, MUT_SF_SC24, MUT_SF_LUMA_SWITCH, MUT_SF_LM, MUT_SF_EXPO, MUT_SF_VE, MUT_SF_LE,
    MUT_SF_DEPTH, MUT_SF_PASS,
    MUT_RS_Y, MUT_RS_COCG, MUT_RS_AO, MUT_RS_DIST, MUT_RS_BASE, MUT_RS_ALBEDO, MUT_RS_RAD, MUT_RS_FRESNEL,
    MUT_RS_SHADOW_TRACE, MUT_RS_SKY, MUT_RS_MEAN,
    MUT_CMP_D_PI, MUT_CMP_ALPHA, MUT_CMP_SUNVIS, MUT_CMP_SAT, MUT_CMP_SH_L1, MUT_CMP_COCG, MUT_CMP_AO_EXP,
    MUT_CMP_SHADOW, MUT_CMP_SPEC_BORDER, MUT_CMP_KARIS, MUT_CMP_F0, MUT_CMP_SKY, MUT_CMP_KD,
    MUT_COUNT
};

static const char* const kMutationNames[MUT_COUNT] = {
    "none",
    "prespatial: K(1) = 1/2 instead of 2/3",
    "prespatial: colour phi 5 instead of 4",
    "prespatial: exp(-lw + nw) instead of exp(-lw - nw)",
    "prespatial: normal term dropped, exp(-lw)",
    "prespatial: weight floor 0.001 instead of 0.01",
    "prespatial: position test |dP|^2 < 2 instead of < 1",
    "temporal: centre tap weight 3/32 instead of 9/64",
    "temporal: blend floor 0.1 instead of 0.05",
    "temporal: tolerance 2.0 instead of 1.414 for 8 <= d < 16",
    "temporal: tolerance 3.5 instead of 4.2 for 48 <= d < 64",
    "temporal: last band (tolerance 14, no block / normal test) extends past 200",
    "temporal: block test kept up to d < 48 instead of 32",
    "temporal: normal test kept up to d < 96 instead of 64",
    "temporal: screen-edge threshold 0.001 instead of 0.0035",
    "temporal: moment blends luma instead of luma^2",
    "temporal: CoCg lower clamp -100 instead of -10",
    "temporal: history length clamp 255 instead of 150",
    "temporal: sky / non-sky mismatch not rejected",
    "variance: short-history threshold 8 instead of 12",
    "variance: short-history factor 1 instead of 3",
    "variance: colour phi 4 instead of 5",
    "variance: depth weight exp(-|dt|) not squared",
    "variance: 7x7 window instead of 9x9",
    "variance: no 12 / frames scaling",
    "variance: upper clamp 10 instead of 50",
    "variance: normal weight ignored",
    "a-trous: K(1) = 1/2 instead of 2/3",
    "a-trous: tap spacing without the resolution scale",
    "a-trous: jitter not halved",
    "a-trous: jitter amplitude step * 1.0 instead of step * 0.8",
    "a-trous: strong filtering while history <= 12 instead of <= 8",
    "a-trous: strong filtering at step 16 too",
    "a-trous: phi multiplied by the bias instead of divided",
    "a-trous: TweakVariance curve 96 instead of 128 below 0.01",
    "a-trous: variance pre-blur kernel 0.5 / 0.25",
    "a-trous: depth weight exp(-d) not squared",
    "a-trous: depth weight floor 1e-3 instead of 1e-4",
    "a-trous: normal weight floor 0.01 instead of 0.001",
    "a-trous: weight floor 0.01 instead of 0.001",
    "a-trous: variance propagated with w instead of w^2",
    "a-trous: AO filtered at step 8",
    "a-trous: sky pixels filtered",
    "a-trous: luminance weight without the sqrt of the variance",
    "shadow filter: 7x7 switch at sqrt(2) instead of 1.414 sqrt(2)",
    "shadow filter: tap spacing 1.0 instead of 1.2",
    "shadow filter: scale 2.0 instead of 2.4 above 16",
    "shadow filter: luminance switch at 4 instead of 7.5",
    "shadow filter: short-history exponent factor from 0.2 instead of 0.1",
    "shadow filter: exponent factor 1.0 instead of 0.9",
    "shadow filter: variance estimate +2 / +5 swapped",
    "shadow filter: luminance error / 2 instead of / 3",
    "shadow filter: depth weight squared instead of cubed",
    "shadow filter: pass-through below sqrt(2) dropped",
    "reflection shading: SHToIrradianceA Y = sqrt(pi) L0 instead of 2 sqrt(pi) L0",
    "reflection shading: Co and Cg swapped in SHToIrradianceA",
    "reflection shading: ambient times AO instead of AO^0.75",
    "reflection shading: AO applied beyond distance 40 too",
    "reflection shading: reprojected hits keep the ambient at the pixel",
    "reflection shading: ambient not multiplied by the albedo",
    "reflection shading: direct radiance without the 0.6 factor",
    "reflection shading: Fresnel dropped from the diffuse term (kd = 1)",
    "reflection shading: reprojected hits trace the shadow ray",
    "reflection shading: misses add nothing instead of the sky",
    "reflection shading: colour is the sum over samples, not the mean",
    "composite: GGX D without the pi",
    "composite: GGX alpha = r instead of r^2",
    "composite: SunVisibility clamped to [0, 1]",
    "composite: BasicSaturation weight 0.1 instead of 0.9",
    "composite: SH L1 constant 0.511664 instead of 1.023326",
    "composite: Co and Cg swapped in SHToIrridiance",
    "composite: AO^1 instead of AO^2.2",
    "composite: shadow ignored",
    "composite: reflection sampled up to the screen edge",
    "composite: Karis bias term 1.04 a004 sign flipped",
    "composite: F0 0.02 instead of 0.04",
    "composite: sky pixels 0 instead of the sky radiance",
    "composite: diffuse not weighted by 1 - F",
};

static int g_mutation = MUT_NONE;
static inline bool mut(int id) { return g_mutation == id; }

struct Frame {
    const float* inv_view;
    const float* inv_proj;
    int W, H;
    V3 origin() const { return v3(inv_view[12], inv_view[13], inv_view[14]); }
};

/* GetPositionAt(tex, txc) with Dist already fetched. */
static inline V3 position_from(const Frame& f, float sx, float sy, float dist)
{
    V3 o = f.origin();
    V3 dir = normalize3(ray_direction_at(f.inv_view, f.inv_proj, sx, sy));
    return v3(o.x + dir.x * dist, o.y + dir.y * dist, o.z + dir.z * dist);
}

static inline float r16(float x) { return vpt_round_f16(x); }
static inline float r8(float x) { return vpt_unorm8_decode(vpt_unorm8_encode(x)); }

} /* namespace */

OAPI int oracle_mutation_count(void) { return MUT_COUNT; }
OAPI const char* oracle_mutation_name(int id) { return id >= 0 && id < MUT_COUNT ? kMutationNames[id] : ""; }
OAPI void oracle_set_mutation(int id) { g_mutation = id >= 0 && id < MUT_COUNT ? id : MUT_NONE; }

/* Spatial3x3Initial.main (Spatial3x3Initial.glsl:96-176).
 * in:  sh[4], cocg[2], utility[1], ao[2]  (raw GI planes)
 * out: sh[4], cocg[2], utility[1], ao[2]  (DiffusePreTemporal_SpatialFBO) */
OAPI void oracle_svgf_prespatial(const float* inv_view, const float* inv_proj, int W, int H,
                                 const float* gb_t, const uint8_t* gb_face,
                                 const float* sh, const float* cocg, const float* util, const float* ao,
                                 float* o_sh, float* o_cocg, float* o_util, float* o_ao)
{
    Frame fr = {inv_view, inv_proj, W, H};
    const float atrous[3] = {1.0f, mut(MUT_PRE_K1) ? 0.5f : 2.0f / 3.0f, 1.0f / 6.0f};
    const float tex_w = 1.0f / (float)W, tex_h = 1.0f / (float)H;
#pragma omp parallel for schedule(static)
    for (int py = 0; py < H; py++) {
        for (int px = 0; px < W; px++) {
            size_t i = (size_t)py * W + px;
            float uvx = texcoord(px, W), uvy = texcoord(py, H);
            V3 base_pos = position_from(fr, uvx, uvy, gb_t[i]);
            V3 base_n = face_normal(gb_face[i], 1.0f);
            float base_sh[4] = {sh[4 * i], sh[4 * i + 1], sh[4 * i + 2], sh[4 * i + 3]};
            float base_luma = sh_to_y(base_sh[3]);
            float tsh[4] = {base_sh[0], base_sh[1], base_sh[2], base_sh[3]};
            float tcocg[2] = {cocg[2 * i], cocg[2 * i + 1]};
            float tw = 1.0f;
            float tao[2] = {ao[2 * i], ao[2 * i + 1]};
            float taow = 1.0f;
            const float phi_color = mut(MUT_PRE_PHI) ? 5.0f : 4.0f;
            for (int x = -1; x <= 1; x++) {
                for (int y = -1; y <= 1; y++) {
                    if (x == 0 && y == 0) continue;
                    float sx = uvx + ((float)x * 1.0f) * tex_w;
                    float sy = uvy + ((float)y * 1.0f) * tex_h;
                    if (!in_screen(sx, sy)) continue;
                    size_t j = (size_t)(py + y) * W + (px + x);
                    V3 sp = position_from(fr, sx, sy, gb_t[j]);
                    V3 pd = v3(fabsf(sp.x - base_pos.x), fabsf(sp.y - base_pos.y), fabsf(sp.z - base_pos.z));
                    float dist_sqr = dot3(pd, pd);
                    if (dist_sqr < (mut(MUT_PRE_ACCEPT) ? 2.0f : 1.0f)) {
                        float ssh[4] = {sh[4 * j], sh[4 * j + 1], sh[4 * j + 2], sh[4 * j + 3]};
                        V3 sn = face_normal(gb_face[j], 1.0f);
                        float sluma = sh_to_y(ssh[3]);
                        float nw = vpt_powi(fmaxf(dot3(base_n, sn), 0.0f), 16);
                        float lw = fabsf(sluma - base_luma) / phi_color;
                        float w = vpt_expf(mut(MUT_PRE_NW_SIGN) ? -lw + nw : (mut(MUT_PRE_NW_DROP) ? -lw : -lw - nw));
                        const float floor_w = mut(MUT_PRE_FLOOR) ? 0.001f : 0.01f;
                        w = fmaxf(w, floor_w);
                        w = (atrous[abs(x)] * atrous[abs(y)]) * w;
                        w = fmaxf(w, floor_w);
                        w = fclamp(w, 0.0f, 1.0f);
                        for (int k = 0; k < 4; k++) tsh[k] = tsh[k] + ssh[k] * w;
                        tcocg[0] = tcocg[0] + cocg[2 * j] * w;
                        tcocg[1] = tcocg[1] + cocg[2 * j + 1] * w;
                        tw = tw + w;
                        tao[0] = tao[0] + ao[2 * j] * w;
                        tao[1] = tao[1] + ao[2 * j + 1] * w;
                        taow = taow + w;
                    }
                }
            }
            tw = fmaxf(tw, 0.01f);
            for (int k = 0; k < 4; k++) o_sh[4 * i + k] = r16(tsh[k] / tw);
            o_cocg[2 * i] = r16(tcocg[0] / tw);
            o_cocg[2 * i + 1] = r16(tcocg[1] / tw);
            float aw = fmaxf(taow, 0.01f);
            o_ao[2 * i] = r8(tao[0] / aw);
            o_ao[2 * i + 1] = r8(tao[1] / aw);
            o_util[i] = r16(util[i]);
        }
    }
}

/* TemporalFilter.main (SVGF/TemporalFilter.glsl:132-360), u_BeUseful = true.
 * cur_*: pre-spatial planes of this frame; prev_*: temporal output of the
 * previous frame (sh 4, cocg 2, util 3 = spp/moment/luma, ao 2).
 * Previous positions are rebuilt with the *current* camera origin and ray
 * directions, as GetPositionAt does in the reference (v_RayOrigin and
 * u_InverseView are the current frame's). */
OAPI void oracle_svgf_temporal(const float* inv_view, const float* inv_proj, const float* prev_viewproj,
                               int W, int H,
                               const float* cur_t, const uint8_t* cur_face, const uint8_t* cur_block,
                               const float* prev_t, const uint8_t* prev_face, const uint8_t* prev_block,
                               const float* cur_sh, const float* cur_cocg, const float* cur_luma, const float* cur_ao,
                               const float* prev_sh, const float* prev_cocg, const float* prev_util, const float* prev_ao,
                               float* o_sh, float* o_cocg, float* o_util, float* o_ao)
{
    Frame fr = {inv_view, inv_proj, W, H};
    Tex Tpt = {prev_t, W, H, 1};
    Tex Tsh = {prev_sh, W, H, 4}, Tcc = {prev_cocg, W, H, 2}, Tut = {prev_util, W, H, 3}, Tao = {prev_ao, W, H, 2};
    const float weights[5] = {3.0f / 32.0f, 3.0f / 32.0f, mut(MUT_TMP_CENTRE) ? 3.0f / 32.0f : 9.0f / 64.0f,
                              3.0f / 32.0f, 3.0f / 32.0f};
    const float offs[5][2] = {{1, 0}, {0, 1}, {0, 0}, {-1, 0}, {0, -1}};
    const float tex_w = 1.0f / (float)W, tex_h = 1.0f / (float)H;
    const float b = mut(MUT_TMP_EDGE) ? 0.001f : 0.0035f;
#pragma omp parallel for schedule(static)
    for (int py = 0; py < H; py++) {
        for (int px = 0; px < W; px++) {
            size_t i = (size_t)py * W + px;
            float uvx = texcoord(px, W), uvy = texcoord(py, H);
            float base_w = cur_t[i];
            V3 base_pos = position_from(fr, uvx, uvy, base_w);
            V3 base_n = face_normal(cur_face[i], 1.0f);
            float base_sh[4] = {cur_sh[4 * i], cur_sh[4 * i + 1], cur_sh[4 * i + 2], cur_sh[4 * i + 3]};
            float base_cocg[2] = {cur_cocg[2 * i], cur_cocg[2 * i + 1]};
            float base_ao[2] = {cur_ao[2 * i], cur_ao[2 * i + 1]};

            /* Reprojection (:61-73) */
            float p4[4] = {base_pos.x, base_pos.y, base_pos.z, 1.0f};
            float q[4];
            mat_vec4(prev_viewproj, p4, q);
            float rx = (q[0] / q[3]) * 0.5f + 0.5f;
            float ry = (q[1] / q[3]) * 0.5f + 0.5f;
            float base_lum = cur_luma[i];

            int base_block = std::min(std::max((int)cur_block[i], 0), 127);
            /* Jitter = ivec2((GradientNoise() - 0.5) * 1.0) is always 0. */

            bool do_block = true, do_normal = true;
            float tol = 0.75f;
            V3 cam = fr.origin();
            V3 dv = v3(base_pos.x - cam.x, base_pos.y - cam.y, base_pos.z - cam.z);
            float dtp = sqrtf(dot3(dv, dv));
            if (dtp < 4.0f) { tol = 0.3f; do_block = true; do_normal = true; }
            else if (dtp < 6.0f) { tol = 0.65f; do_block = true; do_normal = true; }
            else if (dtp < 8.0f) { tol = 0.85f; do_block = true; do_normal = true; }
            else if (dtp < 16.0f) { tol = mut(MUT_TMP_TOL_8_16) ? 2.0f : 1.414f; do_block = true; do_normal = true; }
            else if (dtp < 32.0f) { tol = 2.4f; do_block = true; do_normal = true; }
            else if (dtp < 48.0f) { tol = 3.5f; do_block = mut(MUT_TMP_BLOCK_32); do_normal = true; }
            else if (dtp < 64.0f) { tol = mut(MUT_TMP_TOL_48_64) ? 3.5f : 4.2f; do_block = false; do_normal = true; }
            else if (dtp < 96.0f) { tol = 6.25f; do_block = false; do_normal = mut(MUT_TMP_NORMAL_64); }
            else if (dtp < 128.0f) { tol = 9.0f; do_block = false; do_normal = false; }
            else if (dtp < 200.0f || mut(MUT_TMP_LAST_BAND)) { tol = 14.0f; do_block = false; do_normal = false; }

            float tw = 0.0f;
            float ssh[4] = {0, 0, 0, 0}, scc[2] = {0, 0}, sao[2] = {0, 0};
            float slum = 0.0f, sspp = 0.0f, smom = 0.0f;
            int successful = 0;
            for (int k = 0; k < 5; k++) {
                float sx = rx + (offs[k][0] + 0.0f) * tex_w;
                float sy = ry + (offs[k][1] + 0.0f) * tex_h;
                if (!(sx < 1.0f - b && sx > b && sy < 1.0f - b && sy > b)) continue;
                float pw = bilinear(Tpt, sx, sy, 0);
                V3 pp = position_from(fr, sx, sy, pw);
                V3 pn = face_normal(nearest_u8(prev_face, W, H, sx, sy), 1.0f);
                V3 pd = v3(fabsf(base_pos.x - pp.x), fabsf(base_pos.y - pp.y), fabsf(base_pos.z - pp.z));
                float perr = dot3(pd, pd);
                int sblock = std::min(std::max(nearest_u8(prev_block, W, H, sx, sy), 0), 127);
                bool valid = false;
                if (perr < tol && (((pw < 0.0f) == (base_w < 0.0f)) || mut(MUT_TMP_SKY))) {
                    valid = true;
                    if (do_normal && !(pn.x == base_n.x && pn.y == base_n.y && pn.z == base_n.z)) valid = false;
                    if (do_block && base_block != sblock) valid = false;
                }
                if (valid) {
                    float cw = weights[k];
                    for (int c = 0; c < 4; c++) ssh[c] = ssh[c] + bilinear(Tsh, sx, sy, c) * cw;
                    scc[0] = scc[0] + bilinear(Tcc, sx, sy, 0) * cw;
                    scc[1] = scc[1] + bilinear(Tcc, sx, sy, 1) * cw;
                    sspp = sspp + bilinear(Tut, sx, sy, 0) * cw;
                    smom = smom + bilinear(Tut, sx, sy, 1) * cw;
                    slum = slum + bilinear(Tut, sx, sy, 2) * cw;
                    sao[0] = sao[0] + bilinear(Tao, sx, sy, 0) * cw;
                    sao[1] = sao[1] + bilinear(Tao, sx, sy, 1) * cw;
                    tw = tw + cw;
                    successful++;
                }
            }
            if (tw > 0.001f) {
                for (int c = 0; c < 4; c++) ssh[c] = ssh[c] / tw;
                scc[0] = scc[0] / tw; scc[1] = scc[1] / tw;
                smom = smom / tw;
                sspp = sspp / tw;
                slum = slum / tw;
                sao[0] = sao[0] / tw; sao[1] = sao[1] / tw;
            } else {
                successful = 0;
            }
            float increment = 1.0f;
            float spp_inc = sspp + increment;
            if (successful <= 0) spp_inc = 0.01f;
            const float blend_floor = mut(MUT_TMP_FLOOR) ? 0.1f : 0.05f;
            float blend = fmaxf(1.0f / spp_inc, blend_floor);
            float moment_factor = fmaxf(1.0f / spp_inc, blend_floor);
            float util_spp = spp_inc;
            if (successful <= 0) util_spp = 0.0f;
            float util_moment = (1.0f - moment_factor) * smom +
                                moment_factor * (mut(MUT_TMP_MOMENT) ? base_lum : base_lum * base_lum);
            float store_luma = slum * (1.0f - blend) + base_lum * blend;   /* mix(x, y, a) */
            float osh[4], occ[2], oao[2];
            for (int c = 0; c < 4; c++) osh[c] = ssh[c] * (1.0f - blend) + base_sh[c] * blend;
            for (int c = 0; c < 2; c++) occ[c] = scc[c] * (1.0f - blend) + base_cocg[c] * blend;
            for (int c = 0; c < 2; c++) oao[c] = sao[c] * (1.0f - blend) + base_ao[c] * blend;
            if (successful <= 0) {
                for (int c = 0; c < 4; c++) osh[c] = base_sh[c];
                for (int c = 0; c < 2; c++) { occ[c] = base_cocg[c]; oao[c] = base_ao[c]; }
            }
            for (int c = 0; c < 4; c++) o_sh[4 * i + c] = r16(fclamp(osh[c], -100.0f, 100.0f));
            for (int c = 0; c < 2; c++) o_cocg[2 * i + c] = r16(fclamp(occ[c], mut(MUT_TMP_CLAMP_COCG) ? -100.0f : -10.0f, 100.0f));
            for (int c = 0; c < 2; c++) o_ao[2 * i + c] = r8(fclamp(oao[c], 0.0f, 1.0f));
            o_util[3 * i] = r16(fclamp(util_spp, -150.0f, mut(MUT_TMP_SPP_CLAMP) ? 255.0f : 150.0f));
            o_util[3 * i + 1] = r16(fclamp(util_moment, -150.0f, 150.0f));
            o_util[3 * i + 2] = r16(fclamp(store_luma, -150.0f, 150.0f));
        }
    }
}

/* VarianceEstimate.main (SVGF/VarianceEstimate.glsl:73-192), DO_SPATIAL and
 * AGGRESSIVE_DISOCCLUSION_HANDLING true (Pipeline.cpp:83,87). */
OAPI void oracle_svgf_variance(const float* inv_view, const float* inv_proj, int W, int H,
                               const float* gb_t, const uint8_t* gb_face,
                               const float* sh, const float* cocg, const float* util,
                               float* o_sh, float* o_cocg, float* o_var)
{
    const float tex_w = 1.0f / (float)W, tex_h = 1.0f / (float)H;
    const float frame_bias = mut(MUT_VAR_THRESH) ? 0.0f : 4.0f;
    const float thresh = 4.0f + 4.0f + frame_bias;
    const int reach = mut(MUT_VAR_WINDOW) ? 3 : 4;
#pragma omp parallel for schedule(dynamic, 8)
    for (int py = 0; py < H; py++) {
        for (int px = 0; px < W; px++) {
            size_t i = (size_t)py * W + px;
            float uvx = texcoord(px, W), uvy = texcoord(py, H);
            float base_w = gb_t[i];
            V3 base_n = face_normal(gb_face[i], 1.0f);
            float frames = util[3 * i];
            float base_moment = util[3 * i + 1];
            float base_sh[4] = {sh[4 * i], sh[4 * i + 1], sh[4 * i + 2], sh[4 * i + 3]};
            float base_cocg[2] = {cocg[2 * i], cocg[2 * i + 1]};
            float base_lum = sh_to_y(base_sh[3]);
            float osh[4], occ[2];
            float variance;
            if (frames < thresh) {
                const float color_phi = mut(MUT_VAR_PHI) ? 4.0f : 5.0f;
                float tw = 0.0f, tmom = 0.0f, tlum = 0.0f, tw2 = 0.0f;
                float tsh[4] = {0, 0, 0, 0}, tcc[2] = {0, 0};
                for (int x = -reach; x <= reach; x++) {
                    for (int y = -reach; y <= reach; y++) {
                        float sx = uvx + (float)x * tex_w;
                        float sy = uvy + (float)y * tex_h;
                        if (!in_screen(sx, sy)) continue;
                        size_t j = (size_t)(py + y) * W + (px + x);
                        float sw = gb_t[j];
                        V3 sn = face_normal(gb_face[j], 1.0f);
                        float smom = util[3 * j + 1];
                        float s_sh[4] = {sh[4 * j], sh[4 * j + 1], sh[4 * j + 2], sh[4 * j + 3]};
                        float slum = sh_to_y(s_sh[3]);
                        float nw = mut(MUT_VAR_NW) ? 1.0f : vpt_powi(fmaxf(dot3(base_n, sn), 0.0f), 16);
                        float e = vpt_expf(-fabsf(sw - base_w));
                        float dw = mut(MUT_VAR_DEPTH) ? e : e * e;
                        float lw = fabsf(slum - base_lum) / color_phi;
                        float w = (vpt_expf(-lw) * nw) * dw;
                        float w2 = w;
                        w = fmaxf(w, 0.000000015f);
                        w2 = fmaxf(w2, 0.0000000015f);
                        tw = tw + w;
                        tmom = tmom + smom * w2;
                        for (int k = 0; k < 4; k++) tsh[k] = tsh[k] + s_sh[k] * w;
                        tcc[0] = tcc[0] + cocg[2 * j] * w;
                        tcc[1] = tcc[1] + cocg[2 * j + 1] * w;
                        tlum = tlum + slum * w2;
                        tw2 = tw2 + w2;
                    }
                }
                if (tw > 0.0f) {
                    tmom = tmom / tw2;
                    tlum = tlum / tw2;
                    tcc[0] = tcc[0] / tw; tcc[1] = tcc[1] / tw;
                    for (int k = 0; k < 4; k++) tsh[k] = tsh[k] / tw;
                }
                for (int k = 0; k < 4; k++) osh[k] = tsh[k];
                occ[0] = tcc[0]; occ[1] = tcc[1];
                float acc = tlum * tlum;
                variance = tmom - acc;
                variance = variance * (mut(MUT_VAR_FACTOR3) ? 1.0f : 3.0f);
            } else {
                for (int k = 0; k < 4; k++) osh[k] = base_sh[k];
                occ[0] = base_cocg[0]; occ[1] = base_cocg[1];
                variance = base_moment - base_lum * base_lum;
            }
            if (!mut(MUT_VAR_SCALE)) variance = variance * (thresh / frames);
            for (int k = 0; k < 4; k++) o_sh[4 * i + k] = r16(fclamp(osh[k], -100.0f, 100.0f));
            o_cocg[2 * i] = r16(fclamp(occ[0], -10.0f, 100.0f));
            o_cocg[2 * i + 1] = r16(fclamp(occ[1], -10.0f, 100.0f));
            o_var[i] = r16(fclamp(variance, -1.0f, mut(MUT_VAR_CLAMP) ? 10.0f : 50.0f));
        }
    }
}

/* GradientNoise (SpatialFilter.glsl:160-165) with u_Time = time. */
static inline float gradient_noise(int px, int py, float time)
{
    float t = time * 100.493850275f;
    float m = t - 500.0f * floorf(t / 500.0f);          /* mod(x, 500) */
    float cx = ((float)px + 0.5f) + m;
    float cy = ((float)py + 0.5f) + m;
    float inner = 0.06711056f * cx + 0.00583715f * cy;
    inner = inner - floorf(inner);
    float n = 52.9829189f * inner;
    return n - floorf(n);
}

/* SpatialFilter.main (SVGF/SpatialFilter.glsl:175-355) for one a-trous pass.
 * sh/cocg/var/ao: previous pass output; frames: temporal utility .x
 * (u_TemporalMoment).  color_phi_bias = 3.325 (Pipeline.cpp:85),
 * scale = mix(1, 2.4, u_ResolutionScale). */
OAPI void oracle_svgf_atrous(int W, int H, const float* gb_t, const uint8_t* gb_face,
                             const float* sh, const float* cocg, const float* var, const float* ao,
                             const float* temporal_util, int step, float scale, float color_phi_bias, float time,
                             float* o_sh, float* o_cocg, float* o_var, float* o_ao)
{
    Tex Tt = {gb_t, W, H, 1}, Tsh = {sh, W, H, 4}, Tcc = {cocg, W, H, 2}, Tv = {var, W, H, 1}, Tao = {ao, W, H, 2};
    const float atrous[3] = {1.0f, mut(MUT_AT_K1) ? 0.5f : 2.0f / 3.0f, 1.0f / 6.0f};
    const float gauss[2] = {mut(MUT_AT_GAUSS) ? 0.5f : 0.60283f, mut(MUT_AT_GAUSS) ? 0.25f : 0.198585f};
    const float tex_w = 1.0f / (float)W, tex_h = 1.0f / (float)H;
    if (mut(MUT_AT_SCALE)) scale = 1.0f;
    bool filter_ao = step <= (mut(MUT_AT_AO_STEP) ? 8 : 4);
    bool filter_sky = step <= 6;
    filter_sky = filter_sky || filter_ao;
#pragma omp parallel for schedule(static)
    for (int py = 0; py < H; py++) {
        for (int px = 0; px < W; px++) {
            size_t i = (size_t)py * W + px;
            float uvx = texcoord(px, W), uvy = texcoord(py, H);
            int jit = (int)((gradient_noise(px, py, time) - 0.5f) * ((float)step * (mut(MUT_AT_JIT_AMP) ? 1.0f : 0.8f)));
            float base_depth = gb_t[i];
            V3 base_n = face_normal(gb_face[i], 1.0f);
            float base_sh[4] = {sh[4 * i], sh[4 * i + 1], sh[4 * i + 2], sh[4 * i + 3]};
            float base_cocg[2] = {cocg[2 * i], cocg[2 * i + 1]};
            float base_lum = sh_to_y(base_sh[3]);
            /* GaussianVariance (:94-129) */
            float base_var = 0.0f, vsum = 0.0f, ksum = 0.0f;
            for (int x = -1; x <= 1; x++) {
                for (int y = -1; y <= 1; y++) {
                    float sx = uvx + (float)x * tex_w;
                    float sy = uvy + (float)y * tex_h;
                    if (!in_screen(sx, sy)) continue;
                    float kv = gauss[abs(x)] * gauss[abs(y)];
                    float v = var[(size_t)(py + y) * W + (px + x)];
                    if (x == 0 && y == 0) base_var = v;
                    vsum = vsum + v * kv;
                    ksum = ksum + kv;
                }
            }
            float var_est = vsum / fmaxf(ksum, 0.01f);
            float base_ao[2] = {ao[2 * i], ao[2 * i + 1]};

            float tsh[4] = {base_sh[0], base_sh[1], base_sh[2], base_sh[3]};
            float tcc[2] = {base_cocg[0], base_cocg[1]};
            float tw = 1.0f;
            float tvar = base_var;
            float tao[2] = {base_ao[0], base_ao[1]};
            float taow = 1.0f;

            float frames = temporal_util[3 * i];
            bool strong = frames <= (mut(MUT_AT_STRONG_FRAMES) ? 12.0f : 8.0f) && step <= (mut(MUT_AT_STRONG_STEP) ? 16 : 8);

            float curve = 0.0f;
            if (var_est < 0.01f) curve = mut(MUT_AT_TWEAK) ? 96.0f : 128.0f;
            else if (var_est < 0.025f) curve = 112.0f;
            else if (var_est < 0.05f) curve = 96.0f;
            else if (var_est < 0.075f) curve = 84.0f;
            else if (var_est < 0.1f) curve = 70.0f;
            float tweaked;
            if (var_est < 0.1f) {
                /* TweakVariance (:167-173): V * pow(1 - V, E + 6), E integral */
                float fv = fclamp(var_est, 0.0f, 1.0f);
                float tt = 1.0f - fv;
                tweaked = fv * vpt_powi(tt, (int)(curve + 6.0f));
            } else {
                tweaked = var_est;
            }
            float phi = mut(MUT_AT_LUMA_PHI) ? fmaxf(0.0f, 0.000001f + tweaked) : sqrtf(fmaxf(0.0f, 0.000001f + tweaked));
            phi = mut(MUT_AT_PHI_BIAS) ? phi * fmaxf(color_phi_bias, 0.1f) : phi / fmaxf(color_phi_bias, 0.1f);

            if (!(base_depth < 0.0f) || mut(MUT_AT_SKY)) {   /* (BaseDepth < 0) == (DepthDiff < 0), DepthDiff = abs(..) >= 0 */
                for (int x = -1; x <= 1; x++) {
                    for (int y = -1; y <= 1; y++) {
                        if (x == 0 && y == 0) continue;
                        const float jh = mut(MUT_AT_JIT_HALF) ? 1.0f : 0.5f;
                        float ox = ((float)x * (float)step) * scale + (float)jit * jh;
                        float oy = ((float)y * (float)step) * scale + (float)jit * jh;
                        float sx = uvx + ox * tex_w;
                        float sy = uvy + oy * tex_h;
                        if (!in_screen(sx, sy)) continue;
                        float sdepth = bilinear(Tt, sx, sy, 0);
                        float ddiff = fabsf(sdepth - base_depth);
                        V3 sn = face_normal(nearest_u8(gb_face, W, H, sx, sy), 1.0f);
                        float s_sh[4];
                        for (int c = 0; c < 4; c++) s_sh[c] = bilinear(Tsh, sx, sy, c);
                        float s_cc0 = bilinear(Tcc, sx, sy, 0), s_cc1 = bilinear(Tcc, sx, sy, 1);
                        float slum = sh_to_y(s_sh[3]);
                        float svar = bilinear(Tv, sx, sy, 0);
                        float nw = vpt_powi(fmaxf(dot3(base_n, sn), 0.0f), 32);
                        nw = fclamp(nw, mut(MUT_AT_NW_CLAMP) ? 0.01f : 0.001f, 1.0f);
                        float lw = fabsf(slum - base_lum) / phi;
                        float e = vpt_expf(-fmaxf(ddiff, 0.00001f));
                        float dw = fclamp(mut(MUT_AT_DEPTH) ? e : e * e, mut(MUT_AT_DEPTH_CLAMP) ? 0.001f : 0.0001f, 1.0f);
                        float w = strong ? (nw * dw) : ((vpt_expf(-lw) * nw) * dw);
                        w = fclamp(w, mut(MUT_AT_W_CLAMP) ? 0.01f : 0.001f, 1.0f);
                        float xw = atrous[abs(x)], yw = atrous[abs(y)];
                        w = (xw * yw) * w;
                        w = fmaxf(w, 0.00000001f);
                        for (int c = 0; c < 4; c++) tsh[c] = tsh[c] + s_sh[c] * w;
                        tcc[0] = tcc[0] + s_cc0 * w;
                        tcc[1] = tcc[1] + s_cc1 * w;
                        tvar = tvar + (mut(MUT_AT_VAR_PROP) ? w : w * w) * svar;
                        tw = tw + w;
                        if (filter_sky || filter_ao) {
                            float aow = fclamp(((xw * yw) * nw) * dw, 0.000001f, 1.0f);
                            tao[0] = tao[0] + bilinear(Tao, sx, sy, 0) * aow;
                            tao[1] = tao[1] + bilinear(Tao, sx, sy, 1) * aow;
                            taow = taow + aow;
                        }
                    }
                }
            }
            float osh[4], occ[2], ovar, oao[2];
            for (int c = 0; c < 4; c++) osh[c] = tsh[c] / tw;
            occ[0] = tcc[0] / tw; occ[1] = tcc[1] / tw;
            ovar = tvar / (tw * tw);
            oao[0] = tao[0] / taow; oao[1] = tao[1] / taow;
            if (!filter_ao) oao[0] = base_ao[0];
            for (int c = 0; c < 4; c++) o_sh[4 * i + c] = r16(fclamp(osh[c], -100.0f, 100.0f));
            o_cocg[2 * i] = r16(fclamp(occ[0], -10.0f, 100.0f));
            o_cocg[2 * i + 1] = r16(fclamp(occ[1], -10.0f, 100.0f));
            o_var[i] = r16(fclamp(ovar, -1.0f, 50.0f));
            o_ao[2 * i] = r8(fclamp(oao[0], 0.0f, 1.0f));
            o_ao[2 * i + 1] = r8(fclamp(oao[1], 0.0f, 1.0f));
        }
    }
}

/* SHToIrridiance(shY, CoCg, v) (ColorPassFrag.glsl:356-368) with the
 * G-buffer face normal as v. */
OAPI void oracle_sh_decode(int n, const float* sh, const float* cocg, const uint8_t* face, float* rgb)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        V3 v = face_normal(face[i], 1.0f);
        float x = (sh[4 * i] * v.x + sh[4 * i + 1] * v.y) + sh[4 * i + 2] * v.z;
        float Y = 2.0f * (1.023326f * x + 0.886226f * sh[4 * i + 3]);
        Y = fmaxf(Y, 0.0f);
        float k = (Y * 0.282095f) / (sh[4 * i + 3] + 1e-6f);
        float co = cocg[2 * i] * k, cg = cocg[2 * i + 1] * k;
        float T = Y - cg * 0.5f;
        float G = cg + T;
        float B = T - co * 0.5f;
        float R = B + co;
        rgb[3 * i] = fmaxf(R, 0.0f);
        rgb[3 * i + 1] = fmaxf(G, 0.0f);
        rgb[3 * i + 2] = fmaxf(B, 0.0f);
    }
}

/* ======================================================================= */
/* Cross-checks                                                             */
/* ======================================================================= */

/* World::RaycastDetect (World.cpp:496-546): epsilon stepping, 48 steps,
 * cells with any coordinate <= 0 treated as outside.  Falling off the end
 * (undefined in the reference) returns -1.  out = (x, y, z, id). */
OAPI void oracle_raycast_detect(const uint8_t* grid, int nx, int ny, int nz, const float* pos, const float* dir,
                                int32_t* out)
{
    Dims d = {nx, ny, nz};
    float p[3] = {pos[0], pos[1], pos[2]};
    float sgn[3];
    for (int i = 0; i < 3; i++) sgn[i] = dir[i] > 0 ? 1.0f : 0.0f;
    out[0] = out[1] = out[2] = out[3] = -1;
    for (int it = 0; it < 48; it++) {
        float tv[3];
        for (int i = 0; i < 3; i++) tv[i] = (floorf(p[i] + sgn[i]) - p[i]) / dir[i];
        float t = std::min(tv[0], std::min(tv[1], tv[2]));
        for (int i = 0; i < 3; i++) p[i] = p[i] + dir[i] * (t + 0.001f);
        int fx = (int)floorf(p[0]), fy = (int)floorf(p[1]), fz = (int)floorf(p[2]);
        if (!(fx >= nx || fy >= ny || fz >= nz || fx <= 0 || fy <= 0 || fz <= 0)) {
            int b = grid[lin((int)p[0], (int)p[1], (int)p[2], d)];
            if (b != 0) {
                out[0] = fx; out[1] = fy; out[2] = fz; out[3] = b;
                return;
            }
        }
    }
}

/* Textbook Amanatides & Woo (1987) traversal in double precision: first
 * solid cell along the ray within max_t, and the axis/sign of the face
 * crossed.  out = (x, y, z, face) or -1. */
OAPI void oracle_dda_textbook(const uint8_t* grid, int nx, int ny, int nz, const double* o, const double* dir,
                              double max_t, int32_t* out)
{
    Dims d = {nx, ny, nz};
    out[0] = out[1] = out[2] = out[3] = -1;
    int c[3], step[3];
    double tmax[3], tdelta[3];
    int n[3] = {nx, ny, nz};
    for (int i = 0; i < 3; i++) {
        c[i] = (int)floor(o[i]);
        if (dir[i] > 0) { step[i] = 1; tmax[i] = (floor(o[i]) + 1.0 - o[i]) / dir[i]; tdelta[i] = 1.0 / dir[i]; }
        else if (dir[i] < 0) { step[i] = -1; tmax[i] = (o[i] - floor(o[i])) / -dir[i]; tdelta[i] = -1.0 / dir[i]; }
        else { step[i] = 0; tmax[i] = INFINITY; tdelta[i] = INFINITY; }
    }
    int axis = -1;
    double t = 0.0;
    while (t <= max_t) {
        if (c[0] < 0 || c[1] < 0 || c[2] < 0 || c[0] >= n[0] || c[1] >= n[1] || c[2] >= n[2]) return;
        if (grid[lin(c[0], c[1], c[2], d)] > 0) {
            out[0] = c[0]; out[1] = c[1]; out[2] = c[2];
            out[3] = axis < 0 ? -1 : face_from_axis(axis, step[axis]);
            return;
        }
        axis = (tmax[0] < tmax[1]) ? (tmax[0] < tmax[2] ? 0 : 2) : (tmax[1] < tmax[2] ? 1 : 2);
        t = tmax[axis];
        c[axis] += step[axis];
        tmax[axis] += tdelta[axis];
    }
}

/* ======================================================================= */
/* Camera (FpsCamera.cpp:23-24,150-163; Player.cpp:10)                      */
/* ======================================================================= */

/* glm::perspective / glm::lookAt (right-handed, clip z in [-1, 1]) and
 * their inverses, evaluated in double and rounded to float. */
OAPI void oracle_camera(const float* pos, float yaw_deg, float pitch_deg, float fov_deg, float aspect, float znear,
                        float zfar, float* view, float* proj, float* inv_view, float* inv_proj)
{
    const double PI = 3.14159265358979323846;
    double yaw = yaw_deg * PI / 180.0, pitch = pitch_deg * PI / 180.0;
    double f[3] = {cos(pitch) * cos(yaw), sin(pitch), cos(pitch) * sin(yaw)};
    double fl = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (int i = 0; i < 3; i++) f[i] /= fl;
    double up[3] = {0, 1, 0};
    double s[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
    double sl = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (int i = 0; i < 3; i++) s[i] /= sl;
    double u[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2], s[0] * f[1] - s[1] * f[0]};
    double e[3] = {pos[0], pos[1], pos[2]};
    double V[16] = {0};
    V[0] = s[0]; V[4] = s[1]; V[8] = s[2];
    V[1] = u[0]; V[5] = u[1]; V[9] = u[2];
    V[2] = -f[0]; V[6] = -f[1]; V[10] = -f[2];
    V[12] = -(s[0] * e[0] + s[1] * e[1] + s[2] * e[2]);
    V[13] = -(u[0] * e[0] + u[1] * e[1] + u[2] * e[2]);
    V[14] = (f[0] * e[0] + f[1] * e[1] + f[2] * e[2]);
    V[15] = 1.0;
    double th = tan(fov_deg * PI / 180.0 / 2.0);
    double P[16] = {0};
    P[0] = 1.0 / (aspect * th);
    P[5] = 1.0 / th;
    P[10] = -((double)zfar + znear) / ((double)zfar - znear);
    P[11] = -1.0;
    P[14] = -(2.0 * zfar * znear) / ((double)zfar - znear);
    /* inverses: view is rigid, projection has a closed form */
    double IV[16] = {0};
    IV[0] = s[0]; IV[1] = s[1]; IV[2] = s[2];
    IV[4] = u[0]; IV[5] = u[1]; IV[6] = u[2];
    IV[8] = -f[0]; IV[9] = -f[1]; IV[10] = -f[2];
    IV[12] = e[0]; IV[13] = e[1]; IV[14] = e[2]; IV[15] = 1.0;
    double IP[16] = {0};
    IP[0] = 1.0 / P[0];
    IP[5] = 1.0 / P[5];
    IP[11] = 1.0 / P[14];
    IP[14] = -1.0;
    IP[15] = P[10] / P[14];
    for (int i = 0; i < 16; i++) {
        view[i] = (float)V[i];
        proj[i] = (float)P[i];
        inv_view[i] = (float)IV[i];
        inv_proj[i] = (float)IP[i];
    }
}

/* ======================================================================= */
/* CPU baseline timing                                                      */
/* ======================================================================= */

/* Times DF build and primary tracing on the host cores: one warm-up, then
 * the median of `reps` runs.  out = {df_ms, primary_ms, total_samples,
 * threads}. */
OAPI void oracle_time_plain(const uint8_t* grid, int nx, int ny, int nz, const float* inv_view, const float* inv_proj,
                            int W, int H, int max_iter, int reps, double* out)
{
    std::vector<uint8_t> df((size_t)nx * ny * nz);
    std::vector<float> t((size_t)W * H);
    std::vector<uint8_t> face((size_t)W * H), block((size_t)W * H);
    std::vector<uint32_t> vox((size_t)W * H);
    std::vector<int32_t> smp((size_t)W * H);
    std::vector<double> dms, pms;
    for (int r = 0; r <= reps; r++) {
        auto t0 = std::chrono::steady_clock::now();
        oracle_df_build(grid, nx, ny, nz, df.data());
        auto t1 = std::chrono::steady_clock::now();
        oracle_trace_primary(grid, df.data(), nx, ny, nz, inv_view, inv_proj, W, H, max_iter, t.data(), face.data(),
                             block.data(), vox.data(), smp.data());
        auto t2 = std::chrono::steady_clock::now();
        if (r > 0) {
            dms.push_back(std::chrono::duration<double, std::milli>(t1 - t0).count());
            pms.push_back(std::chrono::duration<double, std::milli>(t2 - t1).count());
        }
    }
    std::sort(dms.begin(), dms.end());
    std::sort(pms.begin(), pms.end());
    double total = 0;
    for (auto v : smp) total += v;
    out[0] = dms.empty() ? 0 : dms[dms.size() / 2];
    out[1] = pms.empty() ? 0 : pms[pms.size() / 2];
    out[2] = total;
    out[3] = (double)omp_get_max_threads();
}

/* OpenMP threads the oracle runs with (cores used by the CPU baseline). */
/* ======================================================================= */
/* Sun-shadow denoise: ShadowTemporalFilter + ShadowFilter                  */
/* ======================================================================= */
/* Planes: mask u8 0/1 (the shadow pass's o_Shadow), transversal binary16
 * (o_IntersectionTransversal), temporal colour u8 unorm (GL_RED
 * UNSIGNED_BYTE) and frame count binary16 (R16F) per pixel
 * (Pipeline.cpp:1152-1154).  texture() at the pixel's own centre or at
 * whole-texel offsets from it reads the texel; other coordinates are GL
 * LINEAR + REPEAT (bilinear above); normal ids are NEAREST.  GLSL pow with
 * an integer-valued literal exponent is vpt_powi, with a run-time exponent
 * vpt_powf. */
namespace {

/* GetShadowSpatial (ShadowTemporalFilter.glsl:98-150) over the raw mask. */
static float shadow_spatial_pre(int px, int py, int W, int H, float tva, float base_depth, const float* cur_t,
                                const uint8_t* cur_face, const uint8_t* mask)
{
    const float unit_diagonal = sqrtf(2.0f);
    if (tva <= unit_diagonal * 2.0f) return 1.0f;
    const size_t i = (size_t)py * W + px;
    const float tex_w = 1.0f / (float)W, tex_h = 1.0f / (float)H;
    const float uvx = texcoord(px, W), uvy = texcoord(py, H);
    float total = (float)mask[i];
    const float base = total;
    float weight = 1.0f;
    const V3 bn = face_normal(cur_face[i], 1.0f);
    const float b = 0.03f;
    for (int x = -1; x <= 1; x++) {
        for (int y = -1; y <= 1; y++) {
            if (x == 0 && y == 0) continue;
            float sx = uvx + (float)x * tex_w, sy = uvy + (float)y * tex_h;
            if (!(sx > b && sx < 1.0f - b && sy > b && sy < 1.0f - b)) continue;
            const size_t j = (size_t)(py + y) * W + (px + x);
            const float sd = cur_t[j];
            const V3 sn = face_normal(cur_face[j], 1.0f);
            if (sn.x == bn.x && sn.y == bn.y && sn.z == bn.z && fabsf(sd - base_depth) < 1.0f) {
                const float smp = (float)mask[j];
                float wa = fclamp(1.0f - fclamp(fabsf(smp - base) / 3.0f, 0.0f, 1.0f), 0.0f, 1.0f);
                wa = fclamp(vpt_powi(wa, 7), 0.000001f, 1.0f);
                total = total + smp * wa;
                weight = weight + wa;
            }
        }
    }
    return total / weight;
}

} /* namespace */

/* ShadowTemporalFilter.main (ShadowTemporalFilter.glsl:187-267) with
 * u_ShadowTemporal true (Pipeline.cpp:2856-2908). */
OAPI void oracle_shadow_temporal(const float* inv_view, const float* inv_proj, const float* prev_viewproj, int W, int H,
                                 const float* cur_t, const uint8_t* cur_face, const float* prev_t, const uint8_t* mask,
                                 const uint16_t* transversal, const uint8_t* prev_color, const uint16_t* prev_frames,
                                 uint8_t* o_color, uint16_t* o_frames)
{
    Frame fr = {inv_view, inv_proj, W, H};
    std::vector<float> pc((size_t)W * H), pf((size_t)W * H);
    for (size_t i = 0; i < pc.size(); i++) {
        pc[i] = vpt_unorm8_decode(prev_color[i]);
        pf[i] = vpt_f16_to_f32(prev_frames[i]);
    }
    Tex Tpc = {pc.data(), W, H, 1}, Tpf = {pf.data(), W, H, 1}, Tpt = {prev_t, W, H, 1};
#pragma omp parallel for schedule(static)
    for (int py = 0; py < H; py++) {
        for (int px = 0; px < W; px++) {
            const size_t i = (size_t)py * W + px;
            const float uvx = texcoord(px, W), uvy = texcoord(py, H);
            const float dist = cur_t[i];
            float oc, of;
            if (dist > 0.0f) {
                const V3 pos = position_from(fr, uvx, uvy, dist);
                float p4[4] = {pos.x, pos.y, pos.z, 1.0f}, q[4];
                mat_vec4(prev_viewproj, p4, q);
                const float rx = (q[0] / q[3]) * 0.5f + 0.5f, ry = (q[1] / q[3]) * 0.5f + 0.5f;
                const float tva = vpt_f16_to_f32(transversal[i]) * 100.0f;
                float cur = shadow_spatial_pre(px, py, W, H, tva, dist, cur_t, cur_face, mask);
                const float bias = 0.005f;
                const bool rejected = !(rx > 0.0f + bias && rx < 1.0f - bias && ry > 0.0f + bias && ry < 1.0f - bias);
                if (!rejected) {
                    const float prev_orig = bilinear(Tpc, rx, ry, 0);
                    float prev = prev_orig;
                    if (tva < 1.414f * 3.0f) {   /* ClipShadow (:166-185) */
                        const int offs[5][2] = {{-1, 0}, {1, 0}, {0, 0}, {0, -1}, {0, 1}};
                        float mn = 100.0f, mx = -100.0f;
                        for (int k = 0; k < 5; k++) {
                            const float smp = (float)mask[(size_t)wrap(py + offs[k][1], H) * W + wrap(px + offs[k][0], W)];
                            mn = fminf(smp, mn);
                            mx = fmaxf(smp, mx);
                        }
                        const float lo = mn - 0.025f, hi = mx + 0.025f;   /* clipAABB (:156-164) */
                        const float pcl = 0.5f * (hi + lo), ecl = 0.5f * (hi - lo);
                        const float vcl = prev_orig - pcl;
                        const float denom = fabsf(vcl / ecl);
                        prev = denom > 1.0f ? pcl + vcl / denom : prev_orig;
                    }
                    const float pw = bilinear(Tpt, rx, ry, 0);
                    const V3 pp = position_from(fr, rx, ry, pw);
                    const V3 dd = v3(pp.x - pos.x, pp.y - pos.y, pp.z - pos.z);
                    const float d = sqrtf(dot3(dd, dd));
                    cur = fclamp(cur, 0.0f, 1.0f);
                    prev = fclamp(prev, 0.0f, 1.0f);
                    const float vx = (uvx - rx) * (float)W, vy = (uvy - ry) * (float)H;
                    const float clip_err = fabsf(prev_orig - prev);
                    const float inc = clip_err < 0.2f ? 1.0f : 0.6f;
                    const float fc = bilinear(Tpf, rx, ry, 0);
                    const float fi = fc + inc;
                    float bf = fclamp((1.0f - (1.0f / fi)) * 1.2f, 0.01f, 0.97f);
                    const float vr = fclamp(vpt_expf(-sqrtf(vx * vx + vy * vy)) * 0.8f + 0.6f, 0.00000001f, 1.0f);
                    bf = bf * vr;
                    float dr = 1.0f;
                    if (d > 0.4f) {
                        dr = vpt_powi(vpt_expf(-d), 48);
                        bf = bf * fclamp(dr, 0.0f, 1.0f);
                    }
                    const float a = fclamp(bf, 0.0f, 0.97f);
                    oc = cur * (1.0f - a) + prev * a;
                    const float mults = dr * vr;
                    of = fc + fclamp(mults * 1.1f, 0.0f, 1.0f);
                    if (mults < 0.1f) of = 0.0f;
                    else if (mults <= 0.2f + 0.001f) of = 2.0f;
                    else if (mults <= 0.3f + 0.001f) of = 3.25f;
                } else {
                    oc = cur;
                    of = 0.0f;
                }
            } else {
                oc = (float)mask[i];
                of = 0.0f;
            }
            of = fclamp(of, 0.0f, 256.0f);
            o_color[i] = vpt_unorm8_encode(oc);
            o_frames[i] = vpt_f32_to_f16(of);
        }
    }
}

/* ShadowFilter.main / ShadowSpatial (ShadowFilter.glsl:69-165) over the
 * temporal output, u_ShadowFilterScale = scale (Pipeline.cpp:2912-2948). */
OAPI void oracle_shadow_filter(const float* inv_view, const float* inv_proj, int W, int H, const float* gb_t,
                               const uint8_t* gb_face, const uint8_t* color, const uint16_t* transversal,
                               const uint16_t* frames, float scale, uint8_t* out)
{
    (void)inv_view;
    (void)inv_proj;
    std::vector<float> cf((size_t)W * H);
    for (size_t i = 0; i < cf.size(); i++) cf[i] = vpt_unorm8_decode(color[i]);
    Tex Tt = {gb_t, W, H, 1}, Tc = {cf.data(), W, H, 1};
    const float tex_w = 1.0f / (float)W, tex_h = 1.0f / (float)H;
#pragma omp parallel for schedule(static)
    for (int py = 0; py < H; py++) {
        for (int px = 0; px < W; px++) {
            const size_t i = (size_t)py * W + px;
            const float uvx = texcoord(px, W), uvy = texcoord(py, H);
            const float fr = vpt_f16_to_f32(frames[i]);
            const bool apply_luma = fr > (mut(MUT_SF_LUMA_SWITCH) ? 4.0f : 7.5f);
            const float dist = gb_t[i];
            const bool sky = dist < 0.0f;
            const V3 cn = face_normal(gb_face[i], 1.0f);
            const float cs = cf[i];
            const float T = vpt_f16_to_f32(transversal[i]) * 100.0f;
            const float cutoff = sqrtf(2.0f);
            if (((T > 0.0f && T < cutoff) && !mut(MUT_SF_PASS)) || sky) {
                out[i] = vpt_unorm8_encode(cs);
                continue;
            }
            const int size = T < cutoff * (mut(MUT_SF_SIZE) ? 1.0f : 1.414f) ? 1 : 3;
            float sc = 1.0f;
            if (T > 6.0f) sc = 2.0f;
            if (T > 16.0f) sc = mut(MUT_SF_SC24) ? 2.0f : 2.4f;
            if (T > 32.0f) sc = 2.6f;
            const float ct = fclamp(T, 0.0f, 10.0f);
            const float a = ct / 10.0f;
            float ve = (20.0f * (1.0f - a) + 6.0f * a) + (T < 6.0f ? (mut(MUT_SF_VE) ? 2.0f : 5.0f) : (mut(MUT_SF_VE) ? 5.0f : 2.0f));
            ve = fclamp(ve - 1.75f, 0.0000001f, 64.0f);
            float lm = 1.0f;
            if (!apply_luma) {
                const float fn = fr / 7.5f;
                lm = (mut(MUT_SF_LM) ? 0.2f : 0.1f) * (1.0f - fn) + 0.5f * fn;
            }
            const float expo = (ve * lm) * (mut(MUT_SF_EXPO) ? 1.0f : 0.9f);
            float tw = 0.0f, ts = 0.0f;
            for (int x = -size; x <= size; x++) {
                for (int y = -size; y <= size; y++) {
                    const float sp = mut(MUT_SF_SPACING) ? 1.0f : 1.2f;
                    const float sx = uvx + (((float)x * tex_w) * sp * sc) * scale;
                    const float sy = uvy + (((float)y * tex_h) * sp * sc) * scale;
                    const float sd = bilinear(Tt, sx, sy, 0);
                    const V3 sn = face_normal(nearest_u8(gb_face, W, H, sx, sy), 1.0f);
                    const float dw = vpt_powi(vpt_expf(-fabsf(dist - sd)), mut(MUT_SF_DEPTH) ? 2 : 3);
                    const float nw = vpt_powi(fmaxf(dot3(cn, sn), 0.000000001f), 32);
                    const float sa = bilinear(Tc, sx, sy, 0);
                    const float le = fclamp(1.0f - fclamp(fabsf(sa - cs) / (mut(MUT_SF_LE) ? 2.0f : 3.0f), 0.0f, 1.0f), 0.0f, 1.0f);
                    float w = 1.0f;
                    w = w * fclamp(vpt_powf(le, expo), 0.0f, 1.0f);
                    w = w * dw;
                    w = w * nw;
                    w = fclamp(w, 0.000000001f, 1.0f);
                    ts = ts + sa * w;
                    tw = tw + w;
                }
            }
            out[i] = vpt_unorm8_encode(ts / fmaxf(tw, 0.01f));
        }
    }
}

OAPI void oracle_logf_n(const float* x, float* y, int n)
{
    for (int i = 0; i < n; i++) y[i] = vpt_logf(x[i]);
}

OAPI void oracle_powf_n(const float* x, const float* e, float* y, int n)
{
    for (int i = 0; i < n; i++) y[i] = vpt_powf(x[i], e[i]);
}

OAPI int oracle_threads(void) { return omp_get_max_threads(); }

/* ======================================================================= */
/* Array wrappers of the shared numerics (include/vpt/vpt_numerics.h), so  */
/* the tests can pin them against libm / numpy.                            */
/* ======================================================================= */

OAPI void oracle_expf_n(const float* x, float* y, int n)
{
    for (int i = 0; i < n; i++) y[i] = vpt_expf(x[i]);
}

OAPI void oracle_sincos2pi_n(const float* u, float* s, float* c, int n)
{
    for (int i = 0; i < n; i++) vpt_sincos_2pi(u[i], &s[i], &c[i]);
}

OAPI void oracle_f32_to_f16_n(const float* x, uint16_t* h, int n)
{
    for (int i = 0; i < n; i++) h[i] = vpt_f32_to_f16(x[i]);
}

OAPI void oracle_f16_to_f32_n(const uint16_t* h, float* x, int n)
{
    for (int i = 0; i < n; i++) x[i] = vpt_f16_to_f32(h[i]);
}

OAPI void oracle_pcg_n(const uint32_t* v, uint32_t* h, int n)
{
    for (int i = 0; i < n; i++) h[i] = vpt_pcg_hash(v[i]);
}

OAPI void oracle_df_step_table(int32_t* out)
{
    for (int k = 0; k < 256; k++) out[k] = vpt_df_step(k);
}

OAPI void oracle_unorm8_n(const float* x, uint8_t* c, int n)
{
    for (int i = 0; i < n; i++) c[i] = vpt_unorm8_encode(x[i]);
}

OAPI void oracle_powi_n(const float* x, int e, float* y, int n)
{
    for (int i = 0; i < n; i++) y[i] = vpt_powi(x[i], e);
}

OAPI void oracle_unorm8_decode_table(float* out)
{
    for (int k = 0; k < 256; k++) out[k] = vpt_unorm8_decode((uint8_t)k);
}

/* ImportanceSampleGGX (ReflectionTraceFrag.glsl:345-366): the GGX half
 * vector around N for roughness r and randoms (x0, x1), phi = 2 pi x0 through
 * vpt_sincos_2pi. */
static inline V3 ggx_half(V3 N, float r, float x0, float x1)
{
    const float alpha = r * r;
    const float alpha2 = alpha * alpha;
    float sp, cp;
    vpt_sincos_2pi(x0, &sp, &cp);
    const float cos_t = sqrtf((1.0f - x1) / (1.0f + (alpha2 - 1.0f) * x1));
    const float sin_t = sqrtf(1.0f - cos_t * cos_t);
    const V3 h = v3(cp * sin_t, sp * sin_t, cos_t);
    const V3 up = fabsf(N.z) < 0.999f ? v3(0.0f, 0.0f, 1.0f) : v3(1.0f, 0.0f, 0.0f);
    const V3 t = normalize3(cross3(up, N));
    const V3 b = cross3(N, t);
    return normalize3(v3((t.x * h.x + b.x * h.y) + N.x * h.z, (t.y * h.x + b.y * h.y) + N.y * h.z,
                         (t.z * h.x + b.z * h.y) + N.z * h.z));
}

/* GetReflectionDirection (:617-640): of three GGX half vectors, the first
 * with the largest dot with N; randoms drawn from the pcg_hash chain. */
static inline V3 reflection_normal(V3 N, float r, uint32_t* state)
{
    r = fmaxf(r, 0.05f);
    float nearest = -100.0f;
    V3 best = N;
    for (int i = 0; i < 3; i++) {
        *state = vpt_pcg_hash(*state);
        const float x0 = vpt_u32_to_unit(*state) * 0.9f;
        *state = vpt_pcg_hash(*state);
        const float x1 = vpt_u32_to_unit(*state) * 0.65f;
        const V3 h = ggx_half(N, r, x0, x1);
        const float d = dot3(h, N);
        if (d > nearest) {
            best = h;
            nearest = d;
        }
    }
    return best;
}

/* SHToIrradianceA (ReflectionTraceFrag.glsl:466-475): the direction-less
 * decode of a diffuse SH / CoCg texel. */
static inline V3 sh_to_irradiance_a(const float sh[4], const float cc[2])
{
    const float Y = fmaxf(0.0f, (mut(MUT_RS_Y) ? 1.7724525f : 3.544905f) * sh[3]);
    const float k = (Y * 0.282095f) / (sh[3] + 1e-6f);
    const int ic = mut(MUT_RS_COCG) ? 1 : 0;
    const float co = cc[ic] * k, cg = cc[1 - ic] * k;
    const float T = Y - cg * 0.5f;
    const float G = cg + T;
    const float B = T - co * 0.5f;
    const float R = B + co;
    return v3(fmaxf(R, 0.0f), fmaxf(G, 0.0f), fmaxf(B, 0.0f));
}

/* CalculateDirectionalLight of the reflection pass (ReflectionTraceFrag.glsl:
 * 313-336) for a flat dielectric (metalness 0): F0 = mix(0.04, albedo, 0) =
 * 0.04, kd = mix(1 - F, 0, 0) = 1 - F.  Its specular term is multiplied by
 * radiance_s = radiance * 0.05 * 0 = 0; D and G are finite for every
 * roughness in (0, 1] (denominators >= alpha^4 > 0 and k > 0), so that term
 * is exactly +0 and is left out.  pow(x, 5.0) is vpt_powi(x, 5). */
static inline V3 reflect_direct(V3 hp, V3 eye, V3 L, V3 rad, V3 alb, V3 N, float shadow)
{
    const float sh = fminf(shadow, 1.0f);
    const V3 Lo = normalize3(v3(eye.x - hp.x, eye.y - hp.y, eye.z - hp.z));
    const V3 Lh = normalize3(v3(L.x + Lo.x, L.y + Lo.y, L.z + Lo.z));
    const float cos_li = fmaxf(0.0f, dot3(N, L));
    const float F = 0.04f + (1.0f - 0.04f) * vpt_powi(1.0f - fmaxf(0.0f, dot3(Lh, Lo)), 5);
    const float kd = mut(MUT_RS_FRESNEL) ? 1.0f : 1.0f - F;
    const float vis = fclamp(1.0f - sh, 0.0f, 1.0f);
    return v3(fmaxf(((kd * alb.x) * rad.x) * cos_li, 0.0f) * vis, fmaxf(((kd * alb.y) * rad.y) * cos_li, 0.0f) * vis,
              fmaxf(((kd * alb.z) * rad.z) * cos_li, 0.0f) * vis);
}

/* Inputs of the shaded reflection pass (see oracle_trace_reflect_shaded). */
struct ReflShade {
    const float* sh;       /* W x H x 4: the frame's denoised diffuse SH (DiffuseDenoiseFBO 0) */
    const float* cocg;     /* W x H x 2: its CoCg (DiffuseDenoiseFBO 1) */
    const float* ao;       /* W x H: its AO channel (DiffuseDenoiseFBO 3, .x) */
    const float* shadow;   /* W x H: the denoised sun shadow (ShadowFiltered) */
    const float* vp;       /* u_Projection * u_View of the frame */
    V3 sun_rad, sky_rad;   /* SAMPLED_COLOR_MIXED and GetAtmosphere substitutes */
    const float* albedo;   /* 128 x RGB per block id */
};

/* REFLECTION_TRACE (ReflectionTraceFrag.glsl:764-1035) for flat materials
 * (the normal-mapped G-buffer normal is the face normal, the roughness one
 * value), no jitter, LPV and player off.
 *
 * The target is (W / div) x (H / div) (ReflectionTraceFBO, Pipeline.cpp:
 * 1141, 1744: the reference's resolution 0.25 is div = 4).  Each target
 * pixel samples the full-resolution G-buffer at its centre: t LINEAR (the
 * R16F position texture, Pipeline.cpp:1094; at div = 1 the texel itself),
 * face NEAREST, decoded by GetNormalFromID (:227-238, ids above 5 -> 1,1,1).
 *   I = normalize(P - eye)                                   (:783)
 *   o = P + N * 0.035                                        (:785)
 *   per sample s < spp (:828-1023):
 *     H = rough ? reflection_normal(N, clamp(roughness, 0.01, 1)) : N
 *     R = reflect(I, H) = I - 2 dot(H, I) H
 *     T = VoxelTraversalDF(o, R), max_iter steps (TRACE_LENGTH = 64)
 *     on a hit (T > 0): the first max(spp / 4, 1) hits cast the sun shadow
 *     GetShadowAt(o + R T + n * 0.055, L) with shadow_iter steps (150 in
 *     the shader, :1098; 1 when it hits, :1327-1345); sum += T
 *   hit distance: sky -1 (:767-773); else (0.001 + sum) / hits, -1 without
 *   hits (:808, 1024-1031), clamped to [-10, 200] (:1036), stored as
 *   binary16 (R16F, Pipeline.cpp:1141).
 * Mirror samples (rough = 0) repeat one ray: it is traced once and its
 * result counted spp times.  Randoms: the chain state starts at
 * pcg_hash(pcg_hash(px + Wr*py + Wr*Hr*frame) + 0x9E3779B9).
 * eye is the inverse view's translation (the reference's u_ViewerPosition,
 * the position its view matrix was built from).  hit_out (optional):
 * face | block << 8 of the first hit | shadow << 16 | hits << 24, 255
 * without a hit.  samples / rays (optional): DF samples and traced rays.
 *
 * With shading inputs (ReprojectReflectionToScreenSpace and the hit
 * shading, :582-593, 843-1023, u_ReprojectToScreenSpace on) each hit is
 * also reprojected with the frame's view-projection: uv = (VP HP).xy / w
 * * 0.5 + 0.5; the G-buffer position there (t LINEAR, full resolution) and
 * face (NEAREST) must lie within |dP|^2 < 0.095 with the hit's normal, and
 * uv inside the 0.0035 screen border.  Reprojected hits take their ambient
 * from the denoised diffuse there (SHToIrradianceA of LINEAR SH / CoCg,
 * times pow(AO, 0.75) when the position there is geometry within 40 of the
 * eye) and, for the shadow-casting samples, the denoised shadow there
 * instead of a shadow ray; other hits take the ambient at the pixel itself
 * (BaseIndirectDiffuse; the reference's LPV fallback is off).  A hit adds
 * ambient * albedo + reflect_direct(sun radiance * 0.6) of its block, a
 * miss the sky radiance; the colour is their mean over spp, clamped to
 * [1e-7, 100], stored as RGBA binary16 (alpha 1; sky pixels 0).  Bit 17 of
 * the hit record: the first hit was reprojected; the shadow bit: the last
 * shadow value >= 0.5. */
static void trace_reflect_impl(const uint8_t* grid, const uint8_t* df, int nx, int ny, int nz,
                               const float* inv_view, const float* inv_proj, int W, int H,
                               const float* gb_t, const uint8_t* gb_face, const float* light_dir, int div,
                               int max_iter, int shadow_iter, int spp, int rough, float roughness, uint32_t frame,
                               const ReflShade* shade, uint16_t* dist_out, uint32_t* hit_out, uint16_t* color_out,
                               int32_t* samples, int32_t* rays)
{
    const Dims d = {nx, ny, nz};
    const int Wr = W / div, Hr = H / div;
    const Frame f = {inv_view, inv_proj, Wr, Hr};
    const Frame ff = {inv_view, inv_proj, W, H};
    const Tex tt = {gb_t, W, H, 1};
    const V3 L = v3(light_dir[0], light_dir[1], light_dir[2]);
    const V3 eye = f.origin();
    const int shadow_rays = std::max(spp / 4, 1);
    const float rr = fclamp(roughness, 0.01f, 1.0f);
    Tex tsh = {}, tcc = {}, tao = {}, tsd = {};
    V3 rad = v3(0.0f, 0.0f, 0.0f);
    if (shade) {
        tsh = {shade->sh, W, H, 4};
        tcc = {shade->cocg, W, H, 2};
        tao = {shade->ao, W, H, 1};
        tsd = {shade->shadow, W, H, 1};
        const float f = mut(MUT_RS_RAD) ? 1.0f : 0.6f;
        rad = v3(shade->sun_rad.x * f, shade->sun_rad.y * f, shade->sun_rad.z * f);
    }
    const float b = 0.0035f;
#pragma omp parallel for schedule(dynamic, 4)
    for (int ry = 0; ry < Hr; ry++) {
        for (int rx = 0; rx < Wr; rx++) {
            const size_t i = (size_t)ry * Wr + rx;
            const float sx = texcoord(rx, Wr), sy = texcoord(ry, Hr);
            float dist;
            int face;
            if (div == 1) {
                dist = gb_t[i];
                face = gb_face[i];
            } else {
                dist = bilinear(tt, sx, sy, 0);
                face = nearest_u8(gb_face, W, H, sx, sy);
            }
            float hd = -1.0f;
            uint32_t rec = 255u;
            int smp = 0, nr = 0;
            float col[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            if (!(dist < 0.0f)) {
                const V3 P = position_from(f, sx, sy, dist);
                const V3 N = face_normal(face, 1.0f);
                const V3 I = normalize3(v3(P.x - eye.x, P.y - eye.y, P.z - eye.z));
                const V3 o = v3(P.x + N.x * 0.035f, P.y + N.y * 0.035f, P.z + N.z * 0.035f);
                uint32_t state = vpt_pcg_hash(vpt_pcg_hash((uint32_t)rx + (uint32_t)Wr * (uint32_t)ry +
                                                           (uint32_t)Wr * (uint32_t)Hr * frame) + 0x9E3779B9u);
                float sum = 0.001f;
                int hits = 0, shadow_itr = 0, shadow = 0;
                float shadow_f = 0.0f;   /* ComputedShadow (:811) */
                bool first_reproj = false;
                V3 base_amb = v3(0.0f, 0.0f, 0.0f);
                if (shade) {   /* BaseIndirectDiffuse (:804-806) at the target pixel */
                    float bs[4], bc[2];
                    for (int c = 0; c < 4; c++) bs[c] = bilinear(tsh, sx, sy, c);
                    for (int c = 0; c < 2; c++) bc[c] = bilinear(tcc, sx, sy, c);
                    base_amb = sh_to_irradiance_a(bs, bc);
                }
                float tot[3] = {0.0f, 0.0f, 0.0f};
                TraceResult tr = {};
                tr.t = -1.0f;
                V3 R = N;
                for (int s = 0; s < spp; s++) {
                    if (rough || s == 0) {
                        const V3 Hn = rough ? reflection_normal(N, rr, &state) : N;
                        const float k = 2.0f * dot3(Hn, I);
                        R = v3(I.x - k * Hn.x, I.y - k * Hn.y, I.z - k * Hn.z);
                        tr = voxel_traversal_df(grid, df, d, o, R, max_iter);
                        smp += tr.samples;
                        nr++;
                    }
                    if (tr.t > 0.0f) {
                        /* the traversal's normal vector (normal[MinIdx] =
                         * -RaySign[MinIdx]), never zero on a hit: a step
                         * along an axis whose direction component is 0
                         * moves the other coordinates by d * inf, and the
                         * ray leaves the volume */
                        const V3 hn = face_normal(tr.face, 0.0f);
                        const V3 hp = v3(o.x + R.x * tr.t, o.y + R.y * tr.t, o.z + R.z * tr.t);
                        bool reproj = false;
                        float ru = 0.0f, rv = 0.0f, t_at = 0.0f;
                        V3 p_at = v3(0.0f, 0.0f, 0.0f);
                        if (shade) {   /* ReprojectReflectionToScreenSpace (:582-593) */
                            float p4[4] = {hp.x, hp.y, hp.z, 1.0f}, q[4];
                            mat_vec4(shade->vp, p4, q);
                            ru = (q[0] / q[3]) * 0.5f + 0.5f;
                            rv = (q[1] / q[3]) * 0.5f + 0.5f;
                            /* the test ANDs the screen border with the
                             * position and normal checks: outside it (also
                             * inf / NaN behind the eye) the fetches are not
                             * needed */
                            if (ru < 1.0f - b && ru > b && rv < 1.0f - b && rv > b) {
                                t_at = bilinear(tt, ru, rv, 0);
                                p_at = position_from(ff, ru, rv, t_at);
                                const V3 n_at = face_normal(nearest_u8(gb_face, W, H, ru, rv), 1.0f);
                                const V3 pd = v3(fabsf(p_at.x - hp.x), fabsf(p_at.y - hp.y), fabsf(p_at.z - hp.z));
                                const float err = dot3(pd, pd);
                                reproj = err < 0.095f && n_at.x == hn.x && n_at.y == hn.y && n_at.z == hn.z;
                            }
                            if (hits == 0) first_reproj = reproj;
                        }
                        if (shadow_itr < shadow_rays) {
                            if (shade && reproj && !mut(MUT_RS_SHADOW_TRACE)) {
                                shadow_f = bilinear(tsd, ru, rv, 0);   /* :934-937 */
                            } else if (rough || shadow_itr == 0) {
                                const V3 so = v3(hp.x + hn.x * 0.055f, hp.y + hn.y * 0.055f, hp.z + hn.z * 0.055f);
                                const TraceResult sr = voxel_traversal_df(grid, df, d, so, L, shadow_iter);
                                smp += sr.samples;
                                nr++;
                                shadow = sr.t > 0.0f ? 1 : 0;
                                shadow_f = (float)shadow;
                            }
                            shadow_itr++;
                        }
                        if (shade) {
                            V3 amb = base_amb;
                            if (reproj && !mut(MUT_RS_BASE)) {
                                float rs[4], rc[2];
                                for (int c = 0; c < 4; c++) rs[c] = bilinear(tsh, ru, rv, c);
                                for (int c = 0; c < 2; c++) rc[c] = bilinear(tcc, ru, rv, c);
                                amb = sh_to_irradiance_a(rs, rc);
                                const float ao_at = bilinear(tao, ru, rv, 0);
                                const float vx = mut(MUT_RS_AO) ? ao_at : vpt_powf(ao_at, 0.75f);
                                if (t_at > 0.0f) {
                                    const V3 dv = v3(p_at.x - eye.x, p_at.y - eye.y, p_at.z - eye.z);
                                    if (sqrtf(dot3(dv, dv)) < 40.0f || mut(MUT_RS_DIST))
                                        amb = v3(amb.x * vx, amb.y * vx, amb.z * vx);
                                }
                            }
                            const int ref = std::min(std::max(tr.block, 0), 127);
                            const V3 alb = v3(shade->albedo[3 * ref], shade->albedo[3 * ref + 1], shade->albedo[3 * ref + 2]);
                            if (!mut(MUT_RS_ALBEDO))   /* (Ambient * 1 * clamp(AO, 0.1, 1)) * Albedo */
                                amb = v3(amb.x * alb.x, amb.y * alb.y, amb.z * alb.z);
                            const V3 dl = reflect_direct(hp, eye, L, rad, alb, hn, shadow_f);
                            tot[0] = tot[0] + (amb.x + dl.x);
                            tot[1] = tot[1] + (amb.y + dl.y);
                            tot[2] = tot[2] + (amb.z + dl.z);
                        }
                        if (hits == 0) rec = (uint32_t)tr.face | ((uint32_t)tr.block << 8);
                        sum += tr.t;
                        hits++;
                    } else if (shade && !mut(MUT_RS_SKY)) {
                        tot[0] = tot[0] + shade->sky_rad.x;
                        tot[1] = tot[1] + shade->sky_rad.y;
                        tot[2] = tot[2] + shade->sky_rad.z;
                    }
                }
                if (shade) shadow = shadow_f >= 0.5f ? 1 : 0;
                if (hits > 0) {
                    hd = sum / fmaxf((float)hits, 0.01f);
                    rec |= ((uint32_t)shadow << 16) | ((uint32_t)hits << 24) | ((uint32_t)first_reproj << 17);
                }
                hd = fclamp(hd, -10.0f, 200.0f);
                const float n = (float)spp;
                for (int c = 0; c < 3; c++) col[c] = fclamp(mut(MUT_RS_MEAN) ? tot[c] : tot[c] / n, 0.0000001f, 100.0f);
                col[3] = fclamp(n / n, 0.0000001f, 100.0f);
            }
            dist_out[i] = vpt_f32_to_f16(hd);
            if (hit_out) hit_out[i] = rec;
            if (color_out)
                for (int c = 0; c < 4; c++) color_out[4 * i + c] = vpt_f32_to_f16(col[c]);
            if (samples) samples[i] = smp;
            if (rays) rays[i] = nr;
        }
    }
}

OAPI void oracle_trace_reflect(const uint8_t* grid, const uint8_t* df, int nx, int ny, int nz,
                               const float* inv_view, const float* inv_proj, int W, int H,
                               const float* gb_t, const uint8_t* gb_face, const float* light_dir, int div,
                               int max_iter, int shadow_iter, int spp, int rough, float roughness, uint32_t frame,
                               uint16_t* dist_out, uint32_t* hit_out, int32_t* samples, int32_t* rays)
{
    trace_reflect_impl(grid, df, nx, ny, nz, inv_view, inv_proj, W, H, gb_t, gb_face, light_dir, div, max_iter,
                       shadow_iter, spp, rough, roughness, frame, nullptr, dist_out, hit_out, nullptr, samples, rays);
}

/* The reflection pass with its hit shading (see trace_reflect_impl). */
OAPI void oracle_trace_reflect_shaded(const uint8_t* grid, const uint8_t* df, int nx, int ny, int nz,
                                      const float* inv_view, const float* inv_proj, int W, int H,
                                      const float* gb_t, const uint8_t* gb_face, const float* light_dir, int div,
                                      int max_iter, int shadow_iter, int spp, int rough, float roughness,
                                      uint32_t frame, const float* sh, const float* cocg, const float* ao,
                                      const float* shadow, const float* viewproj, const float* sun_rad,
                                      const float* sky_rad, const float* albedo, uint16_t* dist_out, uint32_t* hit_out,
                                      uint16_t* color_out, int32_t* samples, int32_t* rays)
{
    ReflShade rs;
    rs.sh = sh;
    rs.cocg = cocg;
    rs.ao = ao;
    rs.shadow = shadow;
    rs.vp = viewproj;
    rs.sun_rad = v3(sun_rad[0], sun_rad[1], sun_rad[2]);
    rs.sky_rad = v3(sky_rad[0], sky_rad[1], sky_rad[2]);
    rs.albedo = albedo;
    trace_reflect_impl(grid, df, nx, ny, nz, inv_view, inv_proj, W, H, gb_t, gb_face, light_dir, div, max_iter,
                       shadow_iter, spp, rough, roughness, frame, &rs, dist_out, hit_out, color_out, samples, rays);
}

/* ======================================================================= */
/* The colour pass's composite (ColorPassFrag.glsl main, :773-946, with    */
/* u_UseDFG, Pipeline.cpp:3683) for flat dielectric materials.             */
/* ======================================================================= */

/* FresnelSchlickRoughness(Eye, N, F0, r) (ColorPassFrag.glsl:1205-1209),
 * scalar F0. */
static inline float fresnel_schlick_roughness(V3 eye, V3 n, float f0, float r)
{
    if (mut(MUT_CMP_F0)) f0 = 0.02f;
    const float c = fclamp(dot3(eye, n), 0.00001f, 1.0f);
    return f0 + (fmaxf(1.0f - r, f0) - f0) * vpt_powi(1.0f - c, 5);
}

/* CalculateDirectionalLight of the colour pass (ColorPassFrag.glsl:419-451),
 * Cook-Torrance with its GGX specular term, metalness 0. */
static inline V3 color_direct(V3 P, V3 eye, V3 Li, V3 rad, V3 alb, V3 N, float r, float shadow)
{
    const float PI = 3.14159265359f;
    const float sh = fminf(shadow, 1.0f);
    const V3 Lo = normalize3(v3(eye.x - P.x, eye.y - P.y, eye.z - P.z));
    const float cos_lo = fmaxf(0.0f, dot3(N, Lo));
    const V3 Lh = normalize3(v3(Li.x + Lo.x, Li.y + Lo.y, Li.z + Lo.z));
    const float cos_li = fmaxf(0.0f, dot3(N, Li));
    const float cos_lh = fmaxf(0.0f, dot3(N, Lh));
    const float F = fresnel_schlick_roughness(Lo, N, 0.04f, r);
    const float alpha = mut(MUT_CMP_ALPHA) ? r : r * r;
    const float alpha2 = alpha * alpha;
    const float denom = (cos_lh * cos_lh) * (alpha2 - 1.0f) + 1.0f;
    const float D = alpha2 / (((mut(MUT_CMP_D_PI) ? 1.0f : PI) * denom) * denom);
    const float rk = r + 1.0f;
    const float k = (rk * rk) / 8.0f;
    const float G = (cos_li / (cos_li * (1.0f - k) + k)) * (cos_lo / (cos_lo * (1.0f - k) + k));
    const float kd = 1.0f - F;
    const float spec = fclamp(((F * D) * G) / fmaxf(0.00001f, (4.0f * cos_li) * cos_lo), 0.0f, 2.0f);
    const float vis = fclamp(1.0f - sh, 0.0f, 1.0f);
    const float a[3] = {alb.x, alb.y, alb.z}, l[3] = {rad.x, rad.y, rad.z};
    float o[3];
    for (int c = 0; c < 3; c++)
        o[c] = fclamp(((kd * a[c]) * l[c]) * cos_li + (spec * l[c]) * cos_li, 0.0f, 2.5f) * vis;
    return v3(o[0], o[1], o[2]);
}

/* KarisEnvBRDFApprox (ColorPassFrag.glsl:385-392). */
static inline void karis_env_brdf(float ndotv, float r, float* ax, float* ay)
{
    const float rx = r * -1.0f + 1.0f, ry = r * -0.0275f + 0.0425f, rz = r * -0.572f + 1.040f,
                rw = r * 0.022f + -0.040f;
    const float a004 = fminf(rx * rx, vpt_exp2f(-9.28f * ndotv)) * rx + ry;
    *ax = -1.04f * a004 + rz;
    *ay = (mut(MUT_CMP_KARIS) ? -1.04f : 1.04f) * a004 + rw;
}

/* Per pixel (uv at its centre): Dist = 1 / (1 / t) (SamplePositionAt reads
 * the R32F 1/t plane, :1122-1126); Dist <= 0 is sky (:803, 1061-1066), the
 * sky pixel radiance (GetAtmosphereAndClouds' substitute).  Otherwise, with
 * the workload's substitutions (albedo per block id for the albedo G-buffer,
 * the face normal for the normal map, PBR = (roughness, 0, -, 1) with no
 * emission, full-resolution buffers for SpatiallyUpscaleBuffers, the
 * reflection colour sampled LINEAR for the upscaled specular, no player):
 *   Albedo = BasicSaturation(albedo, 1 - 0.1)                 (:812)
 *   shadow = clamp(denoised shadow, 0, 1)                    (:837-840, 304-336)
 *   Ed = SHToIrridiance(SH, CoCg, N) * clamp(pow(mix(AO, 1, fade), 2.2),
 *        0.05, 1) inside the 0.00125 border, fade = |P.xz - eye.xz| >= 256
 *                                                             (:848-876)
 *   Direct = max(mix(CDL(sun), CDL(moon), SunVisibility), 1e-6) (:885-891)
 *   Spec = max(0.0001, refl * clamp(pow(max(N.N, 1e-8), 2) 1.5, 0.1, 1)
 *          inside the 0.0035 border) * SpecularAO^1.5 * Horizon^3
 *   F = clamp(FresnelSchlickRoughness(Lo, N, 0.04, r), 0, 1)
 *   out = Direct + ((1 - F) * Ed * Albedo * clamp(1, 0.2, 1.01) +
 *         Spec * (F * Env.x + Env.y))                          (:914-946)
 * The exponent exp2(-16 r - 1) of ComputeSpecularAO (:766-769) is one value
 * per call, (float)exp2((double)(-16 r - 1)) on the host, as the library's
 * launcher computes it.  refl: (W / rdiv) x (H / rdiv) x 4 binary16 values. */
OAPI void oracle_composite(const float* inv_view, const float* inv_proj, int W, int H, const float* gb_t,
                           const uint8_t* gb_face, const uint8_t* gb_block, const float* sh, const float* cocg,
                           const float* ao, const float* shadow, const float* refl, int rdiv, const float* sun_dir,
                           const float* moon_dir, const float* sun_rad, const float* moon_rad, const float* sky_pixel,
                           const float* albedo, float roughness, float* out)
{
    const float spec_ao_exp = (float)exp2((double)(-16.0f * roughness - 1.0f));
    const Frame f = {inv_view, inv_proj, W, H};
    const V3 eye = f.origin();
    const int Wr = std::max(W / rdiv, 1), Hr = std::max(H / rdiv, 1);
    const Tex trf = {refl, Wr, Hr, 4};
    const V3 Ls = v3(sun_dir[0], sun_dir[1], sun_dir[2]), Lm = v3(moon_dir[0], moon_dir[1], moon_dir[2]);
    const V3 Rs = v3(sun_rad[0], sun_rad[1], sun_rad[2]), Rm = v3(moon_rad[0], moon_rad[1], moon_rad[2]);
    const float r = roughness;
    float sun_vis = fclamp(sun_dir[1] + 0.05f, 0.0f, 0.1f) * 12.0f;
    sun_vis = 1.0f - sun_vis;
    if (mut(MUT_CMP_SUNVIS)) sun_vis = fclamp(sun_vis, 0.0f, 1.0f);
#pragma omp parallel for schedule(static)
    for (int py = 0; py < H; py++) {
        for (int px = 0; px < W; px++) {
            const size_t i = (size_t)py * W + px;
            const float u = texcoord(px, W), v = texcoord(py, H);
            const float dist = 1.0f / (1.0f / gb_t[i]);
            float* o = out + 3 * i;
            if (!(dist > 0.0f)) {
                const float z = mut(MUT_CMP_SKY) ? 0.0f : 1.0f;
                o[0] = sky_pixel[0] * z;
                o[1] = sky_pixel[1] * z;
                o[2] = sky_pixel[2] * z;
                continue;
            }
            const V3 P = position_from(f, u, v, dist);
            const V3 N = face_normal(gb_face[i], 1.0f);
            const int ref = gb_block[i] & 127;
            const V3 a0 = v3(albedo[3 * ref], albedo[3 * ref + 1], albedo[3 * ref + 2]);
            const float adj = mut(MUT_CMP_SAT) ? 0.1f : 1.0f - 0.1f;
            const float lum = (a0.x * 0.2125f + a0.y * 0.7154f) + a0.z * 0.0721f;
            const V3 A = v3(lum * (1.0f - adj) + a0.x * adj, lum * (1.0f - adj) + a0.y * adj,
                            lum * (1.0f - adj) + a0.z * adj);
            const float shd = mut(MUT_CMP_SHADOW)
                                  ? 0.0f
                                  : fclamp(fclamp(fclamp(shadow[i], 0.0f, 1.0f) + 0.0f, 0.0f, 1.0f), 0.0f, 1.0f);
            /* SHToIrridiance(SHy, CoCg, N) (:356-368) */
            const float* s4 = sh + 4 * i;
            const float x = (s4[0] * N.x + s4[1] * N.y) + s4[2] * N.z;
            float Y = 2.0f * ((mut(MUT_CMP_SH_L1) ? 0.511664f : 1.023326f) * x + 0.886226f * s4[3]);
            Y = fmaxf(Y, 0.0f);
            const float kk = (Y * 0.282095f) / (s4[3] + 1e-6f);
            const int ic = mut(MUT_CMP_COCG) ? 1 : 0;
            const float co = cocg[2 * i + ic] * kk, cg = cocg[2 * i + 1 - ic] * kk;
            const float T = Y - cg * 0.5f;
            const float G = cg + T;
            const float B = T - co * 0.5f;
            const float Rr = B + co;
            V3 ed = v3(fmaxf(Rr, 0.0f), fmaxf(G, 0.0f), fmaxf(B, 0.0f));
            const float vb = 0.00125f;
            if (u > vb && u < 1.0f - vb && v > vb && v < 1.0f - vb) {
                float aov = ao[i];
                const float dx = P.x - eye.x, dz = P.z - eye.z;
                const float fade = sqrtf(dx * dx + dz * dz) < 256.0f ? 0.0f : 1.0f;
                aov = aov * (1.0f - fade) + 1.0f * fade;
                if (!mut(MUT_CMP_AO_EXP)) aov = vpt_powf(aov, 2.2f);
                const float m = fclamp(aov, 0.05f, 1.0f);
                ed = v3(ed.x * m, ed.y * m, ed.z * m);
            }
            const V3 sd = color_direct(P, eye, Ls, Rs, A, N, r, shd);
            const V3 md = color_direct(P, eye, Lm, Rm, A, N, r, shd);
            V3 dl = v3(sd.x * (1.0f - sun_vis) + md.x * sun_vis, sd.y * (1.0f - sun_vis) + md.y * sun_vis,
                       sd.z * (1.0f - sun_vis) + md.z * sun_vis);
            dl = v3(fmaxf(dl.x, 0.000001f), fmaxf(dl.y, 0.000001f), fmaxf(dl.z, 0.000001f));
            const V3 di = v3(ed.x * A.x, ed.y * A.y, ed.z * A.z);
            V3 si = v3(0.0f, 0.0f, 0.0f);
            const V3 I = normalize3(v3(P.x - eye.x, P.y - eye.y, P.z - eye.z));
            const float kr = 2.0f * dot3(N, I);
            const V3 R = v3(I.x - kr * N.x, I.y - kr * N.y, I.z - kr * N.z);
            const float sb = 0.0035f;
            if ((u > sb && u < 1.0f - sb && v > sb && v < 1.0f - sb) || mut(MUT_CMP_SPEC_BORDER)) {
                const float att = fclamp(vpt_powi(fmaxf(dot3(N, N), 0.00000001f), 2) * 1.5f, 0.1f, 1.0f);
                si = v3(bilinear(trf, u, v, 0) * att, bilinear(trf, u, v, 1) * att, bilinear(trf, u, v, 2) * att);
            }
            si = v3(fmaxf(0.0001f, si.x), fmaxf(0.0001f, si.y), fmaxf(0.0001f, si.z));
            const V3 Lo = normalize3(v3(eye.x - P.x, eye.y - P.y, eye.z - P.z));
            const float nov = fclamp(dot3(N, Lo), 0.0000000001f, 1.0f);
            float sao = fclamp((vpt_powf(nov + 1.0f, spec_ao_exp) - 1.0f) + 1.0f, 0.0f, 1.0f);
            sao = vpt_powf(sao, 1.5f);
            float hz = fminf(1.0f + dot3(R, N), 1.0f);
            hz = hz * (hz * hz);
            si = v3((si.x * sao) * hz, (si.y * sao) * hz, (si.z * sao) * hz);
            const float F = fclamp(fresnel_schlick_roughness(Lo, N, 0.04f, r), 0.0f, 1.0f);
            const float kd = mut(MUT_CMP_KD) ? 1.0f : (1.0f - F) * (1.0f - 0.0f);
            float ex, ey;
            karis_env_brdf(fclamp(fmaxf(-dot3(I, N), 0.000001f), 0.0f, 1.0f), fclamp(r, 0.0f, 1.0f), &ex, &ey);
            const float fs = F * ex + ey;
            const float ao_d = fclamp(1.0f, 0.2f, 1.01f);
            o[0] = dl.x + (kd * (di.x * ao_d) + si.x * fs);
            o[1] = dl.y + (kd * (di.y * ao_d) + si.y * fs);
            o[2] = dl.z + (kd * (di.z * ao_d) + si.z * fs);
        }
    }
}
