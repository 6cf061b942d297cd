/*
 * vpt.h — C ABI of the MI355X (gfx950) voxel path tracer hot path.
 *
 * Replaces the reference's GL-level boundary for three pieces:
 *   - the distance-field build     (World::InitializeDistanceGenerator /
 *                                    World::GenerateDistanceField,
 *                                    Core/World.cpp:48-113, World.h:173-174)
 *   - the DF-accelerated traversal (INITIAL_TRACE / SHADOW_TRACE /
 *                                    DIFFUSE_TRACE passes, selected in
 *                                    Core/ShaderManager.cpp:8-44 and drawn in
 *                                    Core/Pipeline.cpp:1963-2016, 2171-2281,
 *                                    2790-2852)
 *   - the SVGF chain               (SPATIAL_INITIAL + SVGF_TEMPORAL +
 *                                    SVGF_VARIANCE + 5x SVGF_SPATIAL,
 *                                    Core/Pipeline.cpp:2284-2611)
 *
 * Conventions
 *   - Every function returns 0 on success or a negative VPT_E* code;
 *     vpt_last_error() returns a message for the calling thread.  (The
 *     reference reports errors by printf / throw const char*,
 *     ShaderManager.cpp:96,111,126.)
 *   - Pointers named d_* are device pointers; all others are host pointers
 *     borrowed for the duration of the call.
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *     Work is enqueued asynchronously on it; host-pointer outputs are
 *     complete when the call returns.
 *   - One host thread per object at a time; concurrency comes from streams.
 *     Calls on different streams may share a context and its worlds: the
 *     scratch a call needs (the GI pass's shadow-ray jobs) is kept per
 *     stream, and its constants (camera, GI parameters, the albedo table)
 *     travel with its launches, so no call reads what another stream's call
 *     writes.  The caller orders its own buffers (G-buffers, outputs).
 *     (vpt_trace_gi_noalloc is the exception: see there.)
 *   - Matrices are 4x4 float, column-major (GLSL / glm layout).
 *   - Images are row-major W x H; row 0 is the bottom row (v_TexCoords.y
 *     smallest), pixel centre uv = ((px + 0.5) / W, (py + 0.5) / H).
 *   - World grids are uint8 block ids, 0 = air, linear index
 *     x + y*nx + z*nx*ny (World.h:46-70).
 */
#ifndef VPT_H
#define VPT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VPT_API __attribute__((visibility("default")))

enum {
    VPT_OK = 0,
    VPT_EINVAL = -1,     /* bad argument (dims, null pointer, sizes) */
    VPT_ENOMEM = -2,     /* device allocation failed / budget exceeded */
    VPT_EDEVICE = -3,    /* HIP runtime error */
    VPT_ESTATE = -4,     /* object used in the wrong state */
    VPT_ETIMEOUT = -5,   /* a wait ended before the work completed */
    VPT_EBUSY = -6,      /* the object is still in use */
};

typedef struct vpt_ctx vpt_ctx;
typedef struct vpt_world vpt_world;
typedef struct vpt_svgf vpt_svgf;
typedef struct vpt_shadow_denoise vpt_shadow_denoise;

/* Camera as the passes see it: inverse view / inverse projection
 * (u_InverseView, u_InverseProjection; ray origin = inv_view[12..14]). */
typedef struct vpt_camera {
    float inv_view[16];
    float inv_proj[16];
    int32_t width, height;
} vpt_camera;

/* G-buffer written by vpt_trace_primary (InitialTraceFBO, Pipeline.cpp:1094):
 *   t      hit distance (o_HitDistance; -1 on a miss)                 fp32
 *   face   normal id 0..5 (+z,-z,+y,-y,-x,+x), 255 on a miss (o_Normal) u8
 *   block  hit block id, 0 on a miss (o_BlockID)                       u8
 *   voxel  hit cell x | y<<10 | z<<20, 0xffffffff on a miss            u32
 * All four are device pointers of width*height elements. */
typedef struct vpt_gbuffer {
    float* d_t;
    uint8_t* d_face;
    uint8_t* d_block;
    uint32_t* d_voxel;
    int32_t width, height;
} vpt_gbuffer;

/* Frame constants of the diffuse GI pass (DiffuseRayTraceFrag.glsl:822-840)
 * with the workload's substitutions: PCG randoms seeded by
 * pcg_hash(px + W*py + W*H*frame), Lambert sun term, per-id albedo and a
 * constant sky. */
typedef struct vpt_gi_params {
    float sun_dir[3];            /* StrongerLightDirection, normalised */
    float light_color[3];        /* LIGHT_COLOR */
    float sky_radiance[3];       /* radiance of a GI ray that escapes */
    float sky_pixel_radiance[3]; /* radiance written for primary-miss pixels */
    float albedo[128 * 3];       /* Lambert albedo per block id */
    uint32_t frame;
    int32_t gi_iter;             /* traversal cap of the GI ray (48) */
    int32_t shadow_iter;         /* traversal cap of the GI-hit shadow ray (128) */
} vpt_gi_params;

/* Per-pixel SVGF input planes written by vpt_trace_gi when non-NULL
 * (DiffuseRawTraceFBO, Pipeline.cpp:1102): one 16-byte texel per pixel,
 * uint16 halves [SH0 SH1 SH2 SH3 Co Cg luma aosky], where the last half
 * holds the two unorm8 bytes (AO, sky) as (sky << 8) | ao. */
typedef struct vpt_svgf_texel {
    uint16_t h[8];
} vpt_svgf_texel;

/* An arbitrary ray (ray-query API): origin and unit direction. */
typedef struct vpt_ray {
    float origin[3];
    float dir[3];
} vpt_ray;

/* Result of one ray query. */
typedef struct vpt_hit {
    float t;          /* -1 on a miss */
    uint32_t voxel;   /* packed cell as in vpt_gbuffer, ~0 on a miss */
    uint8_t face;     /* 0..5, 255 on a miss */
    uint8_t block;    /* 0 on a miss */
    uint16_t samples; /* DF samples taken (loop iterations inside the volume) */
} vpt_hit;

/* ---- diagnostics ------------------------------------------------------ */

VPT_API const char* vpt_last_error(void);
VPT_API const char* vpt_version(void);

/* Device self-check, with no counterpart in the reference: the shadow
 * filter's pow computes vpt_logf's (m - 1) / (m + 1) through the hardware
 * reciprocal, one Newton step and one correction; this compares that
 * sequence with the IEEE division for every float m in
 * [sqrt(2) / 2, sqrt(2)] on the current device and stores the number of
 * differing m in *mismatches (0 on a conforming device).  Synchronous on
 * `stream`. */
VPT_API int vpt_check_log_division(unsigned long long* mismatches, void* stream);

/* ---- context ----------------------------------------------------------- */

/* Binds `device` (HIP ordinal) for the lifetime of the context. */
VPT_API int vpt_ctx_create(int device, vpt_ctx** out);
VPT_API int vpt_ctx_destroy(vpt_ctx* ctx);
VPT_API int vpt_ctx_device(const vpt_ctx* ctx, int* device);

/* ---- world (World.h:35-222) ----------------------------------------------
 * Device-resident block grid plus distance field, both stored in 8x8x8
 * bricks with Morton order inside a brick.  nx must be a multiple of 16,
 * ny and nz multiples of 8, all in [8, 1024] (nx >= 16). */
VPT_API int vpt_world_create(vpt_ctx* ctx, int nx, int ny, int nz, vpt_world** out);
VPT_API int vpt_world_destroy(vpt_world* w);
VPT_API int vpt_world_dims(const vpt_world* w, int* nx, int* ny, int* nz);
/* Device bytes held by the world (grid + DF + staging). */
VPT_API int vpt_world_bytes(const vpt_world* w, size_t* bytes);

/* World::Buffer (World.h:167-171, Texture3D.cpp:27): upload a linear grid
 * from host memory. */
VPT_API int vpt_world_upload(vpt_world* w, const uint8_t* host_grid, void* stream);
/* Same from a linear grid already in device memory. */
VPT_API int vpt_world_upload_device(vpt_world* w, const uint8_t* d_grid, void* stream);

/* Batched World::SetBlock (World.h:51-70) followed by the single-voxel
 * glTexSubImage3D of World.cpp:367-374.  edits[i] = x | y<<10 | z<<20 and
 * ids[i] the new block id; a voxel named more than once takes its last id,
 * as sequential SetBlock calls would.  Does not rebuild the DF. */
VPT_API int vpt_world_set_blocks(vpt_world* w, const uint32_t* edits, const uint8_t* ids, int n, void* stream);

/* World::GenerateDistanceField (World.cpp:69-113):
 * DF = min(254, L1 distance to the nearest block with id > 0). */
VPT_API int vpt_df_build(vpt_world* w, void* stream);

/* Raw world files (SaveWorld / LoadWorld, WorldFileHandler.cpp:10-83): the
 * nx*ny*nz block-id bytes in the linear layout, no header.  Load checks the
 * file size against the world's dims and uploads; it does not rebuild the
 * DF: call vpt_df_build after it, as the reference does after LoadWorld
 * (Pipeline.cpp:1206, 1293-1294). */
VPT_API int vpt_world_save(vpt_world* w, const char* path);
VPT_API int vpt_world_load(vpt_world* w, const char* path, void* stream);

/* Minecraft world import: MCWorldImporter::ImportWorld (NBT/Importer.cpp:
 * 150-168, ImportRegionFile :86-148) with the enkiMI chunk reader it links
 * (Dependencies/enkiMI/enkimi.c).  Host only, no device work: every *.mca
 * region file in `dir` (in name order) is decoded into host_grid (nx*ny*nz
 * bytes, x fastest, cleared first), for the caller to upload with
 * vpt_world_upload and follow with vpt_df_build, as the reference does
 * (Pipeline.cpp:1236, 1293-1294).
 *   Chunk formats: before 1.13 (Level.Sections Blocks / Data), 1.13-1.17
 *   (Level.Sections Palette / BlockStates) and 1.18+ (sections[].block_states).
 *   Palette names resolve through `names` to legacy (id, data value); a name
 *   not listed is stone (1:0), as in the reference's table lookup.
 *   Only voxels with data value 0 are written, as mc_to_block[id]
 *   (BlockDatabase::GetIDFromMCID; mc_to_block[0] must be 0), at
 *   (x, y, z) - (int)origin + (nx/2, 0, nz/2) when inside the world (the
 *   reference's fixed 384 x 128 x 384 world is nx = 384, nz = 384).
 *   stats (optional, 3 ints): region files read, chunks decoded, sections
 *   imported.  VPT_EINVAL for a missing directory or unreadable file. */
typedef struct vpt_mc_block {
    const char* name;   /* namespaced id, e.g. "minecraft:stone" */
    uint8_t id;         /* legacy block id */
    uint8_t data;       /* legacy data value */
} vpt_mc_block;
VPT_API int vpt_mc_import(const char* dir, int nx, int ny, int nz, const float origin[3],
                          const uint8_t mc_to_block[256], const vpt_mc_block* names, int n_names, uint8_t* host_grid,
                          int32_t* stats);

/* Incremental GenerateDistanceField after vpt_world_set_blocks with the same
 * edit list (the reference rebuilds the whole field after an edit,
 * World.cpp:367-375 -> World.cpp:69-113).  One workgroup finds, on the
 * device, a box around the edits outside which no distance can change
 * (every cell on each of its faces has an old distance below its L1
 * distance to the edits' bounding box) and rebuilds the field inside it
 * from its blocks and the unchanged one-cell shell around it, in LDS.  A box
 * that does not fit (128 KiB with its shell) runs the full rebuild instead,
 * enqueued behind a flag the search sets, so the call never waits for the
 * device unless `region` is given: then it synchronises `stream` and
 * region[6] receives the recomputed box [x0, y0, z0, x1, y1, z1) (the whole
 * world after a full rebuild, all zero for an empty list).  The result
 * equals vpt_df_build's; edits outside the world are VPT_EINVAL.  Calls on
 * one world must not overlap (the world holds one search record). */
VPT_API int vpt_df_update(vpt_world* w, const uint32_t* edits, int n, int32_t* region, void* stream);

/* Read back (parity tests / World::GetBlock): linear x-fastest layout. */
VPT_API int vpt_df_download(vpt_world* w, uint8_t* host_df, void* stream);
VPT_API int vpt_grid_download(vpt_world* w, uint8_t* host_grid, void* stream);

/* ---- rays ---------------------------------------------------------------- */

/* INITIAL_TRACE (InitialRayTraceFrag.glsl:417-468): one primary ray per
 * pixel, at most max_iter traversal iterations (u_RenderDistance). */
VPT_API int vpt_trace_primary(vpt_world* w, const vpt_camera* cam, int max_iter, const vpt_gbuffer* out, void* stream);

/* SHADOW_TRACE (ShadowRayTraceFrag.glsl:414-513) at the G-buffer's
 * resolution, unjittered: d_mask[i] = 1 in shadow, 0 lit or sky. */
VPT_API int vpt_trace_shadow(vpt_world* w, const vpt_camera* cam, const vpt_gbuffer* gb, const float sun_dir[3],
                             int max_iter, uint8_t* d_mask, void* stream);

/* SHADOW_TRACE with its second output (ShadowRayTraceFrag.glsl:7-8,
 * 505-511), the input of the shadow denoiser: the same d_mask, and
 * d_transversal (optional, binary16 bits per pixel) = clamp(T / 100, 1e-5,
 * 196) of the shadow ray, 4.25 / 100 when it misses, 1 / 100 when
 * N.L <= 0.01, 64 for sky.  Unlike vpt_trace_shadow, the ray is also traced
 * when its origin cell is solid, as in the reference. */
VPT_API int vpt_trace_shadow2(vpt_world* w, const vpt_camera* cam, const vpt_gbuffer* gb, const float sun_dir[3],
                              int max_iter, uint8_t* d_mask, uint16_t* d_transversal, void* stream);

/* REFLECTION_TRACE (ReflectionTraceFrag.glsl:764-1035) reduced to its
 * traversal core: flat materials (the reflection normal is the G-buffer
 * face normal, the roughness one value per call), no jitter, screen-space
 * reprojection, LPV and player off.  Textured shading of the hits (albedo,
 * PBR, emissive, normal maps) is not computed.
 *   target   (width / div) x (height / div), div 1..16; the reference's
 *            ReflectionTraceResolution 0.25 (Pipeline.cpp:101, 1744) is
 *            div = 4.  Its pixels sample the G-buffer at their centres (t
 *            LINEAR, face NEAREST).
 *   rays     per sample (u_SPP, 1..16): from P + N * 0.035 in direction
 *            reflect(I, H), I = normalize(P - camera position), at most
 *            max_iter steps (u_ReflectionTraceLength, 64).  H = N for
 *            mirror reflections (rough = 0); with rough = 1
 *            (u_RoughReflections) H is the GGX half vector nearest N of
 *            three (GetReflectionDirection, :617-640, ImportanceSampleGGX
 *            :345-366) at roughness max(clamp(roughness, 0.01, 1), 0.05),
 *            from randoms u = pcg_hash chain of (px + W*py + W*H*frame) in
 *            place of the reference's blue-noise table, scaled by (0.9, 0.65).
 *            The first max(spp / 4, 1) samples that hit cast the sun shadow
 *            ray from hit + normal * 0.055 toward light_dir (shadow_iter
 *            steps, 150 in the shader).
 *   d_dist   binary16 bits per target pixel of o_HitDistance (R16F,
 *            Pipeline.cpp:1141): -1 for sky and when no sample hits, else
 *            (0.001 + sum of T) / hits, clamped to [-10, 200].
 *   d_hit    (optional) face | block << 8 of the first sample that hit |
 *            shadow << 16 (the last shadow ray's result) | hits << 24; 255
 *            without a hit. */
typedef struct vpt_reflect_params {
    float light_dir[3];     /* u_StrongerLightDirection */
    int32_t div;            /* target = G-buffer / div */
    int32_t max_iter;       /* reflected ray: traversal steps */
    int32_t shadow_iter;    /* sun shadow ray: traversal steps */
    int32_t spp;            /* u_SPP */
    int32_t rough;          /* u_RoughReflections */
    float roughness;        /* RoughnessAt * RoughnessBias of the flat material */
    uint32_t frame;         /* seed of the GGX samples */
} vpt_reflect_params;

/* The reference's defaults (Pipeline.cpp:100-114): div 4, 64 / 150 steps,
 * 2 samples, rough reflections at roughness 0.5 * 0.85 (RoughnessBias). */
VPT_API int vpt_reflect_params_default(const float light_dir[3], uint32_t frame, vpt_reflect_params* out);

VPT_API int vpt_trace_reflect(vpt_world* w, const vpt_camera* cam, const vpt_gbuffer* gb,
                              const vpt_reflect_params* prm, uint16_t* d_dist, uint32_t* d_hit, void* stream);

/* Hit shading of the reflection pass (ReflectionTraceFrag.glsl:582-593,
 * 764-1035; run after the diffuse and shadow denoisers, Pipeline.cpp:
 * 2854-3010).  d_diffuse: the frame's denoised diffuse (vpt_svgf_final),
 * d_shadow: its denoised sun shadow (vpt_shadow_denoise_frame's d_out),
 * both at the G-buffer's size. */
typedef struct vpt_reflect_shading {
    const vpt_svgf_texel* d_diffuse;
    const uint8_t* d_shadow;
    float viewproj[16];         /* u_Projection * u_View of the frame (column-major) */
    float sun_radiance[3];      /* SAMPLED_COLOR_MIXED substitute (the hit's direct term uses 0.6x) */
    float sky_radiance[3];      /* GetAtmosphere substitute for a reflected ray that escapes */
    float albedo[128 * 3];      /* albedo per block id */
} vpt_reflect_shading;

/* sun_radiance = gi->light_color, sky_radiance = gi->sky_radiance, albedo =
 * gi->albedo; the planes and viewproj are the caller's. */
VPT_API int vpt_reflect_shading_from_gi(const vpt_gi_params* gi, const float viewproj[16],
                                        const vpt_svgf_texel* d_diffuse, const uint8_t* d_shadow,
                                        vpt_reflect_shading* out);

/* vpt_trace_reflect plus the hit shading.  Each hit is reprojected to the
 * screen with viewproj (ReprojectReflectionToScreenSpace, :582-593): when
 * its uv is inside the 0.0035 border and the G-buffer there holds a
 * position within 0.095 (squared) and the same normal, the hit takes its
 * ambient from the denoised diffuse at uv (SHToIrradianceA, times
 * pow(AO, 0.75) when that position is within 40 of the camera) and, for the
 * shadow-casting samples, the denoised shadow at uv instead of a shadow
 * ray; other hits take the ambient at the pixel.  A hit adds ambient *
 * albedo + the diffuse directional light (sun_radiance * 0.6, F0 0.04), a
 * miss adds sky_radiance.  d_color: RGBA binary16 per target pixel, the
 * mean over spp clamped to [1e-7, 100] with alpha 1 (0 for sky pixels).
 * d_hit adds bit 17 (the first hit was reprojected); its shadow bit is the
 * last shadow value >= 0.5.  Mirror reflections (rough = 0) trace the ray
 * and its shadow ray once and repeat the sample spp times, as the shader's
 * loop does. */
VPT_API int vpt_trace_reflect_shaded(vpt_world* w, const vpt_camera* cam, const vpt_gbuffer* gb,
                                     const vpt_reflect_params* prm, const vpt_reflect_shading* shading,
                                     uint16_t* d_dist, uint32_t* d_hit, uint16_t* d_color, void* stream);

/* Bytes of one sample's record in the deferred form (spp records per target
 * pixel, (W / div) x (H / div) target pixels). */
#define VPT_REFLECT_RECORD_BYTES 16

/* vpt_trace_reflect_shaded in two calls, so that the rays, which need only
 * the G-buffer, can be traced while the denoisers run.
 * vpt_trace_reflect_deferred traces the reflected rays and the shadow rays
 * of the samples whose hits do not reproject with viewproj (which must equal
 * the shading's viewproj given to vpt_shade_reflect), writes d_dist and
 * d_hit (with the shadow bit 0) and each sample's hit to d_records;
 * vpt_shade_reflect then shades the recorded hits from the denoised planes,
 * writes d_color and sets d_hit's shadow bit (d_hit may be NULL in both).
 * Together they give vpt_trace_reflect_shaded's outputs for the same
 * arguments, bit for bit. */
VPT_API int vpt_trace_reflect_deferred(vpt_world* w, const vpt_camera* cam, const vpt_gbuffer* gb,
                                       const vpt_reflect_params* prm, const float viewproj[16], uint16_t* d_dist,
                                       uint32_t* d_hit, void* d_records, void* stream);
VPT_API int vpt_shade_reflect(vpt_ctx* ctx, const vpt_camera* cam, const vpt_gbuffer* gb, const vpt_reflect_params* prm,
                              const vpt_reflect_shading* shading, const void* d_records, uint32_t* d_hit,
                              uint16_t* d_color, void* stream);

/* DIFFUSE_TRACE (DiffuseRayTraceFrag.glsl:822-935), 1 spp, 1 bounce with a
 * shadow ray at the bounce hit.  d_rgb: 3 floats per pixel (clamped to
 * [0, 8]).  d_svgf_in (optional): SVGF input texels.  Three kernels: the
 * bounce rays (regrouped by direction within each 16x16 block) for at most
 * 24 steps, shading their pixels at once unless the hit faces the sun; the
 * rays still running, resumed densely from compacted queues; the shadow
 * rays of the pixels whose hit faces the sun, also from compacted queues.
 * The queues live in scratch the context keeps for `stream` (32 B per
 * pixel + 4 KiB, grown on demand; outgrown buffers are kept until the
 * context is destroyed, since a captured graph may still use them).  The albedo table in prm is passed with the launch and also
 * becomes the context's "last table" for vpt_unet_assemble. */
VPT_API int vpt_trace_gi(vpt_world* w, const vpt_camera* cam, const vpt_gbuffer* gb, const vpt_gi_params* prm,
                         float* d_rgb, vpt_svgf_texel* d_svgf_in, void* stream);

/* No counterpart in the reference: vpt_trace_gi for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, results and
 * kernels; the queues live in a fixed array of
 * 64 MiB + 4 KiB that each GPU the library runs on holds from context
 * creation until the process ends, so images are limited to 2,097,152
 * pixels (1920 x 1080 fits; larger is VPT_EINVAL).  It calls no hipMalloc
 * or hipFree.  Every call on one GPU shares that array: calls that overlap
 * on one GPU (other streams, other contexts, other threads) race; order
 * them on one stream or wait in between. */
VPT_API int vpt_trace_gi_noalloc(vpt_world* w, const vpt_camera* cam, const vpt_gbuffer* gb,
                                 const vpt_gi_params* prm, float* d_rgb, vpt_svgf_texel* d_svgf_in, void* stream);

/* The three trace passes of one frame, one launch each (INITIAL_TRACE,
 * SHADOW_TRACE, DIFFUSE_TRACE, Pipeline.cpp:1963-2852): the primary pass on
 * `stream` fills gb, then the GI pass runs on `stream` and, from the end of
 * its bounce kernel on, the shadow pass on a stream of the context beside
 * it, and `stream` waits for both.  Outputs
 * equal vpt_trace_primary + vpt_trace_shadow + vpt_trace_gi, or, with
 * d_transversal (optional), vpt_trace_shadow2's two outputs in place of
 * vpt_trace_shadow's; the overlap fills the tail of each pass with the
 * other's work.  The side stream and its fork / join events are the
 * context's: calls from one host thread on any streams are ordered
 * correctly, calls from several host threads on one context are not. */
VPT_API int vpt_trace_frame(vpt_world* w, const vpt_camera* cam, const vpt_gi_params* prm, int primary_iter,
                            int shadow_iter, const vpt_gbuffer* gb, uint8_t* d_mask, uint16_t* d_transversal,
                            float* d_rgb, vpt_svgf_texel* d_svgf_in, void* stream);

/* ---- graph replay of a frame's calls ---------------------------------------
 * No counterpart in the reference (its frame is a sequence of GL draws).
 * A host brackets one frame's calls of this library with vpt_graph_begin /
 * vpt_graph_launch: begin starts capturing an internal stream (ordered
 * after the work already on `stream`) and returns it in *capture_stream;
 * the host enqueues the frame's calls on that stream (and on other streams
 * that wait on events recorded in it and are joined back to it); launch
 * ends the capture, instantiates the graph the first time — later frames
 * update the instantiated graph's kernel arguments in place when the
 * sequence has the same shape, and re-instantiate when it does not — and
 * replays it, with `stream` made to wait for it.  Calls made while
 * capturing enqueue nothing until the launch.  Every call of this library
 * used in a frame is capturable (kernels and device-to-device copies only;
 * vpt_trace_gi may allocate its stream's scratch, which the relaxed capture
 * mode used here allows).  vpt_graph_abort
 * ends a capture without launching it (after an error in the frame). */
typedef struct vpt_graph vpt_graph;
typedef struct vpt_graph_stats {
    uint64_t launches;
    uint64_t instantiations;
    uint64_t updates;
    uint32_t nodes;          /* nodes of the last captured graph */
    uint32_t pad;
} vpt_graph_stats;
VPT_API int vpt_graph_create(vpt_ctx* ctx, vpt_graph** out);
VPT_API int vpt_graph_destroy(vpt_graph* g);
VPT_API int vpt_graph_begin(vpt_graph* g, void* stream, void** capture_stream);
VPT_API int vpt_graph_launch(vpt_graph* g, void* stream);
VPT_API int vpt_graph_abort(vpt_graph* g);
VPT_API int vpt_graph_stats_get(const vpt_graph* g, vpt_graph_stats* st);

/* Ray queries (serving API): n rays, traversal cap max_iter. */
VPT_API int vpt_trace_rays(vpt_world* w, const vpt_ray* d_rays, int n, int max_iter, vpt_hit* d_hits, void* stream);

/* GI ray queries: each ray is a diffuse bounce (cap gi_iter); a bounce that
 * hits a block whose face sees the sun continues as the hit's shadow ray
 * from hit + N*0.045 toward sun_dir (cap shadow_iter), as GetShadowAt at
 * DiffuseRayTraceFrag.glsl:1202-1222 is used by CalculateDiffuse (:535-664).
 * d_hits[i] = the bounce's hit; d_vis[i] = 1 when the sun is blocked at the
 * hit, 0 when it is visible, 255 when no shadow ray was traced (miss, or
 * N.L < 0.001).  Two one-launch kernels (bounce, then shadow). */
VPT_API int vpt_trace_gi_rays(vpt_world* w, const vpt_ray* d_rays, int n, int gi_iter, const float sun_dir[3],
                              int shadow_iter, vpt_hit* d_hits, uint8_t* d_vis, void* stream);

/* Work counters for bandwidth accounting: when d_counter is non-NULL (two
 * uint64 in device memory), every later trace call on this world adds its
 * DF samples to d_counter[0] and its traced rays to d_counter[1].  Pass
 * NULL to disable. */
VPT_API int vpt_world_set_sample_counter(vpt_world* w, unsigned long long* d_counter);

/* ---- batched / persistent tracing ------------------------------------------
 * One persistent launch drains a queue of rays from any number of resident
 * worlds: a lane whose ray has finished refills itself from a wave-private
 * chunk of the queue, so divergent rays do not idle their wavefront.
 * Results are bit-identical to vpt_trace_rays. */
typedef struct vpt_batch vpt_batch;

/* A queued ray: origin, unit direction, meta = max_iter (bits 0..15) |
 * world slot (bits 16..23), dst = index of its vpt_hit in the output. */
typedef struct vpt_qray {
    float origin[3];
    float dir[3];
    uint32_t meta;
    uint32_t dst;
} vpt_qray;

/* Launch shape of the persistent kernels (chosen by vpt_batch_autotune). */
typedef struct vpt_launch_cfg {
    int32_t block;           /* threads per workgroup: 64, 128 or 256 */
    int32_t waves_per_cu;    /* resident wavefronts per CU launched */
    int32_t chunk;           /* rays taken from the queue per atomic */
    int32_t refill;          /* a wavefront refills once this many lanes are idle (1..64) */
    int32_t flags;           /* frame path, bit 0: primary rays by the one-launch tile kernel;
                                served rays, bit 1: one lane per ray instead of the persistent kernel */
} vpt_launch_cfg;

VPT_API int vpt_batch_create(vpt_ctx* ctx, int max_worlds, vpt_batch** out);
VPT_API int vpt_batch_destroy(vpt_batch* b);
/* Registers (or with NULL clears) a resident world in `slot`. */
VPT_API int vpt_batch_set_world(vpt_batch* b, int slot, vpt_world* w);
VPT_API int vpt_batch_get_cfg(const vpt_batch* b, vpt_launch_cfg* cfg);
VPT_API int vpt_batch_set_cfg(vpt_batch* b, const vpt_launch_cfg* cfg);
/* Traces n queued rays in one persistent launch: d_hits[ray.dst] = result
 * (d_hits holds n entries; a ray whose dst is outside [0, n) is traced and
 * its result dropped; a ray naming an empty world slot misses). */
VPT_API int vpt_batch_trace(vpt_batch* b, const vpt_qray* d_rays, int n, vpt_hit* d_hits, void* stream);
/* Times the candidate launch shapes on this ray set and keeps the fastest;
 * *best_ms receives its time (may be NULL). */
VPT_API int vpt_batch_autotune(vpt_batch* b, const vpt_qray* d_rays, int n, vpt_hit* d_hits, float* best_ms,
                               void* stream);

/* The per-frame ray passes of one view through the persistent kernels:
 * primary rays into the G-buffer, then one fused launch for the sun-shadow
 * rays and the GI rays (each GI hit continuing as its shadow ray in the
 * same lane), then shading.  Outputs are bit-identical to vpt_trace_primary
 * + vpt_trace_shadow + vpt_trace_gi.  The pass chain can be captured into a
 * HIP graph (vpt_frame_graph) and replayed per frame. */
typedef struct vpt_frame vpt_frame;

VPT_API int vpt_frame_create(vpt_ctx* ctx, int width, int height, vpt_frame** out);
VPT_API int vpt_frame_destroy(vpt_frame* f);
VPT_API int vpt_frame_set_cfg(vpt_frame* f, const vpt_launch_cfg* cfg);
VPT_API int vpt_frame_get_cfg(const vpt_frame* f, vpt_launch_cfg* cfg);
/* Runs the passes for camera `cam` (G-buffer `gb`, shadow mask, GI outputs
 * as in vpt_trace_gi).  use_graph != 0 replays a captured HIP graph
 * (captured on first use, re-captured when buffers change). */
VPT_API int vpt_frame_trace(vpt_frame* f, vpt_world* w, const vpt_camera* cam, const vpt_gi_params* prm,
                            int primary_iter, int shadow_iter, const vpt_gbuffer* gb, uint8_t* d_mask, float* d_rgb,
                            vpt_svgf_texel* d_svgf_in, int use_graph, void* stream);
/* Times candidate launch shapes on one frame and keeps the fastest. */
VPT_API int vpt_frame_autotune(vpt_frame* f, vpt_world* w, const vpt_camera* cam, const vpt_gi_params* prm,
                               int primary_iter, int shadow_iter, const vpt_gbuffer* gb, uint8_t* d_mask,
                               float* d_rgb, vpt_svgf_texel* d_svgf_in, float* best_ms, void* stream);

/* ---- serving ---------------------------------------------------------------
 * A request loop over worlds kept resident in HBM.  Clients submit ray-query
 * requests (primary, shadow or GI rays already in device memory) and frame
 * requests (the primary + shadow + GI passes of one camera) of any size; a
 * worker thread admits them while their result memory fits in the server's
 * arena (first come, first admitted), runs them, and records arrival,
 * admission, start and completion times.
 *
 * mode VPT_SERVE_PLAIN runs one request at a time with the one-launch
 * kernels (vpt_trace_rays / vpt_trace_gi_rays / vpt_trace_primary +
 * vpt_trace_shadow + vpt_trace_gi) and waits for each.  VPT_SERVE_BATCHED
 * does continuous batching: every launch traces the rays of all admitted
 * requests together in one kernel (a request larger than a launch is split
 * over several) — one lane per ray by default, or the persistent refill
 * kernel when the launch shape's bit 1 is clear (vpt_server_set_launch_cfg)
 * — GI bounces continue as their shadow rays in the same lane, up to max_inflight launches are queued on the device at once,
 * and requests that arrive meanwhile join the next launch.  Results are
 * bit-identical in both modes and do not depend on what else is in flight.
 *
 * Results live in server memory until vpt_server_release.  A request's
 * d_rays must hold its rays when it is submitted (the work producing them
 * has completed) and stay valid until the request completes.  The server
 * runs on its own non-blocking stream.  Submit / wait / release may be
 * called from any threads. */
typedef struct vpt_server vpt_server;

enum { VPT_SERVE_PLAIN = 0, VPT_SERVE_BATCHED = 1 };
enum { VPT_REQ_PRIMARY = 0, VPT_REQ_SHADOW = 1, VPT_REQ_GI = 2, VPT_REQ_FRAME = 3 };

typedef struct vpt_server_cfg {
    int32_t mode;             /* VPT_SERVE_PLAIN or VPT_SERVE_BATCHED */
    int32_t max_worlds;       /* world slots, 1..256 */
    uint64_t arena_bytes;     /* request result memory (admission budget) */
    uint64_t world_budget;    /* bytes of resident worlds (and frame scratch) allowed */
    uint32_t max_launch_rays; /* rays per persistent launch (batched) */
    int32_t max_inflight;     /* launches queued on the device at once, 1..4 (batched) */
} vpt_server_cfg;

typedef struct vpt_request {
    int32_t world;            /* slot from vpt_server_add_world */
    int32_t kind;             /* VPT_REQ_* */
    uint32_t n;               /* rays (ray kinds) */
    int32_t max_iter;         /* traversal cap (ray kinds; the bounce cap for GI) */
    const vpt_ray* d_rays;    /* ray kinds: n rays in device memory */
    float sun_dir[3];         /* GI: continuation shadow ray direction */
    int32_t shadow_iter;      /* GI: its cap */
    const vpt_camera* cam;    /* frame: camera (copied at submit) */
    const vpt_gi_params* gi;  /* frame: GI constants (copied at submit) */
    int32_t primary_iter;     /* frame */
    int32_t frame_shadow_iter;/* frame: sun-shadow cap */
    int64_t arrival_ns;       /* arrival time on vpt_now_ns's clock; 0 = at submit */
} vpt_request;

typedef struct vpt_request_result {
    int32_t status;           /* VPT_OK or the error the request ended with */
    int32_t kind;
    uint32_t n;               /* rays, or pixels of a frame */
    vpt_hit* d_hits;          /* ray kinds: n results */
    uint8_t* d_vis;           /* GI: sun blocked (1) / visible (0) / no shadow ray (255) */
    vpt_gbuffer gbuffer;      /* frame: G-buffer */
    uint8_t* d_mask;          /* frame: sun-shadow mask */
    float* d_rgb;             /* frame: GI radiance, 3 floats per pixel */
    int64_t t_arrival_ns, t_admit_ns, t_start_ns, t_done_ns;
} vpt_request_result;

typedef struct vpt_server_stats {
    uint64_t submitted, completed, rays, launches, frames;
    uint64_t arena_used, arena_peak, arena_bytes, world_bytes;
    uint32_t pending, active;
} vpt_server_stats;

/* Steady clock used for request times (nanoseconds). */
VPT_API int64_t vpt_now_ns(void);
VPT_API int vpt_server_create(vpt_ctx* ctx, const vpt_server_cfg* cfg, vpt_server** out);
/* Waits for all submitted work, then frees the server (not the worlds). */
VPT_API int vpt_server_destroy(vpt_server* s);
/* Makes `w` resident in the first free slot (*slot); VPT_ENOMEM when its
 * bytes exceed what is left of world_budget. */
VPT_API int vpt_server_add_world(vpt_server* s, vpt_world* w, int* slot);
/* VPT_EBUSY while a pending or running request names the slot. */
VPT_API int vpt_server_remove_world(vpt_server* s, int slot);
/* Persistent-kernel launch shape (batched mode). */
VPT_API int vpt_server_set_launch_cfg(vpt_server* s, const vpt_launch_cfg* cfg);
/* Queues a request; *ticket identifies it.  VPT_ENOMEM when its result
 * memory exceeds the whole arena (it could never be admitted). */
VPT_API int vpt_server_submit(vpt_server* s, const vpt_request* req, uint64_t* ticket);
/* Blocks until the request completes (timeout_ms < 0: no limit) and fills
 * *res; VPT_ETIMEOUT if it has not completed in time. */
VPT_API int vpt_server_wait(vpt_server* s, uint64_t ticket, int timeout_ms, vpt_request_result* res);
/* Frees a completed request's result memory (the ticket becomes invalid). */
VPT_API int vpt_server_release(vpt_server* s, uint64_t ticket);
VPT_API int vpt_server_stats_get(vpt_server* s, vpt_server_stats* st);

/* ---- SVGF ---------------------------------------------------------------
 * Owns the history and ping-pong planes of one W x H view. */
typedef struct vpt_svgf_params {
    float color_phi_bias;   /* u_ColorPhiBias, 3.325 (Pipeline.cpp:85) */
    float resolution_scale; /* u_ResolutionScale, 0.25 (Pipeline.cpp:78) */
    int32_t pre_spatial;    /* PreTemporalSpatialPass (Pipeline.cpp:239) */
    int32_t steps[5];       /* a-trous steps, 16 8 4 2 1 (Pipeline.cpp:2478-2497) */
} vpt_svgf_params;

VPT_API int vpt_svgf_create(vpt_ctx* ctx, int width, int height, const vpt_svgf_params* prm, vpt_svgf** out);
VPT_API int vpt_svgf_destroy(vpt_svgf* s);
/* Clears the temporal history (first frame after a cut). */
VPT_API int vpt_svgf_reset(vpt_svgf* s, void* stream);

/* One frame of the chain.  cur/prev: this and the previous frame's
 * G-buffer; cam: this frame's camera; prev_viewproj: previous projection *
 * previous view; d_in: this frame's raw GI texels; time: the value used for
 * u_Time (a-trous jitter).  d_rgb_out: decoded irradiance, 3 floats per
 * pixel (SHToIrridiance with the G-buffer normal, ColorPassFrag.glsl:356).
 * d_texel_out (optional): the final SH/CoCg/variance/AO texels. */
VPT_API int vpt_svgf_frame(vpt_svgf* s, const vpt_camera* cam, const float prev_viewproj[16],
                           const vpt_gbuffer* cur, const vpt_gbuffer* prev, const vpt_svgf_texel* d_in, float time,
                           float* d_rgb_out, vpt_svgf_texel* d_texel_out, void* stream);

/* The final a-trous plane of the last frame (what d_texel_out receives),
 * without a copy: valid for work enqueued on the frame's stream before the
 * next vpt_svgf_frame or vpt_svgf_destroy. */
VPT_API int vpt_svgf_final(const vpt_svgf* s, const vpt_svgf_texel** out);

/* Debug access to the planes after the last frame: which = 0 pre-spatial,
 * 1 temporal (history), 2 variance, 3 final a-trous output.  Copies
 * width*height texels to device memory d_dst. */
VPT_API int vpt_svgf_read_plane(vpt_svgf* s, int which, vpt_svgf_texel* d_dst, void* stream);
/* The temporal utility plane of the last frame (o_Utility of
 * TemporalFilter.glsl: accumulated frames, moment, luma as binary16, plus
 * one unused half): 4 halves per pixel. */
VPT_API int vpt_svgf_read_utility(vpt_svgf* s, uint16_t* d_dst, void* stream);

/* ---- sun-shadow denoiser -----------------------------------------------------
 * ShadowTemporalFilter + ShadowFilter (ShadowTemporalFilter.glsl:98-267,
 * ShadowFilter.glsl:69-165; wiring Pipeline.cpp:2854-2948, SoftShadows and
 * DenoiseSunShadows on).  Holds the temporal history (colour unorm8, frame
 * count binary16, ping-pong) for one width x height. */
VPT_API int vpt_shadow_denoise_create(vpt_ctx* ctx, int width, int height, vpt_shadow_denoise** out);
VPT_API int vpt_shadow_denoise_destroy(vpt_shadow_denoise* d);
/* Clears the history (colour 0, frame count 0). */
VPT_API int vpt_shadow_denoise_reset(vpt_shadow_denoise* d, void* stream);
/* One frame.  cur/prev: this and the previous frame's G-buffers (t, face;
 * prev: t); d_mask / d_transversal: vpt_trace_shadow2's outputs for cur;
 * prev_viewproj: the previous frame's projection * view; filter_scale:
 * u_ShadowFilterScale (1 by default, Pipeline.cpp:132).  d_out: the
 * filtered shadow, one unorm8 byte per pixel (ShadowFiltered). */
VPT_API int vpt_shadow_denoise_frame(vpt_shadow_denoise* d, const vpt_camera* cam, const float prev_viewproj[16],
                                     const vpt_gbuffer* cur, const vpt_gbuffer* prev, const uint8_t* d_mask,
                                     const uint16_t* d_transversal, float filter_scale, uint8_t* d_out, void* stream);
/* The last frame's temporal planes: which = 0 colour (1 byte per pixel),
 * 1 frame count (binary16 bits). */
VPT_API int vpt_shadow_denoise_read(vpt_shadow_denoise* d, int which, void* d_dst, void* stream);

/* ---- colour composite -------------------------------------------------------
 * The colour pass's shading of the G-buffer (ColorPassFrag.glsl:773-946,
 * u_UseDFG) for flat dielectric materials: per pixel, the denoised diffuse
 * decoded with the face normal (SHToIrridiance) times VXAO-faded AO, the
 * Cook-Torrance sun and moon light mixed by SunVisibility and shadowed by
 * the denoised shadow, and the reflection colour weighted by specular AO,
 * horizon fade and the Karis environment BRDF.  Stand-ins for what the
 * reference samples from textures outside the hot path: albedo per block
 * id (the block albedo atlas), the face normal (normal maps), one roughness
 * (PBR atlas, metalness 0), constant sun / moon / sky-pixel radiances (the
 * atmosphere cubemap), full-resolution diffuse and shadow (the spatial
 * upscale passes at scale 1), LINEAR sampling of the reflection colour for
 * its upscale. */
typedef struct vpt_composite_params {
    float sun_dir[3];             /* u_SunDirection, normalised */
    float moon_dir[3];            /* u_MoonDirection: normalize(-sun.x, -sun.y, sun.z), Pipeline.cpp:1663 */
    float sun_radiance[3];        /* SAMPLED_SUN_COLOR substitute */
    float moon_radiance[3];       /* SAMPLED_MOON_COLOR substitute */
    float sky_pixel_radiance[3];  /* output of primary-miss pixels */
    float roughness;              /* RoughnessAt of the flat material */
    float albedo[128 * 3];        /* albedo per block id */
} vpt_composite_params;

/* sun_dir, albedo and sky_pixel_radiance from gi, sun_radiance =
 * gi->light_color, moon_radiance = (0.1, 0.1, 1.0) (ATMOSPHERE_MOON_COLOR,
 * ColorPassFrag.glsl:196), roughness 0.5 (the flat material's RoughnessAt
 * before RoughnessBias). */
VPT_API int vpt_composite_params_from_gi(const vpt_gi_params* gi, vpt_composite_params* out);

/* d_rgb: 3 floats per pixel (W x H).  d_diffuse / d_shadow: W x H (as for
 * vpt_reflect_shading); d_reflect: vpt_trace_reflect_shaded's d_color at
 * (W / reflect_div) x (H / reflect_div).  Sky pixels (t <= 0) take
 * sky_pixel_radiance. */
VPT_API int vpt_composite(vpt_ctx* ctx, const vpt_camera* cam, const vpt_gbuffer* gb, const vpt_composite_params* prm,
                          const vpt_svgf_texel* d_diffuse, const uint8_t* d_shadow, const uint16_t* d_reflect,
                          int reflect_div, float* d_rgb, void* stream);

/* ---- learned denoiser -------------------------------------------------------
 * A 3-level U-Net (32 / 64 / 128 channels, 3x3 convolutions, 2x2 average
 * pooling, nearest upsampling with skip concatenation, 1x1 output head)
 * that maps one frame's 1-spp GI irradiance plus albedo, normal and depth
 * to the converged irradiance.  bf16 weights and activations, fp32
 * accumulation on the MFMA units; trained by its own backward kernels and
 * Adam.  No counterpart in the reference (its denoiser is the SVGF chain
 * above, which stays the fallback).
 *
 * Network input: NHWC bf16 with 32 channels per pixel (vpt_unet_assemble):
 * log1p(irradiance) 3, albedo 3, normal 3, log1p(t)/8 1, zeros 22.
 * Output / training target: NHWC fp32, 3 channels, log1p(irradiance)
 * (mode 0) or irradiance (mode 1).  n x h x w with h, w multiples of 4. */
typedef struct vpt_unet vpt_unet;

VPT_API int vpt_unet_create(vpt_ctx* ctx, int max_n, int max_h, int max_w, vpt_unet** out);
VPT_API int vpt_unet_destroy(vpt_unet* u);
VPT_API int vpt_unet_param_count(const vpt_unet* u, size_t* n);
/* He-normal initialisation from `seed`; resets the optimiser. */
VPT_API int vpt_unet_init(vpt_unet* u, uint32_t seed);
/* Parameters as one fp32 vector: per layer [9][cout][cin] weights then
 * [cout] biases, then the head [3][32] and [3]; set resets the optimiser. */
VPT_API int vpt_unet_set_params(vpt_unet* u, const float* host, size_t n);
VPT_API int vpt_unet_get_params(vpt_unet* u, float* host, size_t n);
/* Gradients of the last vpt_unet_backward, same layout. */
VPT_API int vpt_unet_get_grads(vpt_unet* u, float* host, size_t n);
VPT_API int vpt_unet_forward(vpt_unet* u, const uint16_t* d_x, int n, int h, int w, int mode, float* d_out,
                             void* stream);
/* Forward, L1 loss against d_target (log space) and all gradients;
 * *loss (optional) receives the mean absolute error (synchronises). */
VPT_API int vpt_unet_backward(vpt_unet* u, const uint16_t* d_x, const float* d_target, int n, int h, int w,
                              float* loss, void* stream);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) with the last gradients. */
VPT_API int vpt_unet_adam_step(vpt_unet* u, float lr, void* stream);
/* Network input of one frame from the G-buffer and vpt_trace_gi's texels
 * (albedo from the table of the context's last GI pass). */
VPT_API int vpt_unet_assemble(vpt_unet* u, const vpt_gbuffer* gb, const vpt_svgf_texel* d_tex, uint16_t* d_x,
                              void* stream);
/* assemble + forward: denoised irradiance, 3 floats per pixel. */
VPT_API int vpt_unet_denoise(vpt_unet* u, const vpt_gbuffer* gb, const vpt_svgf_texel* d_tex, float* d_rgb,
                             void* stream);
/* Debug / tests: activation `index` of the last forward pass (0 input,
 * 1 e0a, 2 e0b, 3 p0, 4 e1a, 5 e1b, 6 p1, 7 b0, 8 b1, 9 d1a, 10 d1b,
 * 11 d0a, 12 d0b), NHWC bf16 at its level's resolution. */
VPT_API int vpt_unet_read_activation(vpt_unet* u, int index, int n, int h, int w, uint16_t* host);
/* Reference accumulation: d_acc[i][0..5] (+)= SH-Y and CoCg of texel i. */
VPT_API int vpt_texel_accumulate(const vpt_svgf_texel* d_tex, int npx, float* d_acc, int reset, void* stream);
/* SHToIrridiance (ColorPassFrag.glsl:356-368) with the G-buffer face
 * normal, of d_tex or (d_tex NULL) of d_acc * acc_scale. */
VPT_API int vpt_irradiance_decode(const vpt_gbuffer* gb, const vpt_svgf_texel* d_tex, const float* d_acc,
                                  float acc_scale, float* d_rgb, void* stream);

/* ---- host utilities ------------------------------------------------------ */

/* Synthetic terrain: seeded 2D value-noise heightfield in [h_min, h_max],
 * grass (1) on top, 4 dirt (2), stone (3) below (WorldGenerator.cpp:28-47
 * layering), then `random_permille` / 1000 of the solid voxels replaced by
 * ids 1..100 (seed_random).  Host memory, linear layout. */
VPT_API int vpt_generate_terrain(int nx, int ny, int nz, uint32_t seed, uint32_t seed_random, int h_min, int h_max,
                                 int random_permille, uint8_t* host_grid);

/* glm::lookAt / glm::perspective camera (FpsCamera.cpp:23-24,66-72,150-163):
 * front = (cos p cos y, sin p, cos p sin y), up = +y. view/proj/inverses
 * are written when non-NULL. */
VPT_API int vpt_camera_make(const float pos[3], float yaw_deg, float pitch_deg, float fov_deg, int width, int height,
                            float znear, float zfar, vpt_camera* cam, float view[16], float proj[16]);

/* Frame constants of the GI pass for a sun tick (Pipeline.cpp:1652-1671,
 * 1828; DiffuseRayTraceFrag.glsl:181-186, 829-840): sun direction, light
 * colour, sky radiance from a constant sky colour; albedo table seeded. */
VPT_API int vpt_gi_params_default(float sun_tick, uint32_t frame, uint32_t albedo_seed, vpt_gi_params* prm);

VPT_API int vpt_svgf_params_default(vpt_svgf_params* prm);

/* World::RaycastDetect (World.cpp:496-546): CPU block picking over a host
 * grid, 48 epsilon steps of 0.001 past each cell boundary; cells with any
 * coordinate <= 0 count as outside.  out = (x, y, z, id), or all -1 when
 * nothing is hit (the reference falls off the end of the function there). */
VPT_API int vpt_raycast_detect(const uint8_t* host_grid, int nx, int ny, int nz, const float pos[3],
                               const float dir[3], int32_t out[4]);

#ifdef __cplusplus
}
#endif

#endif /* VPT_H */
