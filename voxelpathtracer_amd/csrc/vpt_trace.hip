// Ray kernels, one launch per pass: primary (G-buffer), sun shadow, diffuse
// GI, ray queries.
//
// One lane per ray.  A 256-lane workgroup covers a 16x16 pixel tile made of
// four 8x8 wavefront squares (square footprints keep a wavefront's rays on
// neighbouring bricks), and tiles are dealt to XCDs in contiguous screen
// bands (xcd_remap) so each XCD's L2 holds one region of the world.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>

#include "vpt_device.h"
#include "vpt_gi.h"
#include "vpt_internal.h"

using vpt::CamView;
using vpt::GiConst;
using vpt::set_error;
using vpt::Vec3;
using vpt::WorldView;

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return set_error(VPT_EDEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));      \
    } while (0)

namespace vpt {

// Scratch of one GI pass: the 16 queue counters (continuations, then
// shadow jobs, one per XCD), each on its own 256-byte line, then the
// 16-byte continuation records and the 16-byte shadow jobs, one of each per
// pixel.
constexpr size_t kGiCounterStride = 64;   // words
constexpr size_t kGiCountersBytes = 16 * kGiCounterStride * sizeof(uint32_t);
inline size_t gi_scratch_bytes(size_t npx) { return kGiCountersBytes + 2 * 16 * npx; }

WorldView make_view(const vpt_world* w)
{
    WorldView v;
    v.grid = w->d_grid;
    v.df = w->d_df;
    v.nx = w->nx;
    v.ny = w->ny;
    v.nz = w->nz;
    v.row_stride = (uint32_t)w->nbx * kBrickBytes;
    v.slab_stride = v.row_stride * (uint32_t)w->nby;
    v.xmax = (float)(w->nx - 1);
    v.ymax = (float)(w->ny - 1);
    v.zmax = (float)(w->nz - 1);
    return v;
}

CamView make_cam(const vpt_camera* c)
{
    CamView v;
    for (int i = 0; i < 16; i++) {
        v.iv[i] = c->inv_view[i];
        v.ip[i] = c->inv_proj[i];
    }
    v.W = c->width;
    v.H = c->height;
    return v;
}

GiConst make_gi(const vpt_gi_params* prm)
{
    GiConst c;
    for (int k = 0; k < 3; k++) {
        c.sun[k] = prm->sun_dir[k];
        c.light[k] = prm->light_color[k];
        c.sky[k] = prm->sky_radiance[k];
        c.sky_pixel[k] = prm->sky_pixel_radiance[k];
    }
    c.frame = prm->frame;
    c.gi_iter = prm->gi_iter;
    c.shadow_iter = prm->shadow_iter;
    return c;
}

AlbedoTable take_albedo(vpt_ctx* ctx, const vpt_gi_params* prm)
{
    AlbedoTable t;
    static_assert(sizeof(t.v) == sizeof(prm->albedo), "albedo table size");
    memcpy(t.v, prm->albedo, sizeof(t.v));
    std::lock_guard<std::mutex> lk(ctx->mu);
    memcpy(ctx->albedo_host, t.v, sizeof(t.v));
    ctx->albedo_valid = true;
    return t;
}

static void gi_retire(vpt_ctx* ctx, GiScratch* sc)
{
    if (sc->d_buf) ctx->gi_retired.push_back(sc->d_buf);
    sc->d_buf = nullptr;
    sc->cap = 0;
}

// The GI scratch of `stream`, at least `npx` pixels, with `stream` made to
// wait for the entry's previous use.  Called with ctx->mu held.  While
// `stream` is being captured into a graph the ordering comes from the
// stream the graph is replayed on instead of the event (a capture cannot
// wait for an event recorded outside it), and growing allocates (a relaxed
// capture allows it) without freeing.
static int gi_scratch_for(vpt_ctx* ctx, hipStream_t stream, size_t npx, bool capturing, GiScratch** out)
{
    GiScratch* sc = nullptr;
    for (GiScratch& e : ctx->gi_scratch)
        if (e.stream == (void*)stream) sc = &e;
    if (!sc) {
        if (ctx->gi_scratch.size() < kMaxGiScratch) {
            ctx->gi_scratch.emplace_back();
            sc = &ctx->gi_scratch.back();
        } else {
            sc = &ctx->gi_scratch[0];
            for (GiScratch& e : ctx->gi_scratch)
                if (e.last_use < sc->last_use) sc = &e;
            gi_retire(ctx, sc);   // the evicted stream's graphs may still use them
        }
        sc->stream = (void*)stream;
    }
    if (!sc->done) {
        hipEvent_t ev = nullptr;
        HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        sc->done = ev;
    }
    if (sc->cap < npx) {
        gi_retire(ctx, sc);
        hipError_t e = hipMalloc(&sc->d_buf, gi_scratch_bytes(npx));
        if (e != hipSuccess) {
            gi_retire(ctx, sc);
            return set_error(capturing ? VPT_ESTATE : VPT_ENOMEM,
                             "vpt_trace_gi: GI scratch of %zu pixels: %s%s", npx, hipGetErrorString(e),
                             capturing ? " (this capture mode allows no allocation: run the pass once outside "
                                         "capture first)"
                                       : "");
        }
        sc->cap = npx;
    }
    if (!capturing) HIP_TRY(hipStreamWaitEvent(stream, (hipEvent_t)sc->done, 0));
    sc->last_use = ++ctx->gi_clock;
    *out = sc;
    return VPT_OK;
}

int check_cam(const vpt_camera* cam, const char* fn)
{
    if (!cam) return set_error(VPT_EINVAL, "%s: null camera", fn);
    if (cam->width <= 0 || cam->height <= 0 || cam->width > 16384 || cam->height > 16384)
        return set_error(VPT_EINVAL, "%s: bad image size %dx%d", fn, cam->width, cam->height);
    return VPT_OK;
}

int check_gbuffer(const vpt_gbuffer* gb, const vpt_camera* cam, const char* fn, bool need_block)
{
    if (!gb || !gb->d_t || !gb->d_face || (need_block && !gb->d_block))
        return set_error(VPT_EINVAL, "%s: G-buffer planes missing", fn);
    if (gb->width != cam->width || gb->height != cam->height)
        return set_error(VPT_EINVAL, "%s: G-buffer %dx%d does not match camera %dx%d", fn, gb->width, gb->height,
                         cam->width, cam->height);
    return VPT_OK;
}

int check_iter(int max_iter, const char* fn)
{
    if (max_iter < 0 || max_iter > 65535) return set_error(VPT_EINVAL, "%s: max_iter %d out of range", fn, max_iter);
    return VPT_OK;
}

}  // namespace vpt

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

// Pixel of this lane in its 16x16 tile (four 8x8 wavefront squares).
__device__ inline void tile_pixel(int W, int H, int& px, int& py)
{
    const uint32_t tiles_x = (uint32_t)(W + 15) / 16u;
    const uint32_t tiles_y = (uint32_t)(H + 15) / 16u;
    const uint32_t tile = vpt::xcd_remap(blockIdx.x, tiles_x * tiles_y);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    px = (int)(tile % tiles_x) * 16 + (wave & 1) * 8 + (lane & 7);
    py = (int)(tile / tiles_x) * 16 + (wave >> 1) * 8 + (lane >> 3);
}

// Adds this lane's DF samples to counter[0] and traced rays to counter[1]
// with one atomic each per wavefront.  Every lane of the wavefront must
// reach this call.
__device__ inline void count_samples(unsigned long long* counter, int samples, int rays)
{
    if (!counter) return;
    unsigned long long v = (unsigned long long)samples;
    unsigned long long r = (unsigned long long)rays;
    for (int off = 32; off > 0; off >>= 1) {
        v += __shfl_xor(v, off);
        r += __shfl_xor(r, off);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(counter, v);
        atomicAdd(counter + 1, r);
    }
}

// INITIAL_TRACE main (InitialRayTraceFrag.glsl:417-468).
__global__ void __launch_bounds__(256) primary_kernel(WorldView w, CamView cam, int max_iter, float* __restrict__ out_t,
                                                      uint8_t* __restrict__ out_face, uint8_t* __restrict__ out_block,
                                                      uint32_t* __restrict__ out_voxel,
                                                      unsigned long long* __restrict__ counter,
                                                      uint32_t* __restrict__ gi_counts)
{
    // vpt_trace_frame: the GI pass that follows on this stream finds its
    // queue counters cleared (saves it a launch)
    if (gi_counts && blockIdx.x == 0 && threadIdx.x < 16) gi_counts[threadIdx.x * vpt::kGiCounterStride] = 0u;
    int px, py;
    tile_pixel(cam.W, cam.H, px, py);
    int samples = 0, rays = 0;
    if (px < cam.W && py < cam.H) {
        Vec3 dir = vpt::primary_dir(cam, px, py);
        Vec3 org = vpt::mk3(cam.iv[12], cam.iv[13], cam.iv[14]);
        vpt::Hit h = vpt::traverse(w, org, dir, max_iter);
        bool intersect = h.t > 0.0f && h.block > 0;
        size_t i = (size_t)py * cam.W + px;
        out_t[i] = h.t;
        out_face[i] = (uint8_t)(intersect ? h.face : 255);
        out_block[i] = (uint8_t)(intersect ? h.block : 0);
        if (out_voxel) out_voxel[i] = intersect ? h.voxel : 0xffffffffu;
        samples = h.samples;
        rays = 1;
    }
    count_samples(counter, samples, rays);
}

// SHADOW_TRACE main (ShadowRayTraceFrag.glsl:414-513), no cone jitter, no
// alpha test.
__global__ void __launch_bounds__(256) shadow_kernel(WorldView w, CamView cam, const float* __restrict__ gb_t,
                                                     const uint8_t* __restrict__ gb_face, float lx, float ly, float lz,
                                                     int max_iter, uint8_t* __restrict__ mask,
                                                     unsigned long long* __restrict__ counter)
{
    int px, py;
    tile_pixel(cam.W, cam.H, px, py);
    int samples = 0, rays = 0;
    if (px < cam.W && py < cam.H) {
        size_t i = (size_t)py * cam.W + px;
        Vec3 o;
        uint8_t m;
        if (vpt::shadow_setup(w, cam, px, py, gb_t[i], gb_face[i], vpt::mk3(lx, ly, lz), o, m)) {
            vpt::Hit h = vpt::traverse(w, o, vpt::mk3(lx, ly, lz), max_iter);
            m = h.t > 0.0f ? 1 : 0;
            samples = h.samples;
            rays = 1;
        }
        mask[i] = m;
    }
    count_samples(counter, samples, rays);
}

// SHADOW_TRACE main with both outputs (ShadowRayTraceFrag.glsl:7-8, 414-513):
// the mask as shadow_kernel, and o_IntersectionTransversal as binary16:
// clamp(T / 100, 1e-5, 196), 4.25 / 100 when the ray misses, 1 / 100 when
// N.L <= 0.01, 64 for sky.  The ray is traced also when its origin cell is
// solid (the mask is 1 there either way).
__global__ void __launch_bounds__(256) shadow2_kernel(WorldView w, CamView cam, const float* __restrict__ gb_t,
                                                      const uint8_t* __restrict__ gb_face, float lx, float ly,
                                                      float lz, int max_iter, uint8_t* __restrict__ mask,
                                                      uint16_t* __restrict__ transversal,
                                                      unsigned long long* __restrict__ counter)
{
    int px, py;
    tile_pixel(cam.W, cam.H, px, py);
    int samples = 0, rays = 0;
    if (px < cam.W && py < cam.H) {
        const size_t i = (size_t)py * cam.W + px;
        const Vec3 L = vpt::mk3(lx, ly, lz);
        const float dist = gb_t[i];
        uint8_t m;
        float tv;
        if (dist < 0.0f) {
            m = 0;
            tv = 64.0f;
        } else {
            const Vec3 P = vpt::gbuffer_position(cam, px, py, dist);
            const Vec3 N = vpt::face_normal(gb_face[i], 1.0f);
            const float ndotl = vpt::dot3(N, L);
            if (ndotl <= 0.01f) {
                m = 1;
                tv = 1.0f / 100.0f;
            } else {
                const Vec3 o = vpt::mk3(P.x + N.x * 0.06f, P.y + N.y * 0.06f, P.z + N.z * 0.06f);
                float T = -1.0f;
                const int block_at = vpt::voxel_at(w, o);
                if (dist > 0.0f) {
                    const vpt::Hit h = vpt::traverse(w, o, L, max_iter);
                    T = h.t;
                    samples = h.samples;
                    rays = 1;
                }
                m = (T > 0.0f || block_at > 0) ? 1 : 0;
                tv = vpt::fclamp(T / 100.0f, 0.00001f, 196.0f);
                if (T < 0.0f) tv = 4.25f / 100.0f;
            }
        }
        mask[i] = m;
        if (transversal) transversal[i] = (uint16_t)vpt::hw_f2h(tv);
    }
    count_samples(counter, samples, rays);
}

// DIFFUSE_TRACE main + CalculateDiffuse (DiffuseRayTraceFrag.glsl:535-664,
// 822-935), 1 spp, 1 bounce; Lambert sun term, PCG randoms, per-id albedo,
// constant sky.  The pass is split so that a wavefront's lanes stay busy:
// the bounce rays' lengths vary widely (a sixth hit the terrain in two
// steps, a twentieth run to the 48-step cap), and the sun-shadow rays at
// the bounce hits (up to 128 steps) are cast by only a quarter of the
// pixels.
//   gi_bounce_kernel   one lane per bounce ray: the block's 16x16 pixels
//                      set up their rays, which are regrouped onto lanes by
//                      the ray's y direction (16 bins, through LDS), so that
//                      a wavefront traces rays of similar length (rays
//                      leaving upward escape in a few large steps, grazing
//                      ones walk the terrain).  Each ray runs at most
//                      kGiPhase1 steps here; a ray still running then is
//                      saved as a continuation.  A finished ray's pixel is
//                      shaded at once, or, when its hit faces the sun,
//                      becomes a shadow job (sky pixels are shaded by their
//                      own lanes);
//   gi_cont_kernel     one lane per continuation: the pixel's ray again
//                      (from the G-buffer), resumed from the saved state to
//                      its end, then shaded or made a job;
//   gi_shadow_kernel   one lane per job: the pixel's ray again, the shadow
//                      ray, then the pixel's shading.
// Continuations and jobs are appended with one atomic per wavefront to
// queues kept per XCD (GiQueues), whose counters sit on separate cache
// lines: atomics from every wavefront of the chip on one line serialise,
// and cost the bounce kernel more than its tracing.  The next kernel takes
// the queues' records in order as one dense list, so every XCD gets an
// equal share however the hits fall on the screen.
// Every pixel goes through the same functions as a single-kernel pass, and
// a resumed traversal continues from its exact state, so the outputs are
// identical; only the lane that traces and shades a pixel changes.
constexpr int kGiBins = 16;
constexpr int kGiPhase1 = 24;   // bounce steps before a ray is continued
using vpt::kGiCounterStride;

// A pixel waiting for its shadow ray: index, bounce hit distance, face |
// block << 8.
struct GiJob {
    uint32_t pix;
    float T;
    uint32_t fb;
    uint32_t pad;
};

// A bounce ray after kGiPhase1 steps: its position, and the pixel index |
// MinIdx << 28 | Intersection << 30 (pixel indices are below 2^28).  The
// rest of the traversal state is the ray itself, recomputed from the
// pixel, and the sample count, kGiPhase1.
struct GiCont {
    float ox, oy, oz;
    uint32_t meta;
};

// The per-XCD queues of a W x H pass.  The hardware deals workgroups to the
// 8 XCDs round-robin, so bounce block b runs on XCD b % 8; queue k holds
// what those blocks produce, in slots [base[k], base[k] + cap[k]) of the
// continuation and of the job array, and cap[k] is the number of pixels of
// their tiles, so a queue cannot overflow.
struct GiQueues {
    uint32_t base[8];
    uint32_t cap[8];
};

// a[q] for the block-uniform q, by selects: indexing the by-value kernel
// argument directly would copy it to scratch memory.
__device__ inline uint32_t pick8(const uint32_t (&a)[8], uint32_t q)
{
    uint32_t v = a[0];
#pragma unroll
    for (uint32_t k = 1; k < 8; k++) v = q == k ? a[k] : v;
    return v;
}

// Slot of the k-th record of the queues taken in order (queue q holds
// min(counts[q], cap[q]) records from base[q]), and its queue; false past
// the last one.  The counts are the same for every lane (scalar loads).
__device__ inline bool queue_slot(uint32_t k, const uint32_t* __restrict__ counts, const GiQueues& qs, uint32_t& slot,
                                  uint32_t& queue)
{
    uint32_t start = 0;
    bool found = false;
#pragma unroll
    for (uint32_t q = 0; q < 8; q++) {
        const uint32_t n = min(counts[q * kGiCounterStride], qs.cap[q]);
        const bool in = k >= start && k - start < n;
        slot = in ? qs.base[q] + (k - start) : slot;
        queue = in ? q : queue;
        found = found || in;
        start += n;
    }
    return found;
}

// Pixels of the 16x16 tiles [a, b) of a W x H image (tiles row-major).
size_t tile_range_pixels(uint32_t a, uint32_t b, int W, int H)
{
    const uint32_t tx = (uint32_t)(W + 15) / 16u;
    auto width_upto = [&](uint32_t x1) -> size_t {   // columns of tiles 0 .. x1 - 1 of a row
        return x1 == 0 ? 0 : (size_t)16 * (x1 - 1) + (size_t)std::min(16, W - 16 * (int)(x1 - 1));
    };
    size_t n = 0;
    for (uint32_t t = a; t < b;) {
        const uint32_t y = t / tx, x0 = t % tx, x1 = std::min(tx, x0 + (b - t));
        n += (width_upto(x1) - width_upto(x0)) * (size_t)std::min(16, H - 16 * (int)y);
        t += x1 - x0;
    }
    return n;
}

// The queues for gi_bounce_kernel's grid (tile_grid, tiles dealt by
// xcd_remap: XCD k takes tiles [k * per, (k + 1) * per), per = tiles / 8,
// and the tiles from 8 * per on go one each to XCDs 0, 1, ...).  Returns
// the total capacity, W * H.
uint32_t gi_queues(int W, int H, GiQueues& q)
{
    const uint32_t tiles = ((uint32_t)(W + 15) / 16u) * ((uint32_t)(H + 15) / 16u);
    const uint32_t per = tiles / 8u;
    uint32_t base = 0;
    for (uint32_t k = 0; k < 8; k++) {
        size_t cap = tile_range_pixels(k * per, (k + 1) * per, W, H);
        if (8u * per + k < tiles) cap += tile_range_pixels(8u * per + k, 8u * per + k + 1, W, H);
        q.base[k] = base;
        q.cap[k] = (uint32_t)cap;
        base += (uint32_t)cap;
    }
    return base;
}

__global__ void gi_queues_clear_kernel(uint32_t* __restrict__ counts)
{
    if (threadIdx.x < 16) counts[threadIdx.x * kGiCounterStride] = 0u;
}

// A queue record (GiJob, GiCont) as one 16-byte register quad.
typedef uint32_t Rec16 __attribute__((ext_vector_type(4)));

// Appends the record `rec` of every lane with `take` to the queue (count,
// slots from base, at most cap), one atomic per wavefront.  Every lane of
// the wavefront must reach this call.
__device__ inline void queue_append(bool take, Rec16 rec, Rec16* __restrict__ arr, uint32_t* __restrict__ count,
                                    uint32_t base, uint32_t cap)
{
    const uint64_t m = __ballot(take);
    if (!m) return;
    const int lane = (int)(threadIdx.x & 63), first = __ffsll((long long)m) - 1;
    uint32_t b = 0;
    if (lane == first) b = atomicAdd(count, (uint32_t)__popcll(m));
    b = __shfl(b, first);
    const uint32_t idx = b + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    if (take && idx < cap) arr[base + idx] = rec;
}

// queue_append with each lane's own queue (qi, 0..7): one atomic per
// wavefront and queue present.  Every lane of the wavefront must reach it.
__device__ inline void queue_append_to(bool take, uint32_t qi, Rec16 rec, Rec16* __restrict__ arr,
                                       uint32_t* __restrict__ counts, const GiQueues& qs)
{
#pragma unroll
    for (uint32_t q = 0; q < 8; q++) {
        const bool t = take && qi == q;
        if (__ballot(t) == 0ull) continue;
        queue_append(t, rec, arr, counts + q * kGiCounterStride, qs.base[q], qs.cap[q]);
    }
}

// The bounce's result for its pixel: a hit facing the sun becomes a shadow
// job (returned with job = true), every other pixel is shaded now.
__device__ inline Rec16 gi_bounce_result(const GiConst& g, const float* albedo, Vec3 o, Vec3 dir, const vpt::Hit& h,
                                         uint32_t pix, float* __restrict__ rgb, vpt_svgf_texel* __restrict__ texels,
                                         bool& job)
{
    const bool hit = h.t > 0.0f && h.block > 0;
    const int face = h.face & 0xff, block = h.block & 0xff;
    Vec3 so;
    float ndotl;
    job = hit && vpt::gi_shadow_ray(o, dir, h.t, h.face, g, so, ndotl);
    if (!job)
        vpt::gi_write_hit_pixel(g, albedo, dir, h.t, hit, face, block, 0.0f, rgb + 3 * (size_t)pix,
                                texels ? texels + pix : nullptr);
    return Rec16{pix, __float_as_uint(h.t), (uint32_t)face | ((uint32_t)block << 8), 0u};
}

__global__ void __launch_bounds__(256) gi_bounce_kernel(WorldView w, CamView cam, const float* __restrict__ gb_t,
                                                        const uint8_t* __restrict__ gb_face, GiConst g,
                                                        const vpt::AlbedoTable albedo, float* __restrict__ rgb,
                                                        vpt_svgf_texel* __restrict__ texels, GiQueues qs,
                                                        GiCont* __restrict__ conts, GiJob* __restrict__ jobs,
                                                        uint32_t* __restrict__ counts,
                                                        unsigned long long* __restrict__ counter)
{
    __shared__ float s_ray[6][256];
    __shared__ uint32_t s_pix[256];
    __shared__ int s_bin[kGiBins + 1];   // counts, then first slot of each bin; the last bin: no ray
    int px, py;
    tile_pixel(cam.W, cam.H, px, py);
    int samples = 0, rays = 0;
    bool job = false, cont = false;
    Rec16 jb = {0u, 0u, 0u, 0u}, cr = jb;   // a GiJob, a GiCont
    if (threadIdx.x <= kGiBins) s_bin[threadIdx.x] = 0;
    __syncthreads();
    // set up this pixel's ray and rank it within its y-direction bin; sky
    // pixels are shaded from the G-buffer alone, by their own lane
    Vec3 ro = vpt::mk3(0.0f, 0.0f, 0.0f), rd = ro;
    int bin = kGiBins;
    uint32_t rpix = 0xffffffffu;
    if (px < cam.W && py < cam.H) {
        const size_t i0 = (size_t)py * cam.W + px;
        const float dist = gb_t[i0];
        const Vec3 N = vpt::face_normal(gb_face[i0], 0.5f);
        if (!(dist < 0.0f)) {
            vpt::gi_ray(cam, px, py, dist, N, g.frame, ro, rd);
            bin = min(kGiBins - 1, max(0, (int)((rd.y + 1.0f) * (0.5f * kGiBins))));
            rpix = (uint32_t)i0;
        } else {
            vpt::gi_write_sky_pixel(g, N, rgb + 3 * i0, texels ? texels + i0 : nullptr);
        }
    }
    const int rank = atomicAdd(&s_bin[bin], 1);
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int b = 0; b <= kGiBins; b++) {
            const int c = s_bin[b];
            s_bin[b] = acc;
            acc += c;
        }
    }
    __syncthreads();
    const int slot = s_bin[bin] + rank;   // < 256: the slots of the block's 256 pixels
    s_ray[0][slot] = ro.x;
    s_ray[1][slot] = ro.y;
    s_ray[2][slot] = ro.z;
    s_ray[3][slot] = rd.x;
    s_ray[4][slot] = rd.y;
    s_ray[5][slot] = rd.z;
    s_pix[slot] = rpix;
    __syncthreads();
    // trace the ray in this lane's slot, at most kGiPhase1 steps
    const uint32_t pix = s_pix[threadIdx.x];
    if (pix != 0xffffffffu) {
        const Vec3 o = vpt::mk3(s_ray[0][threadIdx.x], s_ray[1][threadIdx.x], s_ray[2][threadIdx.x]);
        const Vec3 dir = vpt::mk3(s_ray[3][threadIdx.x], s_ray[4][threadIdx.x], s_ray[5][threadIdx.x]);
        vpt::Trav r;
        vpt::trav_init(r, o, dir, min(g.gi_iter, kGiPhase1));
        bool capped = false;   // stopped by the phase's step cap, not by a hit or the volume's end
        if (r.max_iter > 0) {
            uint32_t off;
            vpt::Cell c;
            while (vpt::trav_locate(w, r, off, c)) {
                const int m = vpt::plane_byte(w.df, off);
                if (!vpt::trav_advance(r, m, c)) {
                    capped = m != 0;
                    break;
                }
            }
        }
        rays = 1;
        if (capped && g.gi_iter > kGiPhase1) {
            cont = true;
            samples = r.samples;
            cr = Rec16{__float_as_uint(r.ox), __float_as_uint(r.oy), __float_as_uint(r.oz),
                       pix | ((uint32_t)r.min_idx << 28) | ((uint32_t)r.inter << 30)};
        } else {
            const vpt::Hit h = vpt::trav_finish(w, r);
            samples = h.samples;
            jb = gi_bounce_result(g, albedo.v, o, dir, h, pix, rgb, texels, job);
        }
    }
    const uint32_t q = blockIdx.x & 7u, qbase = pick8(qs.base, q), qcap = pick8(qs.cap, q);
    queue_append(cont, cr, reinterpret_cast<Rec16*>(conts), counts + q * kGiCounterStride, qbase, qcap);
    queue_append(job, jb, reinterpret_cast<Rec16*>(jobs), counts + (8 + q) * kGiCounterStride, qbase, qcap);
    count_samples(counter, samples, rays);
}

__global__ void __launch_bounds__(256) gi_cont_kernel(WorldView w, CamView cam, const float* __restrict__ gb_t,
                                                      const uint8_t* __restrict__ gb_face, GiConst g,
                                                      const vpt::AlbedoTable albedo, float* __restrict__ rgb,
                                                      vpt_svgf_texel* __restrict__ texels, GiQueues qs,
                                                      const GiCont* __restrict__ conts, GiJob* __restrict__ jobs,
                                                      uint32_t* __restrict__ counts,
                                                      unsigned long long* __restrict__ counter)
{
    uint32_t slot = 0, queue = 0;
    int samples = 0;
    bool job = false;
    Rec16 jb = {0u, 0u, 0u, 0u};
    if (queue_slot(blockIdx.x * 256u + threadIdx.x, counts, qs, slot, queue)) {
        const Rec16 cr = reinterpret_cast<const Rec16*>(conts)[slot];   // a GiCont
        const uint32_t meta = cr.w, pix = meta & 0x0fffffffu;
        const int px = (int)(pix % (uint32_t)cam.W), py = (int)(pix / (uint32_t)cam.W);
        Vec3 o, dir;
        vpt::gi_ray(cam, px, py, gb_t[pix], vpt::face_normal(gb_face[pix], 0.5f), g.frame, o, dir);
        vpt::Trav r;
        vpt::trav_init(r, o, dir, g.gi_iter);
        r.ox = __uint_as_float(cr.x);
        r.oy = __uint_as_float(cr.y);
        r.oz = __uint_as_float(cr.z);
        r.samples = kGiPhase1;   // < gi_iter: only such rays are continued
        r.min_idx = (int)((meta >> 28) & 3u);
        r.inter = (int)((meta >> 30) & 1u);
        uint32_t off;
        vpt::Cell c;
        while (vpt::trav_locate(w, r, off, c)) {
            if (!vpt::trav_advance(r, vpt::plane_byte(w.df, off), c)) break;
        }
        const vpt::Hit h = vpt::trav_finish(w, r);
        samples = h.samples - kGiPhase1;   // the first kGiPhase1 were counted by the bounce kernel
        jb = gi_bounce_result(g, albedo.v, o, dir, h, pix, rgb, texels, job);
    }
    // a pixel's job goes to the queue its continuation came from, which
    // holds a slot for every pixel of that queue's tiles
    queue_append_to(job, queue, jb, reinterpret_cast<Rec16*>(jobs), counts + 8 * kGiCounterStride, qs);
    count_samples(counter, samples, 0);
}

__global__ void __launch_bounds__(256) gi_shadow_kernel(WorldView w, CamView cam, const float* __restrict__ gb_t,
                                                        const uint8_t* __restrict__ gb_face, GiConst g,
                                                        const vpt::AlbedoTable albedo, GiQueues qs,
                                                        const GiJob* __restrict__ jobs,
                                                        const uint32_t* __restrict__ counts, float* __restrict__ rgb,
                                                        vpt_svgf_texel* __restrict__ texels,
                                                        unsigned long long* __restrict__ counter)
{
    uint32_t slot = 0, queue = 0;
    int samples = 0, rays = 0;
    if (queue_slot(blockIdx.x * 256u + threadIdx.x, counts + 8 * kGiCounterStride, qs, slot, queue)) {
        const GiJob jb = jobs[slot];
        const int px = (int)(jb.pix % (uint32_t)cam.W), py = (int)(jb.pix / (uint32_t)cam.W);
        const float dist = gb_t[jb.pix];
        Vec3 o, dir;
        vpt::gi_ray(cam, px, py, dist, vpt::face_normal(gb_face[jb.pix], 0.5f), g.frame, o, dir);
        const int face = (int)(jb.fb & 0xffu), block = (int)((jb.fb >> 8) & 0xffu);
        Vec3 so;
        float ndotl;
        vpt::gi_shadow_ray(o, dir, jb.T, face, g, so, ndotl);   // true: the job was made for it
        const vpt::Hit sh = vpt::traverse(w, so, vpt::mk3(g.sun[0], g.sun[1], g.sun[2]), g.shadow_iter);
        samples = sh.samples;
        rays = 1;
        vpt::gi_write_hit_pixel(g, albedo.v, dir, jb.T, true, face, block, sh.t > 0.0f ? 1.0f : 0.0f,
                                rgb + 3 * (size_t)jb.pix, texels ? texels + jb.pix : nullptr);
    }
    count_samples(counter, samples, rays);
}

// Ray queries: one lane per ray.
__global__ void __launch_bounds__(256) rays_kernel(WorldView w, const vpt_ray* __restrict__ rays, int n, int max_iter,
                                                   vpt_hit* __restrict__ hits,
                                                   unsigned long long* __restrict__ counter)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int samples = 0, nrays = 0;
    if (i < n) {
        vpt_ray r = rays[i];
        vpt::Hit h = vpt::traverse(w, vpt::mk3(r.origin[0], r.origin[1], r.origin[2]),
                                   vpt::mk3(r.dir[0], r.dir[1], r.dir[2]), max_iter);
        vpt_hit o;
        o.t = h.t;
        o.voxel = h.voxel;
        o.face = (uint8_t)h.face;
        o.block = (uint8_t)h.block;
        o.samples = (uint16_t)min(h.samples, 65535);
        hits[i] = o;
        samples = h.samples;
        nrays = 1;
    }
    count_samples(counter, samples, nrays);
}

// Shadow rays of GI ray queries (second launch of vpt_trace_gi_rays): the
// bounce hit in hits[i] continues toward the sun when its face sees it.
__global__ void __launch_bounds__(256) gi_rays_shadow_kernel(WorldView w, const vpt_ray* __restrict__ rays, int n,
                                                             const vpt_hit* __restrict__ hits, float lx, float ly,
                                                             float lz, int shadow_iter, uint8_t* __restrict__ vis,
                                                             unsigned long long* __restrict__ counter)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int samples = 0, nrays = 0;
    if (i < n) {
        const vpt_ray r = rays[i];
        const vpt_hit h = hits[i];
        const Vec3 L = vpt::mk3(lx, ly, lz);
        Vec3 so;
        float ndotl;
        uint8_t v = 255;
        if (h.t > 0.0f && h.block > 0 &&
            vpt::gi_shadow_origin(vpt::mk3(r.origin[0], r.origin[1], r.origin[2]),
                                  vpt::mk3(r.dir[0], r.dir[1], r.dir[2]), h.t, h.face, L, so, ndotl)) {
            vpt::Hit s = vpt::traverse(w, so, L, shadow_iter);
            v = s.t > 0.0f ? 1 : 0;
            samples = s.samples;
            nrays = 1;
        }
        vis[i] = v;
    }
    count_samples(counter, samples, nrays);
}

// ---- mirror reflections (ReflectionTraceFrag.glsl:764-1035) ---------------
// The traversal core of the reflection pass: smooth mirror reflections, one
// sample, no jitter, reprojection / LPV / player off, flat materials.  The
// target is (W / div) x (H / div); its pixels sample the full-resolution
// G-buffer at their centres: t GL LINEAR (the texel itself at div = 1),
// face GL NEAREST, both with REPEAT wrap.

__device__ inline int repeat_idx(int i, int n)
{
    const int r = i % n;
    return r < 0 ? r + n : r;
}

// texture(sampler2D LINEAR REPEAT, (s, t)) of a float plane, fp32 weights.
__device__ inline float gl_linear(const float* __restrict__ p, int W, int H, float s, float t)
{
    const float u = s * (float)W - 0.5f, v = t * (float)H - 0.5f;
    const float fu = floorf(u), fv = floorf(v);
    const float a = u - fu, b = v - fv;
    const int x0 = repeat_idx((int)fu, W), x1 = repeat_idx((int)fu + 1, W);
    const size_t r0 = (size_t)repeat_idx((int)fv, H) * W, r1 = (size_t)repeat_idx((int)fv + 1, H) * W;
    const float top = p[r0 + x0] * (1.0f - a) + p[r0 + x1] * a;
    const float bot = p[r1 + x0] * (1.0f - a) + p[r1 + x1] * a;
    return top * (1.0f - b) + bot * b;
}

// texture(sampler2D NEAREST REPEAT, (s, t)) of a byte plane.
__device__ inline int gl_nearest(const uint8_t* __restrict__ p, int W, int H, float s, float t)
{
    const int x = repeat_idx((int)floorf(s * (float)W), W), y = repeat_idx((int)floorf(t * (float)H), H);
    return p[(size_t)y * W + x];
}

struct ReflectConst {
    float lx, ly, lz;
    int div, max_iter, shadow_iter, spp, rough;
    float roughness;   // clamp(roughness, 0.01, 1)
    uint32_t frame;
};

// ImportanceSampleGGX (ReflectionTraceFrag.glsl:345-366): the GGX half
// vector around N for roughness r and randoms (x0, x1).
__device__ inline Vec3 ggx_half(Vec3 N, float r, float x0, float x1)
{
    const float alpha = r * r;
    const float alpha2 = alpha * alpha;
    float sp, cp;
    vpt_sincos_2pi(x0, &sp, &cp);
    const float cos_t = sqrtf((1.0f - x1) / (1.0f + (alpha2 - 1.0f) * x1));
    const float sin_t = sqrtf(1.0f - cos_t * cos_t);
    const float hx = cp * sin_t, hy = sp * sin_t, hz = cos_t;
    const Vec3 up = fabsf(N.z) < 0.999f ? vpt::mk3(0.0f, 0.0f, 1.0f) : vpt::mk3(1.0f, 0.0f, 0.0f);
    const Vec3 t = vpt::normalize3(vpt::cross3(up, N));
    const Vec3 b = vpt::cross3(N, t);
    return vpt::normalize3(vpt::mk3((t.x * hx + b.x * hy) + N.x * hz, (t.y * hx + b.y * hy) + N.y * hz,
                                    (t.z * hx + b.z * hy) + N.z * hz));
}

// GetReflectionDirection (:617-640): of three GGX half vectors, the first
// with the largest dot with N, randoms from the pcg_hash chain.
__device__ inline Vec3 reflection_normal(Vec3 N, float r, uint32_t& state)
{
    r = fmaxf(r, 0.05f);
    float nearest = -100.0f;
    Vec3 best = N;
#pragma unroll 1
    for (int i = 0; i < 3; i++) {
        state = vpt_pcg_hash(state);
        const float x0 = vpt_u32_to_unit(state) * 0.9f;
        state = vpt_pcg_hash(state);
        const float x1 = vpt_u32_to_unit(state) * 0.65f;
        const Vec3 h = ggx_half(N, r, x0, x1);
        const float d = vpt::dot3(h, N);
        if (d > nearest) {
            best = h;
            nearest = d;
        }
    }
    return best;
}

// GetShadowAt(HitPosition + Normal * 0.055) toward L: 1 when it hits.  A
// hit's normal is never zero (a step along an axis with a zero direction
// component moves the other coordinates by d * inf, and the ray leaves the
// volume), so the face normal is the traversal's normal vector.
__device__ inline int reflect_shadow(const WorldView& w, Vec3 o, Vec3 R, const vpt::Hit& h, Vec3 L, int shadow_iter,
                                     int& samples, int& rays)
{
    const Vec3 hn = vpt::face_normal(h.face, 0.0f);
    const Vec3 hp = vpt::mk3(o.x + R.x * h.t, o.y + R.y * h.t, o.z + R.z * h.t);
    const Vec3 so = vpt::mk3(hp.x + hn.x * 0.055f, hp.y + hn.y * 0.055f, hp.z + hn.z * 0.055f);
    const vpt::Hit sh = vpt::traverse(w, so, L, shadow_iter);
    samples += sh.samples;
    rays++;
    return sh.t > 0.0f ? 1 : 0;
}

// Hit shading of the reflection pass (ReflectionTraceFrag.glsl:582-593,
// 843-1023): the frame's denoised diffuse and sun shadow, its
// view-projection, the light and sky radiances and the albedo table.
struct ReflShade {
    const vpt_svgf_texel* diffuse;   // W x H final SVGF texels
    const uint8_t* shadow;           // W x H denoised sun shadow (unorm8)
    float vp[16];                    // u_Projection * u_View
    float rad[3];                    // SAMPLED_COLOR_MIXED substitute * 0.6
    float sky[3];                    // GetAtmosphere substitute
    vpt::AlbedoTable albedo;
};

// texture(LINEAR, REPEAT) of the SVGF texel plane's SH (4) and CoCg (2) at
// (s, t), fp32 weights; and of its AO channel; and of a unorm8 plane.
__device__ inline void texel_linear_sh(const vpt_svgf_texel* __restrict__ p, int W, int H, float s, float t,
                                      float sh[4], float cc[2], float* ao)
{
    const float u = s * (float)W - 0.5f, v = t * (float)H - 0.5f;
    const float fu = floorf(u), fv = floorf(v);
    const float a = u - fu, b = v - fv;
    const int x0 = repeat_idx((int)fu, W), x1 = repeat_idx((int)fu + 1, W);
    const size_t r0 = (size_t)repeat_idx((int)fv, H) * W, r1 = (size_t)repeat_idx((int)fv + 1, H) * W;
    const uint4 q00 = *reinterpret_cast<const uint4*>(p + r0 + x0), q10 = *reinterpret_cast<const uint4*>(p + r0 + x1);
    const uint4 q01 = *reinterpret_cast<const uint4*>(p + r1 + x0), q11 = *reinterpret_cast<const uint4*>(p + r1 + x1);
    auto lin = [&](uint32_t w00, uint32_t w10, uint32_t w01, uint32_t w11) {
        const float top = vpt::hw_h2f(w00) * (1.0f - a) + vpt::hw_h2f(w10) * a;
        const float bot = vpt::hw_h2f(w01) * (1.0f - a) + vpt::hw_h2f(w11) * a;
        return top * (1.0f - b) + bot * b;
    };
    sh[0] = lin(q00.x & 0xffffu, q10.x & 0xffffu, q01.x & 0xffffu, q11.x & 0xffffu);
    sh[1] = lin(q00.x >> 16, q10.x >> 16, q01.x >> 16, q11.x >> 16);
    sh[2] = lin(q00.y & 0xffffu, q10.y & 0xffffu, q01.y & 0xffffu, q11.y & 0xffffu);
    sh[3] = lin(q00.y >> 16, q10.y >> 16, q01.y >> 16, q11.y >> 16);
    cc[0] = lin(q00.z & 0xffffu, q10.z & 0xffffu, q01.z & 0xffffu, q11.z & 0xffffu);
    cc[1] = lin(q00.z >> 16, q10.z >> 16, q01.z >> 16, q11.z >> 16);
    if (ao) {
        auto d8 = [](uint32_t w) { return vpt_unorm8_decode((uint8_t)((w >> 16) & 0xffu)); };
        const float top = d8(q00.w) * (1.0f - a) + d8(q10.w) * a;
        const float bot = d8(q01.w) * (1.0f - a) + d8(q11.w) * a;
        *ao = top * (1.0f - b) + bot * b;
    }
}

__device__ inline float u8_linear(const uint8_t* __restrict__ p, int W, int H, float s, float t)
{
    const float u = s * (float)W - 0.5f, v = t * (float)H - 0.5f;
    const float fu = floorf(u), fv = floorf(v);
    const float a = u - fu, b = v - fv;
    const int x0 = repeat_idx((int)fu, W), x1 = repeat_idx((int)fu + 1, W);
    const size_t r0 = (size_t)repeat_idx((int)fv, H) * W, r1 = (size_t)repeat_idx((int)fv + 1, H) * W;
    const float top = vpt_unorm8_decode(p[r0 + x0]) * (1.0f - a) + vpt_unorm8_decode(p[r0 + x1]) * a;
    const float bot = vpt_unorm8_decode(p[r1 + x0]) * (1.0f - a) + vpt_unorm8_decode(p[r1 + x1]) * a;
    return top * (1.0f - b) + bot * b;
}

// SHToIrradianceA (ReflectionTraceFrag.glsl:466-475).
__device__ inline Vec3 sh_to_irradiance_a(const float sh[4], const float cc[2])
{
    const float Y = fmaxf(0.0f, 3.544905f * sh[3]);
    const float k = (Y * 0.282095f) / (sh[3] + 1e-6f);
    const float co = cc[0] * k, cg = cc[1] * k;
    const float T = Y - cg * 0.5f;
    const float G = cg + T;
    const float B = T - co * 0.5f;
    const float R = B + co;
    return vpt::mk3(fmaxf(R, 0.0f), fmaxf(G, 0.0f), fmaxf(B, 0.0f));
}

// CalculateDirectionalLight of the reflection pass (:313-336), flat
// dielectric: F0 = 0.04, kd = 1 - F; its specular term is multiplied by a
// radiance of exactly 0 with finite D and G, so it is +0 and left out.
__device__ inline Vec3 reflect_direct(Vec3 hp, Vec3 eye, Vec3 L, const float rad[3], Vec3 alb, Vec3 N, float shadow)
{
    const float sh = fminf(shadow, 1.0f);
    const Vec3 Lo = vpt::normalize3(vpt::mk3(eye.x - hp.x, eye.y - hp.y, eye.z - hp.z));
    const Vec3 Lh = vpt::normalize3(vpt::mk3(L.x + Lo.x, L.y + Lo.y, L.z + Lo.z));
    const float cos_li = fmaxf(0.0f, vpt::dot3(N, L));
    const float F = 0.04f + (1.0f - 0.04f) * vpt_powi(1.0f - fmaxf(0.0f, vpt::dot3(Lh, Lo)), 5);
    const float kd = 1.0f - F;
    const float vis = vpt::fclamp(1.0f - sh, 0.0f, 1.0f);
    return vpt::mk3(fmaxf(((kd * alb.x) * rad[0]) * cos_li, 0.0f) * vis,
                    fmaxf(((kd * alb.y) * rad[1]) * cos_li, 0.0f) * vis,
                    fmaxf(((kd * alb.z) * rad[2]) * cos_li, 0.0f) * vis);
}

// Reprojection of a reflection hit onto the frame (ReprojectReflectionToScreenSpace,
// ReflectionTraceFrag.glsl:582-593): uv of the hit through the view-projection;
// inside the 0.0035 border, the G-buffer position there (t LINEAR, the ray
// at uv) and its face (NEAREST) must match the hit (squared distance
// < 0.095, same normal).
struct Reproj {
    bool ok;
    float ru, rv, t_at;
    Vec3 p_at;
};

__device__ inline Reproj reproject(const float vp[16], const CamView& cam, int gW, int gH,
                                   const float* __restrict__ gb_t, const uint8_t* __restrict__ gb_face, Vec3 eye,
                                   Vec3 hp, Vec3 hn)
{
    Reproj r;
    r.ok = false;
    r.t_at = 0.0f;
    r.p_at = vpt::mk3(0.0f, 0.0f, 0.0f);
    const float q0 = vpt::mat_row(vp, 0, hp.x, hp.y, hp.z, 1.0f);
    const float q1 = vpt::mat_row(vp, 1, hp.x, hp.y, hp.z, 1.0f);
    const float q3 = vpt::mat_row(vp, 3, hp.x, hp.y, hp.z, 1.0f);
    r.ru = (q0 / q3) * 0.5f + 0.5f;
    r.rv = (q1 / q3) * 0.5f + 0.5f;
    const float b = 0.0035f;
    if (r.ru < 1.0f - b && r.ru > b && r.rv < 1.0f - b && r.rv > b) {
        r.t_at = gl_linear(gb_t, gW, gH, r.ru, r.rv);
        const Vec3 dir = vpt::normalize3(vpt::ray_dir_at(cam.iv, cam.ip, r.ru, r.rv));
        r.p_at = vpt::mk3(eye.x + dir.x * r.t_at, eye.y + dir.y * r.t_at, eye.z + dir.z * r.t_at);
        const Vec3 n_at = vpt::face_normal(gl_nearest(gb_face, gW, gH, r.ru, r.rv), 1.0f);
        const float dx = fabsf(r.p_at.x - hp.x), dy = fabsf(r.p_at.y - hp.y), dz = fabsf(r.p_at.z - hp.z);
        const float err = (dx * dx + dy * dy) + dz * dz;
        r.ok = err < 0.095f && n_at.x == hn.x && n_at.y == hn.y && n_at.z == hn.z;
    }
    return r;
}

// The colour one reflection hit adds: the ambient (the denoised diffuse at
// the reprojected uv, times pow(AO, 0.75) within 40 of the camera, else the
// ambient at the pixel) times the block's albedo, plus reflect_direct().
__device__ inline void shade_hit(const ReflShade& shd, int gW, int gH, const Reproj& rp, Vec3 base_amb, Vec3 eye,
                                 Vec3 L, Vec3 hp, Vec3 hn, int block, float shadow_f, float tot[3])
{
    Vec3 amb = base_amb;
    if (rp.ok) {
        float rs[4], rcc[2], ao;
        texel_linear_sh(shd.diffuse, gW, gH, rp.ru, rp.rv, rs, rcc, &ao);
        amb = sh_to_irradiance_a(rs, rcc);
        const float vx = vpt_powf(ao, 0.75f);
        if (rp.t_at > 0.0f) {
            const float ex = rp.p_at.x - eye.x, ey = rp.p_at.y - eye.y, ez = rp.p_at.z - eye.z;
            if (sqrtf((ex * ex + ey * ey) + ez * ez) < 40.0f) amb = vpt::mk3(amb.x * vx, amb.y * vx, amb.z * vx);
        }
    }
    const int ref = min(max(block, 0), 127);
    const Vec3 alb = vpt::mk3(shd.albedo.v[3 * ref], shd.albedo.v[3 * ref + 1], shd.albedo.v[3 * ref + 2]);
    amb = vpt::mk3(amb.x * alb.x, amb.y * alb.y, amb.z * alb.z);
    const Vec3 dl = reflect_direct(hp, eye, L, shd.rad, alb, hn, shadow_f);
    tot[0] = tot[0] + (amb.x + dl.x);
    tot[1] = tot[1] + (amb.y + dl.y);
    tot[2] = tot[2] + (amb.z + dl.z);
}

// A sample's record in the deferred form: the hit position, and face |
// block << 8 | hit << 16 | shadow ray traced << 17 | its result << 18.
__device__ inline uint4 reflect_record(Vec3 hp, int face, int block, bool hit, bool traced, int shadow)
{
    return make_uint4(__float_as_uint(hp.x), __float_as_uint(hp.y), __float_as_uint(hp.z),
                      (uint32_t)(face & 0xff) | ((uint32_t)(block & 0xff) << 8) | ((uint32_t)hit << 16) |
                          ((uint32_t)traced << 17) | ((uint32_t)shadow << 18));
}

// G-buffer distance and face at a target pixel (its centre sampled LINEAR /
// NEAREST when the target is smaller).
__device__ inline void reflect_gbuffer(const CamView& cam, int gW, int gH, const float* __restrict__ gb_t,
                                       const uint8_t* __restrict__ gb_face, int div, int px, int py, float& dist,
                                       int& face)
{
    if (div == 1) {
        const size_t i = (size_t)py * cam.W + px;
        dist = gb_t[i];
        face = gb_face[i];
    } else {
        const float sx = vpt::texcoord(px, cam.W), sy = vpt::texcoord(py, cam.H);
        dist = gl_linear(gb_t, gW, gH, sx, sy);
        face = gl_nearest(gb_face, gW, gH, sx, sy);
    }
}

// One lane per target pixel: spp reflected rays (max_iter steps each) and,
// for the first max(spp / 4, 1) that hit, the sun shadow ray (shadow_iter
// steps), all in the lane.
//   dist_out: binary16 o_HitDistance, -1 for sky and misses.
//   hit_out:  face | block << 8 of the first hit | shadow << 16 |
//             hits << 24, 255 without a hit.
// kMode 1 adds the hit shading (see ReflShade): each hit is reprojected
// with the frame's view-projection (reproject()); reprojected hits take
// their ambient from the denoised diffuse there and, for the
// shadow-casting samples, the denoised shadow there instead of a shadow
// ray; other hits take the ambient at the pixel itself (shade_hit()).
// color_out: the mean over spp clamped to [1e-7, 100], RGBA binary16 (sky
// pixels 0); hit_out bit 17: the first hit was reprojected; the shadow bit:
// the last shadow >= 0.5.  kMode 2 is kMode 1's traversal half: the same
// rays, reprojection tests and shadow rays, with each sample's hit written
// to `records` (reflect_record, sample-major) and the shadow bit left for
// reflect_shade_kernel, which does the shading from them.
template <bool kRough, int kMode>
__global__ void __launch_bounds__(256) reflect_kernel(WorldView w, CamView cam, int gW, int gH,
                                                      const float* __restrict__ gb_t,
                                                      const uint8_t* __restrict__ gb_face, ReflectConst rc,
                                                      uint16_t* __restrict__ dist_out, uint32_t* __restrict__ hit_out,
                                                      unsigned long long* __restrict__ counter, const ReflShade shd,
                                                      uint2* __restrict__ color_out, uint4* __restrict__ records)
{
    constexpr bool kShade = kMode == 1, kReproj = kMode >= 1;
    int px, py;
    tile_pixel(cam.W, cam.H, px, py);
    int samples = 0, rays = 0;
    if (px < cam.W && py < cam.H) {
        const size_t i = (size_t)py * cam.W + px;
        const size_t npx = (size_t)cam.W * cam.H;
        float dist;
        int face;
        reflect_gbuffer(cam, gW, gH, gb_t, gb_face, rc.div, px, py, dist, face);
        float hd = -1.0f;
        uint32_t rec = 255u;
        float col[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        if (!(dist < 0.0f)) {   // sky: o_HitDistance -1
            const Vec3 P = vpt::gbuffer_position(cam, px, py, dist);
            const Vec3 N = vpt::face_normal(face, 1.0f);   // GetNormalFromID: ids above 5 -> (1, 1, 1)
            const Vec3 eye = vpt::mk3(cam.iv[12], cam.iv[13], cam.iv[14]);
            const Vec3 I = vpt::normalize3(vpt::mk3(P.x - eye.x, P.y - eye.y, P.z - eye.z));
            const Vec3 o = vpt::mk3(P.x + N.x * 0.035f, P.y + N.y * 0.035f, P.z + N.z * 0.035f);
            const Vec3 L = vpt::mk3(rc.lx, rc.ly, rc.lz);
            const int shadow_rays = max(rc.spp / 4, 1);
            float sum = 0.001f;
            int hits = 0, shadow = 0;
            float shadow_f = 0.0f;   // ComputedShadow
            bool first_reproj = false;
            float tot[3] = {0.0f, 0.0f, 0.0f};
            Vec3 base_amb = vpt::mk3(0.0f, 0.0f, 0.0f);
            if (kShade) {   // BaseIndirectDiffuse at the target pixel
                float bs[4], bc[2];
                texel_linear_sh(shd.diffuse, gW, gH, vpt::texcoord(px, cam.W), vpt::texcoord(py, cam.H), bs, bc,
                                nullptr);
                base_amb = sh_to_irradiance_a(bs, bc);
            }
            // one hit of sample s: reprojection, shadow, shading or record
            auto on_hit = [&](int s, const Vec3& R, const vpt::Hit& h, bool trace_shadow) {
                const Vec3 hn = vpt::face_normal(h.face, 0.0f);
                const Vec3 hp = vpt::mk3(o.x + R.x * h.t, o.y + R.y * h.t, o.z + R.z * h.t);
                Reproj rp;
                rp.ok = false;
                if (kReproj) {
                    rp = reproject(shd.vp, cam, gW, gH, gb_t, gb_face, eye, hp, hn);
                    if (hits == 0) first_reproj = rp.ok;
                }
                bool traced = false;
                if (trace_shadow) {
                    if (kReproj && rp.ok) {
                        if (kShade) shadow_f = u8_linear(shd.shadow, gW, gH, rp.ru, rp.rv);
                    } else {
                        shadow = reflect_shadow(w, o, R, h, L, rc.shadow_iter, samples, rays);
                        shadow_f = (float)shadow;
                        traced = true;
                    }
                }
                if (kShade) shade_hit(shd, gW, gH, rp, base_amb, eye, L, hp, hn, h.block, shadow_f, tot);
                if (kMode == 2) records[(size_t)s * npx + i] = reflect_record(hp, h.face, h.block, true, traced, shadow);
            };
            auto on_miss = [&](int s) {
                if (kShade) {
                    tot[0] = tot[0] + shd.sky[0];
                    tot[1] = tot[1] + shd.sky[1];
                    tot[2] = tot[2] + shd.sky[2];
                }
                if (kMode == 2) records[(size_t)s * npx + i] = reflect_record(vpt::mk3(0.0f, 0.0f, 0.0f), 255, 0, false, false, 0);
            };
            if (kRough) {
                uint32_t state = vpt_pcg_hash(vpt_pcg_hash((uint32_t)px + (uint32_t)cam.W * (uint32_t)py +
                                                           (uint32_t)cam.W * (uint32_t)cam.H * rc.frame) +
                                              0x9E3779B9u);
                int shadow_itr = 0;
                for (int s = 0; s < rc.spp; s++) {
                    const Vec3 Hn = reflection_normal(N, rc.roughness, state);
                    const float k = 2.0f * vpt::dot3(Hn, I);
                    const Vec3 R = vpt::mk3(I.x - k * Hn.x, I.y - k * Hn.y, I.z - k * Hn.z);   // reflect(I, H)
                    const vpt::Hit h = vpt::traverse(w, o, R, rc.max_iter);
                    samples += h.samples;
                    rays++;
                    if (h.t > 0.0f) {
                        on_hit(s, R, h, shadow_itr < shadow_rays);
                        if (shadow_itr < shadow_rays) shadow_itr++;
                        if (hits == 0) rec = (uint32_t)(h.face & 0xff) | ((uint32_t)(h.block & 0xff) << 8);
                        sum = sum + h.t;
                        hits++;
                    } else {
                        on_miss(s);
                    }
                }
            } else if (kReproj) {
                // the reference repeats the mirror sample spp times: the ray
                // and, when not reprojected, its shadow ray are traced once
                // (the deferred form records it once, as sample 0)
                const float k = 2.0f * vpt::dot3(N, I);
                const Vec3 R = vpt::mk3(I.x - k * N.x, I.y - k * N.y, I.z - k * N.z);   // reflect(I, N)
                const vpt::Hit h = vpt::traverse(w, o, R, rc.max_iter);
                samples += h.samples;
                rays++;
                int shadow_itr = 0;
                for (int s = 0; s < rc.spp; s++) {
                    if (h.t > 0.0f) {
                        if (kMode == 1 || s == 0) on_hit(s, R, h, shadow_itr == 0);
                        if (shadow_itr < shadow_rays) shadow_itr++;
                        if (hits == 0) rec = (uint32_t)(h.face & 0xff) | ((uint32_t)(h.block & 0xff) << 8);
                        sum = sum + h.t;
                        hits++;
                    } else {
                        if (kMode == 1 || s == 0) on_miss(s);
                    }
                }
            } else {
                const float k = 2.0f * vpt::dot3(N, I);
                const Vec3 R = vpt::mk3(I.x - k * N.x, I.y - k * N.y, I.z - k * N.z);   // reflect(I, N)
                const vpt::Hit h = vpt::traverse(w, o, R, rc.max_iter);
                samples += h.samples;
                rays++;
                if (h.t > 0.0f) {
                    shadow = reflect_shadow(w, o, R, h, L, rc.shadow_iter, samples, rays);
                    rec = (uint32_t)(h.face & 0xff) | ((uint32_t)(h.block & 0xff) << 8);
                    for (int s = 0; s < rc.spp; s++) sum = sum + h.t;
                    hits = rc.spp;
                }
            }
            if (kShade) shadow = shadow_f >= 0.5f ? 1 : 0;
            if (kMode == 2) shadow = 0;   // set by reflect_shade_kernel
            if (hits > 0) {
                hd = sum / fmaxf((float)hits, 0.01f);
                rec |= ((uint32_t)shadow << 16) | ((uint32_t)hits << 24) | ((uint32_t)first_reproj << 17);
            }
            hd = vpt::fclamp(hd, -10.0f, 200.0f);
            if (kShade) {
                const float n = (float)rc.spp;
                for (int c = 0; c < 3; c++) col[c] = vpt::fclamp(tot[c] / n, 0.0000001f, 100.0f);
                col[3] = vpt::fclamp(n / n, 0.0000001f, 100.0f);
            }
        }
        dist_out[i] = (uint16_t)vpt::hw_f2h(hd);
        if (hit_out) hit_out[i] = rec;
        if (kShade)
            color_out[i] = make_uint2(vpt::hw_f2h(col[0]) | (vpt::hw_f2h(col[1]) << 16),
                                      vpt::hw_f2h(col[2]) | (vpt::hw_f2h(col[3]) << 16));
    }
    count_samples(counter, samples, rays);
}

// reflect_kernel<true, kMode> (kMode 0 or 2) for spp = 2 with a lane per
// (pixel, sample): the reference's settings give one 480x270 target pixel
// per 16 screen pixels, about two wavefronts per SIMD with a lane per pixel,
// and each lane then traced both samples in turn.  Here a pair of adjacent
// lanes traces the two samples of one pixel at once (the second lane first
// advances the pixel's PCG state past the first sample's six draws), and
// the pair exchanges what the sequential loop carries from sample 0 to
// sample 1: whether sample 0 hit (then sample 1 casts no shadow ray),
// sample 0's shadow result (the value sample 1's record holds when it casts
// none), and, for the pixel's outputs, the hits' distances, faces, blocks,
// reprojection flags and shadow results, which the sample-0 lane combines in
// sample order.  Records, distances, hit records and ray / sample counts
// equal reflect_kernel<true, kMode>'s.
template <int kMode>
__global__ void __launch_bounds__(256) reflect_rough2_kernel(WorldView w, CamView cam, int gW, int gH,
                                                                      const float* __restrict__ gb_t,
                                                                      const uint8_t* __restrict__ gb_face,
                                                                      ReflectConst rc, uint16_t* __restrict__ dist_out,
                                                                      uint32_t* __restrict__ hit_out,
                                                                      unsigned long long* __restrict__ counter,
                                                                      const ReflShade shd, uint4* __restrict__ records)
{
    // 16x8 pixels per block: each wavefront 8x4 pixels, two lanes per pixel
    const uint32_t tiles_x = (uint32_t)(cam.W + 15) / 16u, tiles_y = (uint32_t)(cam.H + 7) / 8u;
    const uint32_t tile = vpt::xcd_remap(blockIdx.x, tiles_x * tiles_y);
    const int wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
    const int sample = lane & 1, pl = lane >> 1;
    const int px = (int)(tile % tiles_x) * 16 + (wave & 1) * 8 + (pl & 7);
    const int py = (int)(tile / tiles_x) * 8 + (wave >> 1) * 4 + (pl >> 3);
    const bool valid = px < cam.W && py < cam.H;   // the same for both lanes of a pair
    const size_t npx = (size_t)cam.W * cam.H;
    const size_t i = valid ? (size_t)py * cam.W + px : 0;
    int samples = 0, rays = 0;
    bool sky = true, hit = false, reproj = false, traced = false;
    int shadow_mine = 0;
    float ht = 0.0f;
    uint32_t fb = 255u;   // face | block << 8 of this sample's hit
    Vec3 hp = vpt::mk3(0.0f, 0.0f, 0.0f);
    Vec3 o = hp, R = hp, L = vpt::mk3(rc.lx, rc.ly, rc.lz);
    vpt::Hit h{};
    if (valid) {
        float dist;
        int face;
        reflect_gbuffer(cam, gW, gH, gb_t, gb_face, rc.div, px, py, dist, face);
        if (!(dist < 0.0f)) {
            sky = false;
            const Vec3 P = vpt::gbuffer_position(cam, px, py, dist);
            const Vec3 N = vpt::face_normal(face, 1.0f);
            const Vec3 eye = vpt::mk3(cam.iv[12], cam.iv[13], cam.iv[14]);
            const Vec3 I = vpt::normalize3(vpt::mk3(P.x - eye.x, P.y - eye.y, P.z - eye.z));
            o = vpt::mk3(P.x + N.x * 0.035f, P.y + N.y * 0.035f, P.z + N.z * 0.035f);
            uint32_t state = vpt_pcg_hash(vpt_pcg_hash((uint32_t)px + (uint32_t)cam.W * (uint32_t)py +
                                                       (uint32_t)cam.W * (uint32_t)cam.H * rc.frame) +
                                          0x9E3779B9u);
            if (sample == 1)
                for (int k = 0; k < 6; k++) state = vpt_pcg_hash(state);   // sample 0's draws
            const Vec3 Hn = reflection_normal(N, rc.roughness, state);
            const float k = 2.0f * vpt::dot3(Hn, I);
            R = vpt::mk3(I.x - k * Hn.x, I.y - k * Hn.y, I.z - k * Hn.z);
            h = vpt::traverse(w, o, R, rc.max_iter);
            samples += h.samples;
            rays++;
            hit = h.t > 0.0f;
            if (hit) {
                const Vec3 hn = vpt::face_normal(h.face, 0.0f);
                hp = vpt::mk3(o.x + R.x * h.t, o.y + R.y * h.t, o.z + R.z * h.t);
                if (kMode == 2) reproj = reproject(shd.vp, cam, gW, gH, gb_t, gb_face, eye, hp, hn).ok;
                ht = h.t;
                fb = (uint32_t)(h.face & 0xff) | ((uint32_t)(h.block & 0xff) << 8);
            }
        }
    }
    // sample 0's hit decides whether sample 1 casts the pixel's one shadow ray
    const bool hit_other = __shfl_xor((int)hit, 1) != 0;
    const bool hit0 = sample == 0 ? hit : hit_other;
    if (hit && (sample == 0 || !hit0) && !reproj) {
        shadow_mine = reflect_shadow(w, o, R, h, L, rc.shadow_iter, samples, rays);
        traced = true;
    }
    const int shadow_other = __shfl_xor(shadow_mine, 1);
    const float ht_other = __shfl_xor(ht, 1);
    const uint32_t fb_other = (uint32_t)__shfl_xor((int)fb, 1);
    const bool reproj_other = __shfl_xor((int)reproj, 1) != 0;
    const bool traced_other = __shfl_xor((int)traced, 1) != 0;
    if (kMode == 2 && valid && !sky) {
        // the record's shadow field is the loop's running value: sample 1
        // without a ray of its own keeps sample 0's
        const int shadow_rec = (sample == 0 || traced) ? shadow_mine : shadow_other;
        records[(size_t)sample * npx + i] = hit ? reflect_record(hp, (int)(fb & 0xffu), (int)(fb >> 8), true, traced,
                                                                 shadow_rec)
                                                : reflect_record(vpt::mk3(0.0f, 0.0f, 0.0f), 255, 0, false, false, 0);
    }
    if (valid && sample == 0) {
        float hd = -1.0f;
        uint32_t rec = 255u;
        if (!sky) {
            float sum = 0.001f;
            int hits = 0;
            bool first_reproj = false;
            if (hit) {
                rec = fb;
                first_reproj = reproj;
                sum = sum + ht;
                hits = 1;
            }
            if (hit_other) {
                if (hits == 0) {
                    rec = fb_other;
                    first_reproj = reproj_other;
                }
                sum = sum + ht_other;
                hits++;
            }
            // the loop's last shadow value (kMode 0; kMode 2 leaves the bit to
            // reflect_shade_kernel): at most one of the two samples casts a ray
            const int shadow = kMode == 2 ? 0 : (traced ? shadow_mine : (traced_other ? shadow_other : 0));
            if (hits > 0) {
                hd = sum / fmaxf((float)hits, 0.01f);
                rec |= ((uint32_t)shadow << 16) | ((uint32_t)hits << 24) | ((uint32_t)first_reproj << 17);
            }
            hd = vpt::fclamp(hd, -10.0f, 200.0f);
        }
        dist_out[i] = (uint16_t)vpt::hw_f2h(hd);
        if (hit_out) hit_out[i] = rec;
    }
    count_samples(counter, samples, rays);
}

// The shading half of the deferred form: per target pixel, the samples of
// reflect_kernel<., 2>'s records in order -- the same reprojection (from
// the recorded hit position), the same choice of which samples take a
// shadow and from where (the denoised plane when reprojected, else the
// recorded shadow ray), the same shade_hit() -- then the colour and the hit
// record's shadow bit, exactly as reflect_kernel<., 1> computes them.
template <bool kRough>
__global__ void __launch_bounds__(256) reflect_shade_kernel(CamView cam, int gW, int gH,
                                                            const float* __restrict__ gb_t,
                                                            const uint8_t* __restrict__ gb_face, ReflectConst rc,
                                                            const ReflShade shd, const uint4* __restrict__ records,
                                                            uint32_t* __restrict__ hit_out,
                                                            uint2* __restrict__ color_out)
{
    int px, py;
    tile_pixel(cam.W, cam.H, px, py);
    if (px >= cam.W || py >= cam.H) return;
    const size_t i = (size_t)py * cam.W + px;
    const size_t npx = (size_t)cam.W * cam.H;
    float dist;
    int face;
    reflect_gbuffer(cam, gW, gH, gb_t, gb_face, rc.div, px, py, dist, face);
    float col[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (!(dist < 0.0f)) {
        const Vec3 eye = vpt::mk3(cam.iv[12], cam.iv[13], cam.iv[14]);
        const Vec3 L = vpt::mk3(rc.lx, rc.ly, rc.lz);
        const int shadow_rays = max(rc.spp / 4, 1);
        float bs[4], bc[2];
        texel_linear_sh(shd.diffuse, gW, gH, vpt::texcoord(px, cam.W), vpt::texcoord(py, cam.H), bs, bc, nullptr);
        const Vec3 base_amb = sh_to_irradiance_a(bs, bc);
        float shadow_f = 0.0f;
        float tot[3] = {0.0f, 0.0f, 0.0f};
        int hits = 0, shadow_itr = 0;
        for (int s = 0; s < rc.spp; s++) {
            const uint4 r = records[(size_t)(kRough ? s : 0) * npx + i];
            if ((r.w >> 16) & 1u) {
                const int hface = (int)(r.w & 0xffu), hblock = (int)((r.w >> 8) & 0xffu);
                const Vec3 hp = vpt::mk3(__uint_as_float(r.x), __uint_as_float(r.y), __uint_as_float(r.z));
                const Vec3 hn = vpt::face_normal(hface, 0.0f);
                const Reproj rp = reproject(shd.vp, cam, gW, gH, gb_t, gb_face, eye, hp, hn);
                const bool trace_shadow = kRough ? shadow_itr < shadow_rays : shadow_itr == 0;
                if (trace_shadow)
                    shadow_f = rp.ok ? u8_linear(shd.shadow, gW, gH, rp.ru, rp.rv) : (float)((r.w >> 18) & 1u);
                if (shadow_itr < shadow_rays) shadow_itr++;
                shade_hit(shd, gW, gH, rp, base_amb, eye, L, hp, hn, hblock, shadow_f, tot);
                hits++;
            } else {
                tot[0] = tot[0] + shd.sky[0];
                tot[1] = tot[1] + shd.sky[1];
                tot[2] = tot[2] + shd.sky[2];
            }
        }
        if (hits > 0 && hit_out) hit_out[i] = (hit_out[i] & ~(1u << 16)) | ((shadow_f >= 0.5f ? 1u : 0u) << 16);
        const float n = (float)rc.spp;
        for (int c = 0; c < 3; c++) col[c] = vpt::fclamp(tot[c] / n, 0.0000001f, 100.0f);
        col[3] = vpt::fclamp(n / n, 0.0000001f, 100.0f);
    }
    color_out[i] = make_uint2(vpt::hw_f2h(col[0]) | (vpt::hw_f2h(col[1]) << 16),
                              vpt::hw_f2h(col[2]) | (vpt::hw_f2h(col[3]) << 16));
}

dim3 tile_grid(int W, int H) { return dim3((unsigned)(((W + 15) / 16) * ((H + 15) / 16))); }

}  // namespace

static int primary_launch(vpt_world* w, const vpt_camera* cam, int max_iter, const vpt_gbuffer* out, hipStream_t s,
                          uint32_t* gi_counts)
{
    hipLaunchKernelGGL(primary_kernel, tile_grid(cam->width, cam->height), dim3(256), 0, s, vpt::make_view(w),
                       vpt::make_cam(cam), max_iter, out->d_t, out->d_face, out->d_block, out->d_voxel, w->d_counter,
                       gi_counts);
    HIP_TRY(hipGetLastError());
    return VPT_OK;
}

extern "C" VPT_API int vpt_trace_primary(vpt_world* w, const vpt_camera* cam, int max_iter, const vpt_gbuffer* out,
                                         void* stream)
{
    if (!w) return set_error(VPT_EINVAL, "vpt_trace_primary: null world");
    if (int rc = vpt::check_cam(cam, "vpt_trace_primary")) return rc;
    if (int rc = vpt::check_gbuffer(out, cam, "vpt_trace_primary", true)) return rc;
    if (int rc = vpt::check_iter(max_iter, "vpt_trace_primary")) return rc;
    DeviceGuard g(w->device);
    return primary_launch(w, cam, max_iter, out, (hipStream_t)stream, nullptr);
}

extern "C" VPT_API int vpt_trace_shadow(vpt_world* w, const vpt_camera* cam, const vpt_gbuffer* gb,
                                        const float sun_dir[3], int max_iter, uint8_t* d_mask, void* stream)
{
    if (!w || !sun_dir || !d_mask) return set_error(VPT_EINVAL, "vpt_trace_shadow: null argument");
    if (int rc = vpt::check_cam(cam, "vpt_trace_shadow")) return rc;
    if (int rc = vpt::check_gbuffer(gb, cam, "vpt_trace_shadow", false)) return rc;
    if (int rc = vpt::check_iter(max_iter, "vpt_trace_shadow")) return rc;
    DeviceGuard g(w->device);
    hipLaunchKernelGGL(shadow_kernel, tile_grid(cam->width, cam->height), dim3(256), 0, (hipStream_t)stream,
                       vpt::make_view(w), vpt::make_cam(cam), gb->d_t, gb->d_face, sun_dir[0], sun_dir[1], sun_dir[2],
                       max_iter, d_mask, w->d_counter);
    HIP_TRY(hipGetLastError());
    return VPT_OK;
}

extern "C" VPT_API int vpt_trace_shadow2(vpt_world* w, const vpt_camera* cam, const vpt_gbuffer* gb,
                                         const float sun_dir[3], int max_iter, uint8_t* d_mask,
                                         uint16_t* d_transversal, void* stream)
{
    if (!w || !sun_dir || !d_mask) return set_error(VPT_EINVAL, "vpt_trace_shadow2: null argument");
    if (int rc = vpt::check_cam(cam, "vpt_trace_shadow2")) return rc;
    if (int rc = vpt::check_gbuffer(gb, cam, "vpt_trace_shadow2", false)) return rc;
    if (int rc = vpt::check_iter(max_iter, "vpt_trace_shadow2")) return rc;
    DeviceGuard g(w->device);
    hipLaunchKernelGGL(shadow2_kernel, tile_grid(cam->width, cam->height), dim3(256), 0, (hipStream_t)stream,
                       vpt::make_view(w), vpt::make_cam(cam), gb->d_t, gb->d_face, sun_dir[0], sun_dir[1], sun_dir[2],
                       max_iter, d_mask, d_transversal, w->d_counter);
    HIP_TRY(hipGetLastError());
    return VPT_OK;
}

static int reflect_check(const char* fn, const vpt_camera* cam, const vpt_gbuffer* gb, const vpt_reflect_params* prm)
{
    if (int rc = vpt::check_cam(cam, fn)) return rc;
    if (int rc = vpt::check_gbuffer(gb, cam, fn, false)) return rc;
    if (int rc = vpt::check_iter(prm->max_iter, fn)) return rc;
    if (int rc = vpt::check_iter(prm->shadow_iter, fn)) return rc;
    const int div = prm->div;
    if (div < 1 || div > 16 || cam->width / div < 1 || cam->height / div < 1)
        return set_error(VPT_EINVAL, "%s: div must be 1..16 and leave at least one pixel", fn);
    if (prm->spp < 1 || prm->spp > 16) return set_error(VPT_EINVAL, "%s: spp %d not in 1..16", fn, prm->spp);
    if (prm->rough != 0 && prm->rough != 1) return set_error(VPT_EINVAL, "%s: rough must be 0 or 1", fn);
    if (!(prm->roughness == prm->roughness)) return set_error(VPT_EINVAL, "%s: roughness is NaN", fn);
    return VPT_OK;
}

static ReflectConst reflect_const(const vpt_reflect_params* prm)
{
    return {prm->light_dir[0], prm->light_dir[1], prm->light_dir[2], prm->div, prm->max_iter,
            prm->shadow_iter, prm->spp, prm->rough, fminf(fmaxf(prm->roughness, 0.01f), 1.0f),
            prm->frame};
}

static ReflShade reflect_shade_of(const vpt_reflect_shading* shading)
{
    ReflShade shd{};
    shd.diffuse = shading->d_diffuse;
    shd.shadow = shading->d_shadow;
    memcpy(shd.vp, shading->viewproj, sizeof(shd.vp));
    for (int c = 0; c < 3; c++) {
        shd.rad[c] = shading->sun_radiance[c] * 0.6f;
        shd.sky[c] = shading->sky_radiance[c];
    }
    static_assert(sizeof(shd.albedo.v) == sizeof(shading->albedo), "albedo table size");
    memcpy(shd.albedo.v, shading->albedo, sizeof(shd.albedo.v));
    return shd;
}

// mode 0: plain, 1: with the hit shading, 2: the deferred traversal half
// (viewproj from `vp`, records to d_records)
static int reflect_launch(const char* fn, vpt_world* w, const vpt_camera* cam, const vpt_gbuffer* gb,
                          const vpt_reflect_params* prm, int mode, const vpt_reflect_shading* shading,
                          const float* vp, uint16_t* d_dist, uint32_t* d_hit, uint16_t* d_color, void* d_records,
                          void* stream)
{
    if (!w || !prm || !d_dist) return set_error(VPT_EINVAL, "%s: null argument", fn);
    if (int rc = reflect_check(fn, cam, gb, prm)) return rc;
    ReflShade shd{};
    if (mode == 1) {
        if (!shading->d_diffuse || !shading->d_shadow || !d_color)
            return set_error(VPT_EINVAL, "%s: diffuse, shadow and colour planes are required", fn);
        shd = reflect_shade_of(shading);
    } else if (mode == 2) {
        if (!vp || !d_records) return set_error(VPT_EINVAL, "%s: viewproj and records are required", fn);
        memcpy(shd.vp, vp, sizeof(shd.vp));
    }
    DeviceGuard g(w->device);
    CamView cv = vpt::make_cam(cam);
    cv.W = cam->width / prm->div;
    cv.H = cam->height / prm->div;
    const ReflectConst rc = reflect_const(prm);
    uint2* col = reinterpret_cast<uint2*>(d_color);
    uint4* recs = reinterpret_cast<uint4*>(d_records);
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, tile_grid(cv.W, cv.H), dim3(256), 0, (hipStream_t)stream, vpt::make_view(w), cv,
                           cam->width, cam->height, gb->d_t, gb->d_face, rc, d_dist, d_hit, w->d_counter, shd, col,
                           recs);
    };
    if (mode == 1) {
        rc.rough ? go(reflect_kernel<true, 1>) : go(reflect_kernel<false, 1>);
    } else if (mode != 1 && rc.rough && rc.spp == 2) {
        // the reference's setting: a lane per (pixel, sample), 16x8 pixels per block
        const unsigned blocks = (unsigned)(((cv.W + 15) / 16) * ((cv.H + 7) / 8));
        hipLaunchKernelGGL(mode == 2 ? reflect_rough2_kernel<2> : reflect_rough2_kernel<0>, dim3(blocks), dim3(256), 0,
                           (hipStream_t)stream, vpt::make_view(w), cv, cam->width, cam->height, gb->d_t, gb->d_face, rc,
                           d_dist, d_hit, w->d_counter, shd, recs);
    } else if (mode == 2) {
        rc.rough ? go(reflect_kernel<true, 2>) : go(reflect_kernel<false, 2>);
    } else {
        rc.rough ? go(reflect_kernel<true, 0>) : go(reflect_kernel<false, 0>);
    }
    HIP_TRY(hipGetLastError());
    return VPT_OK;
}

extern "C" VPT_API int vpt_trace_reflect(vpt_world* w, const vpt_camera* cam, const vpt_gbuffer* gb,
                                         const vpt_reflect_params* prm, uint16_t* d_dist, uint32_t* d_hit,
                                         void* stream)
{
    return reflect_launch("vpt_trace_reflect", w, cam, gb, prm, 0, nullptr, nullptr, d_dist, d_hit, nullptr, nullptr,
                          stream);
}

extern "C" VPT_API int vpt_trace_reflect_shaded(vpt_world* w, const vpt_camera* cam, const vpt_gbuffer* gb,
                                                const vpt_reflect_params* prm, const vpt_reflect_shading* shading,
                                                uint16_t* d_dist, uint32_t* d_hit, uint16_t* d_color, void* stream)
{
    if (!shading) return set_error(VPT_EINVAL, "vpt_trace_reflect_shaded: null shading");
    return reflect_launch("vpt_trace_reflect_shaded", w, cam, gb, prm, 1, shading, nullptr, d_dist, d_hit, d_color,
                          nullptr, stream);
}

extern "C" VPT_API int vpt_trace_reflect_deferred(vpt_world* w, const vpt_camera* cam, const vpt_gbuffer* gb,
                                                  const vpt_reflect_params* prm, const float viewproj[16],
                                                  uint16_t* d_dist, uint32_t* d_hit, void* d_records, void* stream)
{
    return reflect_launch("vpt_trace_reflect_deferred", w, cam, gb, prm, 2, nullptr, viewproj, d_dist, d_hit, nullptr,
                          d_records, stream);
}

extern "C" VPT_API int vpt_shade_reflect(vpt_ctx* ctx, const vpt_camera* cam, const vpt_gbuffer* gb,
                                         const vpt_reflect_params* prm, const vpt_reflect_shading* shading,
                                         const void* d_records, uint32_t* d_hit, uint16_t* d_color, void* stream)
{
    const char* fn = "vpt_shade_reflect";
    if (!ctx || !prm || !shading || !d_records || !d_color) return set_error(VPT_EINVAL, "%s: null argument", fn);
    if (int rc = reflect_check(fn, cam, gb, prm)) return rc;
    if (!shading->d_diffuse || !shading->d_shadow)
        return set_error(VPT_EINVAL, "%s: diffuse and shadow planes are required", fn);
    DeviceGuard g(ctx->device);
    CamView cv = vpt::make_cam(cam);
    cv.W = cam->width / prm->div;
    cv.H = cam->height / prm->div;
    const ReflectConst rc = reflect_const(prm);
    const ReflShade shd = reflect_shade_of(shading);
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, tile_grid(cv.W, cv.H), dim3(256), 0, (hipStream_t)stream, cv, cam->width,
                           cam->height, gb->d_t, gb->d_face, rc, shd, reinterpret_cast<const uint4*>(d_records), d_hit,
                           reinterpret_cast<uint2*>(d_color));
    };
    rc.rough ? go(reflect_shade_kernel<true>) : go(reflect_shade_kernel<false>);
    HIP_TRY(hipGetLastError());
    return VPT_OK;
}

extern "C" VPT_API int vpt_reflect_shading_from_gi(const vpt_gi_params* gi, const float viewproj[16],
                                                   const vpt_svgf_texel* d_diffuse, const uint8_t* d_shadow,
                                                   vpt_reflect_shading* out)
{
    if (!gi || !viewproj || !out) return set_error(VPT_EINVAL, "vpt_reflect_shading_from_gi: null argument");
    vpt_reflect_shading r{};
    r.d_diffuse = d_diffuse;
    r.d_shadow = d_shadow;
    memcpy(r.viewproj, viewproj, sizeof(r.viewproj));
    memcpy(r.sun_radiance, gi->light_color, sizeof(r.sun_radiance));
    memcpy(r.sky_radiance, gi->sky_radiance, sizeof(r.sky_radiance));
    memcpy(r.albedo, gi->albedo, sizeof(r.albedo));
    *out = r;
    return VPT_OK;
}

// The GI pass's kernels over the given scratch (the job counter is cleared
// by a kernel, so the sequence can be captured in a graph).
// The GI pass's kernels over the given scratch: 16 queue counters, then
// continuation and job arrays of at least one record per pixel each.  The
// counters are cleared by a kernel, so the sequence can be captured in a
// graph.
// Where vpt_trace_frame starts the shadow pass beside the GI pass: after
// the bounce kernel, so that it fills the chip beside the continuation and
// shadow-job kernels, whose few long rays leave most CUs idle, rather than
// competing with the bounce kernel, which fills the chip by itself.
// 0: before the GI pass, 1: after the bounce kernel, 2: after the
// continuation kernel.
constexpr int kFrameForkStage = 1;

static int gi_enqueue(vpt_world* w, const vpt_camera* cam, const vpt_gbuffer* gb, const vpt_gi_params* prm,
                      const vpt::AlbedoTable& albedo, float* d_rgb, vpt_svgf_texel* d_svgf_in, uint32_t* counts,
                      GiCont* conts, GiJob* jobs, hipStream_t s, hipEvent_t fork = nullptr, bool cleared = false)
{
    const GiConst gc = vpt::make_gi(prm);
    const CamView cv = vpt::make_cam(cam);
    GiQueues qs;
    const uint32_t total = gi_queues(cam->width, cam->height, qs);   // = W * H
    const dim3 qgrid((total + 255u) / 256u);
    if (fork && kFrameForkStage == 0) HIP_TRY(hipEventRecord(fork, s));
    if (!cleared) {
        hipLaunchKernelGGL(gi_queues_clear_kernel, dim3(1), dim3(64), 0, s, counts);
        HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(gi_bounce_kernel, tile_grid(cam->width, cam->height), dim3(256), 0, s, vpt::make_view(w), cv,
                       gb->d_t, gb->d_face, gc, albedo, d_rgb, d_svgf_in, qs, conts, jobs, counts, w->d_counter);
    HIP_TRY(hipGetLastError());
    if (fork && kFrameForkStage == 1) HIP_TRY(hipEventRecord(fork, s));
    if (prm->gi_iter > kGiPhase1) {
        hipLaunchKernelGGL(gi_cont_kernel, qgrid, dim3(256), 0, s, vpt::make_view(w), cv, gb->d_t, gb->d_face, gc,
                           albedo, d_rgb, d_svgf_in, qs, conts, jobs, counts, w->d_counter);
        HIP_TRY(hipGetLastError());
    }
    if (fork && kFrameForkStage == 2) HIP_TRY(hipEventRecord(fork, s));
    hipLaunchKernelGGL(gi_shadow_kernel, qgrid, dim3(256), 0, s, vpt::make_view(w), cv, gb->d_t, gb->d_face, gc,
                       albedo, qs, jobs, counts, d_rgb, d_svgf_in, w->d_counter);
    HIP_TRY(hipGetLastError());
    return VPT_OK;
}

// The GI pass on `s` over the context's scratch for `s` (arguments already
// checked); `fork`, when given, is recorded on `s` at kFrameForkStage.
// With `primary` (vpt_trace_frame), the primary pass is launched on `s`
// first, once the scratch is this call's, and clears the queue counters.
struct PrimaryFirst {
    const vpt_gbuffer* gb;
    int max_iter;
};

static int gi_pass(vpt_world* w, const vpt_camera* cam, const vpt_gbuffer* gb, const vpt_gi_params* prm, float* d_rgb,
                   vpt_svgf_texel* d_svgf_in, hipStream_t s, hipEvent_t fork, const PrimaryFirst* primary = nullptr)
{
    vpt_ctx* ctx = w->ctx;
    const vpt::AlbedoTable albedo = vpt::take_albedo(ctx, prm);
    const size_t npx = (size_t)cam->width * cam->height;
    // the scratch stays this stream's until the launches below are enqueued
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIP_TRY(hipStreamIsCapturing(s, &cs));
    const bool capturing = cs != hipStreamCaptureStatusNone;
    std::lock_guard<std::mutex> lk(ctx->mu);
    vpt::GiScratch* sc = nullptr;
    if (int rc = gi_scratch_for(ctx, s, npx, capturing, &sc)) return rc;
    uint32_t* counts = reinterpret_cast<uint32_t*>(sc->d_buf);
    if (primary)
        if (int rc = primary_launch(w, cam, primary->max_iter, primary->gb, s, counts)) return rc;
    if (int rc = gi_enqueue(w, cam, gb, prm, albedo, d_rgb, d_svgf_in, counts,
                            reinterpret_cast<GiCont*>(sc->d_buf + vpt::kGiCountersBytes),
                            reinterpret_cast<GiJob*>(sc->d_buf + vpt::kGiCountersBytes + sizeof(GiCont) * sc->cap), s,
                            fork, primary != nullptr))
        return rc;
    if (!capturing) HIP_TRY(hipEventRecord((hipEvent_t)sc->done, s));
    return VPT_OK;
}

extern "C" VPT_API int vpt_trace_gi(vpt_world* w, const vpt_camera* cam, const vpt_gbuffer* gb,
                                    const vpt_gi_params* prm, float* d_rgb, vpt_svgf_texel* d_svgf_in, void* stream)
{
    if (!w || !prm || !d_rgb) return set_error(VPT_EINVAL, "vpt_trace_gi: null argument");
    if (int rc = vpt::check_cam(cam, "vpt_trace_gi")) return rc;
    if (int rc = vpt::check_gbuffer(gb, cam, "vpt_trace_gi", false)) return rc;
    if (int rc = vpt::check_iter(prm->gi_iter, "vpt_trace_gi")) return rc;
    if (int rc = vpt::check_iter(prm->shadow_iter, "vpt_trace_gi")) return rc;
    DeviceGuard g(w->device);
    return gi_pass(w, cam, gb, prm, d_rgb, d_svgf_in, (hipStream_t)stream, nullptr);
}

extern "C" VPT_API int vpt_trace_frame(vpt_world* w, const vpt_camera* cam, const vpt_gi_params* prm,
                                       int primary_iter, int shadow_iter, const vpt_gbuffer* gb, uint8_t* d_mask,
                                       uint16_t* d_transversal, float* d_rgb, vpt_svgf_texel* d_svgf_in,
                                       void* stream)
{
    if (!w || !prm || !d_mask || !d_rgb) return set_error(VPT_EINVAL, "vpt_trace_frame: null argument");
    if (int rc = vpt::check_cam(cam, "vpt_trace_frame")) return rc;
    if (int rc = vpt::check_gbuffer(gb, cam, "vpt_trace_frame", true)) return rc;
    if (int rc = vpt::check_iter(primary_iter, "vpt_trace_frame")) return rc;
    if (int rc = vpt::check_iter(shadow_iter, "vpt_trace_frame")) return rc;
    if (int rc = vpt::check_iter(prm->gi_iter, "vpt_trace_frame")) return rc;
    if (int rc = vpt::check_iter(prm->shadow_iter, "vpt_trace_frame")) return rc;
    DeviceGuard g(w->device);
    vpt_ctx* ctx = w->ctx;
    if (!ctx->side_stream) {
        hipStream_t side = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        hipError_t e = hipStreamCreateWithFlags(&side, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&e0, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&e1, hipEventDisableTiming);
        if (e != hipSuccess) {
            if (e0) (void)hipEventDestroy(e0);
            if (side) (void)hipStreamDestroy(side);
            return set_error(VPT_EDEVICE, "vpt_trace_frame: side stream: %s", hipGetErrorString(e));
        }
        ctx->side_stream = side;
        ctx->ev_fork = e0;
        ctx->ev_join = e1;
    }
    hipStream_t s = (hipStream_t)stream, side = (hipStream_t)ctx->side_stream;
    // the primary pass, then the GI pass on `stream`, which records the fork
    // event partway (after its bounce kernel); the shadow pass, which like
    // GI only reads the G-buffer, starts there on the side stream; join
    // before anything later on `stream`
    const PrimaryFirst primary = {gb, primary_iter};
    if (int rc = gi_pass(w, cam, gb, prm, d_rgb, d_svgf_in, s, (hipEvent_t)ctx->ev_fork, &primary)) return rc;
    HIP_TRY(hipStreamWaitEvent(side, (hipEvent_t)ctx->ev_fork, 0));
    const int rc_shadow = d_transversal
                              ? vpt_trace_shadow2(w, cam, gb, prm->sun_dir, shadow_iter, d_mask, d_transversal, side)
                              : vpt_trace_shadow(w, cam, gb, prm->sun_dir, shadow_iter, d_mask, side);
    HIP_TRY(hipEventRecord((hipEvent_t)ctx->ev_join, side));
    HIP_TRY(hipStreamWaitEvent(s, (hipEvent_t)ctx->ev_join, 0));
    return rc_shadow;
}


// ---------------------------------------------------------------------------
// Allocation-free GI pass (vpt_trace_gi_noalloc): the same kernels over a
// fixed scratch array instead of the context's per-stream scratch.
// Layout: the queue counters in the first 4 KiB, then the
// continuations and the shadow jobs (16 B per pixel each), for images of at
// most kNoAllocGiMaxPixels pixels: 64 MiB + 4 KiB.  Each GPU the library
// runs on holds its own copy until the process ends.
// ---------------------------------------------------------------------------

namespace {

constexpr size_t kNoAllocGiMaxPixels = (size_t)1 << 21;   // 2,097,152: 1920 x 1080 fits
constexpr size_t kNoAllocGiContsOff = vpt::kGiCountersBytes;
constexpr size_t kNoAllocGiJobsOff = kNoAllocGiContsOff + kNoAllocGiMaxPixels * sizeof(GiCont);
constexpr size_t kNoAllocGiBytes = kNoAllocGiJobsOff + kNoAllocGiMaxPixels * sizeof(GiJob);
static_assert(kNoAllocGiBytes == ((size_t)64 << 20) + 4096, "64 MiB + 4 KiB");
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kNoAllocGiBytes];

// The device address of `scratch` on each GPU (hipGetSymbolAddress with
// that GPU current).  In host code the array's name is only a host-side
// stand-in, never passed to a kernel.
std::mutex g_scratch_mu;
void* g_scratch_ptr[kNoAllocMaxDevices] = {};

}  // namespace

namespace vpt {

int gi_noalloc_init(int device)
{
    if (device < 0 || device >= kNoAllocMaxDevices)
        return set_error(VPT_EINVAL, "vpt_trace_gi_noalloc: device %d beyond %d", device, kNoAllocMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return VPT_OK;
    int cur = -1;
    HIP_TRY(hipGetDevice(&cur));
    if (cur != device) return set_error(VPT_EINVAL, "gi_noalloc_init: device %d is not current (%d)", device, cur);
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return set_error(VPT_EDEVICE, "gi_noalloc_init: no address for the scratch array on device %d", device);
    g_scratch_ptr[device] = p;
    return VPT_OK;
}

}  // namespace vpt

extern "C" VPT_API int vpt_trace_gi_noalloc(vpt_world* w, const vpt_camera* cam, const vpt_gbuffer* gb,
                                            const vpt_gi_params* prm, float* d_rgb, vpt_svgf_texel* d_svgf_in,
                                            void* stream)
{
    if (!w || !prm || !d_rgb) return set_error(VPT_EINVAL, "vpt_trace_gi_noalloc: null argument");
    if (int rc = vpt::check_cam(cam, "vpt_trace_gi_noalloc")) return rc;
    if (int rc = vpt::check_gbuffer(gb, cam, "vpt_trace_gi_noalloc", false)) return rc;
    if (int rc = vpt::check_iter(prm->gi_iter, "vpt_trace_gi_noalloc")) return rc;
    if (int rc = vpt::check_iter(prm->shadow_iter, "vpt_trace_gi_noalloc")) return rc;
    const size_t npx = (size_t)cam->width * cam->height;
    if (npx > kNoAllocGiMaxPixels)
        return set_error(VPT_EINVAL, "vpt_trace_gi_noalloc: %dx%d = %zu pixels exceeds its fixed scratch (%zu pixels)",
                         cam->width, cam->height, npx, kNoAllocGiMaxPixels);
    if (w->device < 0 || w->device >= kNoAllocMaxDevices)
        return set_error(VPT_EINVAL, "vpt_trace_gi_noalloc: device %d beyond %d", w->device, kNoAllocMaxDevices);
    DeviceGuard g(w->device);
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        base = g_scratch_ptr[w->device];
    }
    if (!base) {   // a context created on this device sets it up; this is the fallback
        if (int rc = vpt::gi_noalloc_init(w->device)) return rc;
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        base = g_scratch_ptr[w->device];
    }
    uint8_t* b = static_cast<uint8_t*>(base);
    const vpt::AlbedoTable albedo = vpt::take_albedo(w->ctx, prm);
    return gi_enqueue(w, cam, gb, prm, albedo, d_rgb, d_svgf_in, reinterpret_cast<uint32_t*>(b),
                      reinterpret_cast<GiCont*>(b + kNoAllocGiContsOff), reinterpret_cast<GiJob*>(b + kNoAllocGiJobsOff),
                      (hipStream_t)stream);
}

extern "C" VPT_API int vpt_trace_rays(vpt_world* w, const vpt_ray* d_rays, int n, int max_iter, vpt_hit* d_hits,
                                      void* stream)
{
    if (!w || (n > 0 && (!d_rays || !d_hits))) return set_error(VPT_EINVAL, "vpt_trace_rays: null argument");
    if (n < 0) return set_error(VPT_EINVAL, "vpt_trace_rays: negative ray count");
    if (int rc = vpt::check_iter(max_iter, "vpt_trace_rays")) return rc;
    if (n == 0) return VPT_OK;
    DeviceGuard g(w->device);
    hipLaunchKernelGGL(rays_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       vpt::make_view(w), d_rays, n, max_iter, d_hits, w->d_counter);
    HIP_TRY(hipGetLastError());
    return VPT_OK;
}

extern "C" VPT_API int vpt_trace_gi_rays(vpt_world* w, const vpt_ray* d_rays, int n, int gi_iter, const float sun_dir[3],
                                         int shadow_iter, vpt_hit* d_hits, uint8_t* d_vis, void* stream)
{
    if (!w || !sun_dir || (n > 0 && (!d_rays || !d_hits || !d_vis)))
        return set_error(VPT_EINVAL, "vpt_trace_gi_rays: null argument");
    if (n < 0) return set_error(VPT_EINVAL, "vpt_trace_gi_rays: negative ray count");
    if (int rc = vpt::check_iter(gi_iter, "vpt_trace_gi_rays")) return rc;
    if (int rc = vpt::check_iter(shadow_iter, "vpt_trace_gi_rays")) return rc;
    if (n == 0) return VPT_OK;
    DeviceGuard g(w->device);
    const WorldView wv = vpt::make_view(w);
    const dim3 grid((unsigned)((n + 255) / 256));
    hipLaunchKernelGGL(rays_kernel, grid, dim3(256), 0, (hipStream_t)stream, wv, d_rays, n, gi_iter, d_hits,
                       w->d_counter);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(gi_rays_shadow_kernel, grid, dim3(256), 0, (hipStream_t)stream, wv, d_rays, n, d_hits,
                       sun_dir[0], sun_dir[1], sun_dir[2], shadow_iter, d_vis, w->d_counter);
    HIP_TRY(hipGetLastError());
    return VPT_OK;
}
