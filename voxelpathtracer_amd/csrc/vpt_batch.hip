// Persistent wavefront tracing: one launch drains a queue of rays of any
// kind and any resident world.
//
// Every lane holds one ray.  Each loop turn the wavefront refills its idle
// lanes from a wave-private chunk of the queue (one atomicAdd per `chunk`
// rays) and advances every active ray by one iteration of the reference's
// loop (trav_prepare / trav_advance, the same code as the one-launch
// kernels, so hits are bit-identical).  A
// finished ray writes its result by kind; a diffuse-GI ray that hits a
// sun-facing block continues in the same slot as the hit's shadow ray.
//
// The frame path runs: persistent primary rays (pixel index -> ray) ->
// secondary-ray generation with wave-level compaction (ballot + popcount +
// one atomicAdd per wavefront) -> one persistent launch over the sun-shadow
// and GI rays together -> shading.  The chain reads its per-frame constants
// from device memory, so it can be captured once into a HIP graph and
// replayed.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "vpt_device.h"
#include "vpt_gi.h"
#include "vpt_internal.h"

using vpt::CamView;
using vpt::GiConst;
using vpt::set_error;
using vpt::Trav;
using vpt::Vec3;
using vpt::WorldView;

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return set_error(VPT_EDEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));      \
    } while (0)

namespace {

enum : uint32_t { KIND_QUERY = 0, KIND_SHADOW = 1, KIND_GI = 2, KIND_GI_SHADOW = 3 };

__host__ __device__ inline uint32_t make_meta(int max_iter, int world, uint32_t kind)
{
    return (uint32_t)max_iter | ((uint32_t)world << 16) | (kind << 24);
}

// Per-frame constants read by the frame kernels from device memory.
struct FrameConst {
    CamView cam;
    GiConst gi;
    int primary_iter;
    int shadow_iter;
    vpt::AlbedoTable albedo;   // this frame's, uploaded with the other constants
};

// Where finished rays write.
struct Outputs {
    vpt_hit* hits;       // KIND_QUERY: hits[dst]
    uint8_t* mask;       // KIND_SHADOW: mask[dst] = T > 0
    float* gi_t;         // KIND_GI: T
    uint16_t* gi_fb;     // KIND_GI: face | block << 8
    uint8_t* gi_vis;     // KIND_GI_SHADOW: T > 0 (0 when no shadow ray)
    float* gb_t;         // camera source: G-buffer
    uint8_t* gb_face;
    uint8_t* gb_block;
    uint32_t* gb_voxel;
    const vpt::ServeSegment* segs;   // served requests: segment table
    const uint32_t* seg_begin;       // their first ids, then the launch's ray count
    uint32_t nseg;
};

struct Slot {
    Trav r;
    uint32_t dst;
    uint32_t kind;
    int active;   // 32-bit so that the slot stays in registers
};

__device__ inline uint64_t lanemask_lt() { return (1ull << (threadIdx.x & 63)) - 1ull; }

__device__ inline void wave_count(unsigned long long* counter, unsigned long long samples, unsigned long long rays)
{
    if (!counter) return;
    for (int off = 32; off > 0; off >>= 1) {
        samples += __shfl_xor(samples, off);
        rays += __shfl_xor(rays, off);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(counter, samples);
        atomicAdd(counter + 1, rays);
    }
}

// Wave-uniform queue cursor of one wavefront.
struct Cursor {
    uint32_t next, end;
    bool exhausted;
};

// Refills the idle lanes from the wave's chunk, taking a new chunk (one
// atomicAdd) when the current one is used up.  Runs only once at least
// `refill_min` lanes are idle (or all of them), so the set-up cost of new
// rays is paid for a group of lanes at a time.
template <bool MULTI, int SRC>
__device__ __attribute__((always_inline)) inline void refill_slot(Slot& s, WorldView& wv, const WorldView& w0,
                                                                  const WorldView* __restrict__ worlds,
                                                                  uint32_t nworlds, const vpt_qray* __restrict__ rays,
                                                                  uint32_t n, uint32_t* __restrict__ head,
                                                                  uint32_t chunk, uint32_t refill_min,
                                                                  const FrameConst* __restrict__ fc, int tiles_x,
                                                                  const Outputs& out, Cursor& c,
                                                                  unsigned long long& n_rays)
{
    const int lane = threadIdx.x & 63;
    bool need = !s.active;
    uint64_t idle = __ballot(need);
    if (c.exhausted || ((uint32_t)__popcll(idle) < refill_min && idle != __ballot(true))) return;
    while (idle != 0ull && !c.exhausted) {
        if (c.next >= c.end) {
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(head, chunk);
            base = __shfl(base, 0);
            if (base >= n) {
                c.exhausted = true;
                break;
            }
            c.next = base;
            c.end = min(base + chunk, n);
        }
        const uint32_t avail = c.end - c.next;
        const uint32_t cnt = min((uint32_t)__popcll(idle), avail);
        const uint32_t rank = (uint32_t)__popcll(idle & lanemask_lt());
        if (need && rank < cnt) {
            const uint32_t id = c.next + rank;
            need = false;
            if (SRC == 0) {
                const vpt_qray q = rays[id];
                const int wi = (int)((q.meta >> 16) & 0xffu);
                if (MULTI) {
                    if ((uint32_t)wi < nworlds) {
                        wv = worlds[wi];
                    } else {
                        // no such world: an empty volume, the ray misses at once
                        wv = w0;
                        wv.xmax = wv.ymax = wv.zmax = -1.0f;
                    }
                }
                vpt::trav_init(s.r, vpt::mk3(q.origin[0], q.origin[1], q.origin[2]),
                               vpt::mk3(q.dir[0], q.dir[1], q.dir[2]), (int)(q.meta & 0xffffu));
                s.dst = q.dst;
                // caller-supplied rays (MULTI) are plain queries; the
                // continuation kinds belong to the frame path
                s.kind = MULTI ? KIND_QUERY : (q.meta >> 24);
                s.active = true;
                n_rays++;
            } else if (SRC == 2) {
                // the segment holding launch-wide id `id`: seg_begin[lo] <= id < seg_begin[lo + 1]
                uint32_t lo = 0, hi = out.nseg;
                while (hi - lo > 1u) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (out.seg_begin[mid] <= id) lo = mid;
                    else hi = mid;
                }
                const vpt::ServeSegment& sg = out.segs[lo];
                const uint32_t local = sg.first + (id - sg.begin);
                const vpt_ray q = sg.rays[local];
                if (sg.world < nworlds) {
                    wv = worlds[sg.world];
                } else {
                    wv = w0;
                    wv.xmax = wv.ymax = wv.zmax = -1.0f;
                }
                vpt::trav_init(s.r, vpt::mk3(q.origin[0], q.origin[1], q.origin[2]),
                               vpt::mk3(q.dir[0], q.dir[1], q.dir[2]), sg.max_iter);
                s.dst = local;
                s.kind = (sg.kind == VPT_REQ_GI ? KIND_GI : KIND_QUERY) | (lo << 8);
                s.active = true;
                n_rays++;
            } else {
                const uint32_t tile = id >> 6, l = id & 63u;
                const int px = (int)(tile % (uint32_t)tiles_x) * 8 + (int)(l & 7u);
                const int py = (int)(tile / (uint32_t)tiles_x) * 8 + (int)(l >> 3);
                if (px < fc->cam.W && py < fc->cam.H) {
                    Vec3 dir = vpt::primary_dir(fc->cam, px, py);
                    vpt::trav_init(s.r, vpt::mk3(fc->cam.iv[12], fc->cam.iv[13], fc->cam.iv[14]), dir,
                                   fc->primary_iter);
                    s.dst = (uint32_t)py * (uint32_t)fc->cam.W + (uint32_t)px;
                    s.kind = KIND_QUERY;
                    s.active = true;
                    n_rays++;
                }
            }
        }
        c.next += cnt;
        idle = __ballot(need);
    }
}

// A ray whose loop has ended: write its result by kind, or turn a GI hit
// into its shadow ray.
template <int SRC>
__device__ __attribute__((always_inline)) inline void finish_slot(Slot& s, const WorldView& w, uint32_t n, const Outputs& out,
                                   const FrameConst* __restrict__ fc, unsigned long long& n_samples,
                                   unsigned long long& n_rays)
{
    vpt::Hit h = vpt::trav_finish(w, s.r);
    n_samples += (unsigned long long)h.samples;
    const uint32_t dst = s.dst;
    s.active = false;
    if (SRC == 1) {
        bool intersect = h.t > 0.0f && h.block > 0;
        out.gb_t[dst] = h.t;
        out.gb_face[dst] = (uint8_t)(intersect ? h.face : 255);
        out.gb_block[dst] = (uint8_t)(intersect ? h.block : 0);
        if (out.gb_voxel) out.gb_voxel[dst] = intersect ? h.voxel : 0xffffffffu;
    } else if (SRC == 2) {
        const uint32_t seg = s.kind >> 8, k = s.kind & 0xffu;
        const vpt::ServeSegment& sg = out.segs[seg];
        if (k == KIND_GI_SHADOW) {
            sg.vis[dst] = h.t > 0.0f ? 1 : 0;
        } else {
            vpt_hit o;
            o.t = h.t;
            o.voxel = h.voxel;
            o.face = (uint8_t)h.face;
            o.block = (uint8_t)h.block;
            o.samples = (uint16_t)min(h.samples, 65535);
            sg.hits[dst] = o;
            if (k == KIND_GI) {
                const Vec3 L = vpt::mk3(sg.sun[0], sg.sun[1], sg.sun[2]);
                Vec3 so;
                float ndotl;
                if (h.t > 0.0f && h.block > 0 &&
                    vpt::gi_shadow_origin(vpt::mk3(s.r.o0x, s.r.o0y, s.r.o0z), vpt::mk3(s.r.dx, s.r.dy, s.r.dz),
                                          h.t, h.face, L, so, ndotl)) {
                    vpt::trav_init(s.r, so, L, sg.shadow_iter);
                    s.kind = KIND_GI_SHADOW | (seg << 8);
                    s.active = true;
                    n_rays++;
                } else {
                    sg.vis[dst] = 255;
                }
            }
        }
    } else if (s.kind == KIND_QUERY) {
        vpt_hit o;
        o.t = h.t;
        o.voxel = h.voxel;
        o.face = (uint8_t)h.face;
        o.block = (uint8_t)h.block;
        o.samples = (uint16_t)min(h.samples, 65535);
        if (dst < n) out.hits[dst] = o;   // dst comes from the caller: results outside [0, n) are dropped
    } else if (s.kind == KIND_SHADOW) {
        out.mask[dst] = h.t > 0.0f ? 1 : 0;
    } else if (s.kind == KIND_GI) {
        out.gi_t[dst] = h.t;
        out.gi_fb[dst] = (uint16_t)((h.face & 0xff) | (h.block << 8));
        Vec3 so;
        float ndotl;
        bool hit = h.t > 0.0f && h.block > 0;
        if (hit && vpt::gi_shadow_ray(vpt::mk3(s.r.o0x, s.r.o0y, s.r.o0z), vpt::mk3(s.r.dx, s.r.dy, s.r.dz), h.t,
                                      h.face, fc->gi, so, ndotl)) {
            vpt::trav_init(s.r, so, vpt::mk3(fc->gi.sun[0], fc->gi.sun[1], fc->gi.sun[2]), fc->gi.shadow_iter);
            s.kind = KIND_GI_SHADOW;
            s.active = true;
            n_rays++;
        } else {
            out.gi_vis[dst] = 0;
        }
    } else {
        out.gi_vis[dst] = h.t > 0.0f ? 1 : 0;
    }
}

// SRC = 0: queue of vpt_qray (rays[id]); SRC = 1: primary ray of the id-th
// pixel in 8x8 tile order (fc->cam); SRC = 2: served requests' rays through
// the segment table in `out`.
template <bool MULTI, int SRC>
__global__ void __launch_bounds__(256) persistent_kernel(WorldView w0, const WorldView* __restrict__ worlds,
                                                         uint32_t nworlds, const vpt_qray* __restrict__ rays,
                                                         const uint32_t* __restrict__ d_n, uint32_t n_host,
                                                         uint32_t* __restrict__ head, uint32_t chunk,
                                                         uint32_t refill_min, Outputs out,
                                                         const FrameConst* __restrict__ fc,
                                                         unsigned long long* __restrict__ counter)
{
    const uint32_t n = d_n ? *d_n : n_host;
    Slot s;
    s.active = false;
    WorldView wv = w0;
    Cursor c;
    c.next = 0;
    c.end = 0;
    c.exhausted = false;
    unsigned long long n_samples = 0, n_rays = 0;
    const int tiles_x = SRC == 1 ? (fc->cam.W + 7) / 8 : 1;

    while (true) {
        refill_slot<MULTI, SRC>(s, wv, w0, worlds, nworlds, rays, n, head, chunk, refill_min, fc, tiles_x, out, c,
                                n_rays);
        if (!__any(s.active)) {
            if (c.exhausted) break;
            continue;
        }
        // one iteration of the lane's ray (a single-world launch reads the
        // world from the kernel arguments)
        const WorldView& w = (MULTI || SRC == 2) ? wv : w0;
        uint32_t off = 0;
        vpt::Cell cell{};
        const bool ld = s.active && vpt::trav_prepare(w, s.r, off, cell);
        const int m = ld ? (int)vpt::plane_byte(w.df, off) : 0;
        if (s.active && !(ld && vpt::trav_advance(s.r, m, cell)))
            finish_slot<SRC>(s, w, n, out, fc, n_samples, n_rays);
    }
    wave_count(counter, n_samples, n_rays);
}


// Served requests, one lane per ray (the launch-wide id selects the
// segment); a GI bounce that hits a sun-facing block traces its shadow ray
// in the same lane.  Same traversal code as every other path.
__global__ void __launch_bounds__(256) serve_direct_kernel(const WorldView* __restrict__ worlds, uint32_t nworlds,
                                                           Outputs out, uint32_t total)
{
    const uint32_t id = blockIdx.x * 256u + threadIdx.x;
    if (id >= total) return;
    uint32_t lo = 0, hi = out.nseg;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (out.seg_begin[mid] <= id) lo = mid;
        else hi = mid;
    }
    const vpt::ServeSegment& sg = out.segs[lo];
    const uint32_t local = sg.first + (id - sg.begin);
    const vpt_ray q = sg.rays[local];
    WorldView wv;
    if (sg.world < nworlds) {
        wv = worlds[sg.world];
    } else {
        memset(&wv, 0, sizeof(wv));
        wv.xmax = wv.ymax = wv.zmax = -1.0f;
    }
    const Vec3 o = vpt::mk3(q.origin[0], q.origin[1], q.origin[2]);
    const Vec3 d = vpt::mk3(q.dir[0], q.dir[1], q.dir[2]);
    const vpt::Hit h = vpt::traverse(wv, o, d, sg.max_iter);
    vpt_hit r;
    r.t = h.t;
    r.voxel = h.voxel;
    r.face = (uint8_t)h.face;
    r.block = (uint8_t)h.block;
    r.samples = (uint16_t)min(h.samples, 65535);
    sg.hits[local] = r;
    if (sg.kind == VPT_REQ_GI) {
        const Vec3 L = vpt::mk3(sg.sun[0], sg.sun[1], sg.sun[2]);
        Vec3 so;
        float ndotl;
        uint8_t v = 255;
        if (h.t > 0.0f && h.block > 0 && vpt::gi_shadow_origin(o, d, h.t, h.face, L, so, ndotl))
            v = vpt::traverse(wv, so, L, sg.shadow_iter).t > 0.0f ? 1 : 0;
        sg.vis[local] = v;
    }
}

__global__ void reset_kernel(uint32_t* counters, int n)
{
    if ((int)threadIdx.x < n) counters[threadIdx.x] = 0u;
}

// Secondary-ray generation (one lane per pixel, 16x16 tiles): the sun-shadow
// ray of ShadowRayTraceFrag.glsl:414-513 and the GI ray of
// DiffuseRayTraceFrag.glsl:541-549, compacted into one queue.
__global__ void __launch_bounds__(256) gen_kernel(WorldView w, const FrameConst* __restrict__ fc,
                                                  const float* __restrict__ gb_t, const uint8_t* __restrict__ gb_face,
                                                  uint8_t* __restrict__ mask, vpt_qray* __restrict__ queue,
                                                  uint32_t* __restrict__ qlen)
{
    const CamView& cam = fc->cam;
    const uint32_t tiles_x = (uint32_t)(cam.W + 15) / 16u;
    const uint32_t tiles_y = (uint32_t)(cam.H + 15) / 16u;
    const uint32_t tile = vpt::xcd_remap(blockIdx.x, tiles_x * tiles_y);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int px = (int)(tile % tiles_x) * 16 + (wave & 1) * 8 + (lane & 7);
    const int py = (int)(tile / tiles_x) * 16 + (wave >> 1) * 8 + (lane >> 3);
    bool emit_s = false, emit_g = false;
    vpt_qray rs, rg;
    if (px < cam.W && py < cam.H) {
        const uint32_t i = (uint32_t)py * (uint32_t)cam.W + (uint32_t)px;
        const float dist = gb_t[i];
        const int face = gb_face[i];
        const Vec3 L = vpt::mk3(fc->gi.sun[0], fc->gi.sun[1], fc->gi.sun[2]);
        Vec3 o;
        uint8_t m;
        if (vpt::shadow_setup(w, cam, px, py, dist, face, L, o, m)) {
            rs.origin[0] = o.x;
            rs.origin[1] = o.y;
            rs.origin[2] = o.z;
            rs.dir[0] = L.x;
            rs.dir[1] = L.y;
            rs.dir[2] = L.z;
            rs.meta = make_meta(fc->shadow_iter, 0, KIND_SHADOW);
            rs.dst = i;
            emit_s = true;
        } else {
            mask[i] = m;
        }
        if (!(dist < 0.0f)) {
            Vec3 go, gd;
            vpt::gi_ray(cam, px, py, dist, vpt::face_normal(face, 0.5f), fc->gi.frame, go, gd);
            rg.origin[0] = go.x;
            rg.origin[1] = go.y;
            rg.origin[2] = go.z;
            rg.dir[0] = gd.x;
            rg.dir[1] = gd.y;
            rg.dir[2] = gd.z;
            rg.meta = make_meta(fc->gi.gi_iter, 0, KIND_GI);
            rg.dst = i;
            emit_g = true;
        }
    }
    // compaction: wave counts by ballot + popcount, one atomicAdd per
    // workgroup (a single queue counter shared by every wave serialises)
    __shared__ uint32_t wave_total[4], group_base;
    const uint64_t bs = __ballot(emit_s), bg = __ballot(emit_g);
    const uint32_t ns = (uint32_t)__popcll(bs), total = ns + (uint32_t)__popcll(bg);
    if (lane == 0) wave_total[wave] = total;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t all = (wave_total[0] + wave_total[1]) + (wave_total[2] + wave_total[3]);
        group_base = all ? atomicAdd(qlen, all) : 0u;
    }
    __syncthreads();
    uint32_t base = group_base;
    for (int v = 0; v < wave; v++) base += wave_total[v];
    if (emit_s) queue[base + (uint32_t)__popcll(bs & lanemask_lt())] = rs;
    if (emit_g) queue[base + ns + (uint32_t)__popcll(bg & lanemask_lt())] = rg;
}

// Shading (DiffuseRayTraceFrag.glsl main :822-935 after the traces).
__global__ void __launch_bounds__(256) shade_kernel(const FrameConst* __restrict__ fc, const float* __restrict__ gb_t,
                                                    const uint8_t* __restrict__ gb_face,
                                                    const float* __restrict__ gi_t,
                                                    const uint16_t* __restrict__ gi_fb,
                                                    const uint8_t* __restrict__ gi_vis, float* __restrict__ rgb,
                                                    vpt_svgf_texel* __restrict__ texels)
{
    const CamView& cam = fc->cam;
    const uint32_t tiles_x = (uint32_t)(cam.W + 15) / 16u;
    const uint32_t tiles_y = (uint32_t)(cam.H + 15) / 16u;
    const uint32_t tile = vpt::xcd_remap(blockIdx.x, tiles_x * tiles_y);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int px = (int)(tile % tiles_x) * 16 + (wave & 1) * 8 + (lane & 7);
    const int py = (int)(tile / tiles_x) * 16 + (wave >> 1) * 8 + (lane >> 3);
    if (px >= cam.W || py >= cam.H) return;
    const uint32_t i = (uint32_t)py * (uint32_t)cam.W + (uint32_t)px;
    const float dist = gb_t[i];
    const Vec3 N = vpt::face_normal(gb_face[i], 0.5f);
    vpt_svgf_texel* tx = texels ? texels + i : nullptr;
    if (dist < 0.0f) {
        vpt::gi_write_sky_pixel(fc->gi, N, rgb + 3 * (size_t)i, tx);
        return;
    }
    Vec3 o, dir;
    vpt::gi_ray(cam, px, py, dist, N, fc->gi.frame, o, dir);
    const float T = gi_t[i];
    const uint32_t fb = gi_fb[i];
    const int hit_face = (int)(fb & 0xffu), hit_block = (int)(fb >> 8);
    const bool hit = T > 0.0f && hit_block > 0;
    const float shadow_at = hit ? (float)gi_vis[i] : 0.0f;
    vpt::gi_write_hit_pixel(fc->gi, fc->albedo.v, dir, T, hit, hit_face, hit_block, shadow_at, rgb + 3 * (size_t)i,
                            tx);
}

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

int cu_count(int device)
{
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device) != hipSuccess) return 256;
    return p.multiProcessorCount > 0 ? p.multiProcessorCount : 256;
}

bool cfg_valid(const vpt_launch_cfg& c)
{
    return (c.block == 64 || c.block == 128 || c.block == 256) && c.waves_per_cu >= 1 && c.waves_per_cu <= 64 &&
           c.chunk >= 1 && c.chunk <= 65536 && c.refill >= 1 && c.refill <= 64 && (c.flags & ~3) == 0;
}

vpt_launch_cfg default_cfg() { return vpt_launch_cfg{256, 16, 256, 32, 1}; }

// The autotuner's search space: workgroup size x resident wavefronts per CU
// x queue chunk per atomic x refill threshold (x primary-ray kernel for the
// frame path).
std::vector<vpt_launch_cfg> candidates(bool frame)
{
    std::vector<vpt_launch_cfg> c;
    for (int fl : {0, 1}) {
        if (fl && !frame) continue;
        for (int b : {64, 256})
            for (int wpc : {8, 16, 32})
                for (int ch : {128, 1024})
                    for (int rf : {8, 32, 64}) c.push_back(vpt_launch_cfg{b, wpc, ch, rf, fl});
    }
    return c;
}

// Primary rays by the one-launch tile kernel (16x16 tiles of four 8x8
// wavefront squares) with the camera read from the frame constants.
__global__ void __launch_bounds__(256) frame_primary_kernel(WorldView w, const FrameConst* __restrict__ fc, Outputs out,
                                                            unsigned long long* __restrict__ counter)
{
    const CamView& cam = fc->cam;
    const uint32_t tiles_x = (uint32_t)(cam.W + 15) / 16u;
    const uint32_t tiles_y = (uint32_t)(cam.H + 15) / 16u;
    const uint32_t tile = vpt::xcd_remap(blockIdx.x, tiles_x * tiles_y);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int px = (int)(tile % tiles_x) * 16 + (wave & 1) * 8 + (lane & 7);
    const int py = (int)(tile / tiles_x) * 16 + (wave >> 1) * 8 + (lane >> 3);
    unsigned long long samples = 0, rays = 0;
    if (px < cam.W && py < cam.H) {
        Vec3 dir = vpt::primary_dir(cam, px, py);
        vpt::Hit h = vpt::traverse(w, vpt::mk3(cam.iv[12], cam.iv[13], cam.iv[14]), dir, fc->primary_iter);
        const uint32_t i = (uint32_t)py * (uint32_t)cam.W + (uint32_t)px;
        bool intersect = h.t > 0.0f && h.block > 0;
        out.gb_t[i] = h.t;
        out.gb_face[i] = (uint8_t)(intersect ? h.face : 255);
        out.gb_block[i] = (uint8_t)(intersect ? h.block : 0);
        if (out.gb_voxel) out.gb_voxel[i] = intersect ? h.voxel : 0xffffffffu;
        samples = (unsigned long long)h.samples;
        rays = 1;
    }
    wave_count(counter, samples, rays);
}

// Launches the persistent kernel variant for (MULTI, SRC) with shape cfg.
template <bool MULTI, int SRC>
int launch_persistent(const vpt_launch_cfg& cfg, int cus, hipStream_t s, WorldView w0, const WorldView* worlds,
                      uint32_t nworlds, const vpt_qray* rays, const uint32_t* d_n, uint32_t n_host, uint32_t* head,
                      const Outputs& out, const FrameConst* fc, unsigned long long* counter)
{
    const int waves_per_block = cfg.block / 64;
    const int blocks = std::max(1, cfg.waves_per_cu * cus / waves_per_block);
    hipLaunchKernelGGL((persistent_kernel<MULTI, SRC>), dim3(blocks), dim3(cfg.block), 0, s, w0, worlds, nworlds,
                       rays, d_n, n_host, head, (uint32_t)cfg.chunk, (uint32_t)cfg.refill, out, fc, counter);
    HIP_TRY(hipGetLastError());
    return VPT_OK;
}

}  // namespace

int vpt::device_cus(int device) { return cu_count(device); }

bool vpt::launch_cfg_valid(const vpt_launch_cfg& cfg) { return cfg_valid(cfg); }

int vpt::serve_launch(int cus, const vpt_launch_cfg& cfg, const void* d_worlds, uint32_t nworlds,
                      const ServeSegment* d_segs, const uint32_t* d_begin, uint32_t nseg, uint32_t total,
                      uint32_t* d_head, void* stream)
{
    if (!cfg_valid(cfg) || nseg == 0 || total == 0) return set_error(VPT_EINVAL, "serve_launch: bad launch");
    hipStream_t s = (hipStream_t)stream;
    Outputs out;
    memset(&out, 0, sizeof(out));
    out.segs = d_segs;
    out.seg_begin = d_begin;
    out.nseg = nseg;
    if (cfg.flags & 2) {
        hipLaunchKernelGGL(serve_direct_kernel, dim3((total + 255) / 256), dim3(256), 0, s, (const WorldView*)d_worlds,
                           nworlds, out, total);
        HIP_TRY(hipGetLastError());
        return VPT_OK;
    }
    hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(64), 0, s, d_head, 1);
    HIP_TRY(hipGetLastError());
    // a launch smaller than the resident lanes gives each wavefront at most
    // one ray per lane, so no ray waits behind another in its lane
    vpt_launch_cfg c = cfg;
    const uint64_t waves = (uint64_t)std::max(1, c.waves_per_cu * cus / (c.block / 64)) * (c.block / 64);
    const uint64_t per_wave = (total + waves - 1) / waves;
    c.chunk = (int)std::min<uint64_t>((uint64_t)c.chunk, std::max<uint64_t>(64, (per_wave + 63) / 64 * 64));
    WorldView w0;
    memset(&w0, 0, sizeof(w0));
    return launch_persistent<false, 2>(c, cus, s, w0, (const WorldView*)d_worlds, nworlds, nullptr, nullptr, total,
                                       d_head, out, nullptr, nullptr);
}

// ===========================================================================
// Batch (ray queries over resident worlds)
// ===========================================================================

struct vpt_batch {
    int device = 0;
    int cus = 256;
    int max_worlds = 0;
    std::vector<vpt_world*> worlds;
    WorldView* d_worlds = nullptr;
    uint32_t* d_head = nullptr;
    vpt_launch_cfg cfg = default_cfg();
    bool table_dirty = true;
};

extern "C" VPT_API int vpt_batch_create(vpt_ctx* ctx, int max_worlds, vpt_batch** out)
{
    if (!ctx || !out) return set_error(VPT_EINVAL, "vpt_batch_create: null argument");
    if (max_worlds < 1 || max_worlds > 256) return set_error(VPT_EINVAL, "vpt_batch_create: max_worlds in [1, 256]");
    DeviceGuard g(ctx->device);
    vpt_batch* b = new vpt_batch();
    b->device = ctx->device;
    b->cus = cu_count(ctx->device);
    b->max_worlds = max_worlds;
    b->worlds.assign(max_worlds, nullptr);
    if (hipMalloc(&b->d_worlds, sizeof(WorldView) * max_worlds) != hipSuccess ||
        hipMalloc(&b->d_head, sizeof(uint32_t) * 4) != hipSuccess) {
        (void)hipFree(b->d_worlds);
        delete b;
        return set_error(VPT_ENOMEM, "vpt_batch_create: hipMalloc failed");
    }
    *out = b;
    return VPT_OK;
}

extern "C" VPT_API int vpt_batch_destroy(vpt_batch* b)
{
    if (!b) return VPT_OK;
    DeviceGuard g(b->device);
    (void)hipFree(b->d_worlds);
    (void)hipFree(b->d_head);
    delete b;
    return VPT_OK;
}

extern "C" VPT_API int vpt_batch_set_world(vpt_batch* b, int slot, vpt_world* w)
{
    if (!b) return set_error(VPT_EINVAL, "vpt_batch_set_world: null batch");
    if (slot < 0 || slot >= b->max_worlds) return set_error(VPT_EINVAL, "vpt_batch_set_world: slot %d out of range", slot);
    if (w && w->device != b->device) return set_error(VPT_EINVAL, "vpt_batch_set_world: world on another device");
    b->worlds[slot] = w;
    b->table_dirty = true;
    return VPT_OK;
}

extern "C" VPT_API int vpt_batch_get_cfg(const vpt_batch* b, vpt_launch_cfg* cfg)
{
    if (!b || !cfg) return set_error(VPT_EINVAL, "vpt_batch_get_cfg: null argument");
    *cfg = b->cfg;
    return VPT_OK;
}

extern "C" VPT_API int vpt_batch_set_cfg(vpt_batch* b, const vpt_launch_cfg* cfg)
{
    if (!b || !cfg) return set_error(VPT_EINVAL, "vpt_batch_set_cfg: null argument");
    if (!cfg_valid(*cfg)) return set_error(VPT_EINVAL, "vpt_batch_set_cfg: invalid launch shape");
    b->cfg = *cfg;
    return VPT_OK;
}

static int batch_sync_table(vpt_batch* b, hipStream_t s)
{
    if (!b->table_dirty) return VPT_OK;
    std::vector<WorldView> host(b->max_worlds);
    for (int i = 0; i < b->max_worlds; i++) {
        if (b->worlds[i]) {
            host[i] = vpt::make_view(b->worlds[i]);
        } else {
            // an empty slot: a 0-voxel volume, every ray leaves it at once
            memset(&host[i], 0, sizeof(WorldView));
            host[i].xmax = host[i].ymax = host[i].zmax = -1.0f;
        }
    }
    HIP_TRY(hipMemcpyAsync(b->d_worlds, host.data(), sizeof(WorldView) * b->max_worlds, hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    b->table_dirty = false;
    return VPT_OK;
}

static int batch_launch(vpt_batch* b, const vpt_launch_cfg& cfg, const vpt_qray* d_rays, int n, vpt_hit* d_hits,
                        hipStream_t s)
{
    hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(64), 0, s, b->d_head, 4);
    HIP_TRY(hipGetLastError());
    Outputs out;
    memset(&out, 0, sizeof(out));
    out.hits = d_hits;
    WorldView w0;
    memset(&w0, 0, sizeof(w0));
    return launch_persistent<true, 0>(cfg, b->cus, s, w0, b->d_worlds, (uint32_t)b->max_worlds, d_rays, nullptr,
                                      (uint32_t)n, b->d_head, out, nullptr, nullptr);
}

extern "C" VPT_API int vpt_batch_trace(vpt_batch* b, const vpt_qray* d_rays, int n, vpt_hit* d_hits, void* stream)
{
    if (!b) return set_error(VPT_EINVAL, "vpt_batch_trace: null batch");
    if (n < 0 || (n > 0 && (!d_rays || !d_hits))) return set_error(VPT_EINVAL, "vpt_batch_trace: bad ray list");
    if (n == 0) return VPT_OK;
    DeviceGuard g(b->device);
    hipStream_t s = (hipStream_t)stream;
    if (int rc = batch_sync_table(b, s)) return rc;
    return batch_launch(b, b->cfg, d_rays, n, d_hits, s);
}

extern "C" VPT_API int vpt_batch_autotune(vpt_batch* b, const vpt_qray* d_rays, int n, vpt_hit* d_hits, float* best_ms,
                                          void* stream)
{
    if (!b || n <= 0 || !d_rays || !d_hits) return set_error(VPT_EINVAL, "vpt_batch_autotune: bad arguments");
    DeviceGuard g(b->device);
    hipStream_t s = (hipStream_t)stream;
    if (int rc = batch_sync_table(b, s)) return rc;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    float best = 1e30f;
    vpt_launch_cfg best_cfg = b->cfg;
    for (const vpt_launch_cfg& c : candidates(false)) {
        if (int rc = batch_launch(b, c, d_rays, n, d_hits, s)) return rc;  // warm-up
        HIP_TRY(hipEventRecord(e0, s));
        for (int r = 0; r < 3; r++)
            if (int rc = batch_launch(b, c, d_rays, n, d_hits, s)) return rc;
        HIP_TRY(hipEventRecord(e1, s));
        HIP_TRY(hipEventSynchronize(e1));
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        ms /= 3.0f;
        if (ms < best) {
            best = ms;
            best_cfg = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->cfg = best_cfg;
    if (best_ms) *best_ms = best;
    return VPT_OK;
}

// ===========================================================================
// Frame path
// ===========================================================================

struct vpt_frame {
    int device = 0;
    int cus = 256;
    int W = 0, H = 0;
    vpt_launch_cfg cfg = default_cfg();
    FrameConst* d_fc = nullptr;
    uint32_t* d_counters = nullptr;   // [0] primary head, [1] queue head, [2] queue length
    vpt_qray* d_queue = nullptr;      // 2 rays per pixel at most
    float* d_gi_t = nullptr;
    uint16_t* d_gi_fb = nullptr;
    uint8_t* d_gi_vis = nullptr;
    // captured graph and the buffers it was captured with
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    uintptr_t key[10] = {};
    // graphs are captured and replayed on an internal stream (the caller's
    // may be the legacy null stream, which cannot be captured), ordered
    // against the caller's stream with events
    hipStream_t cap = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
};

extern "C" VPT_API int vpt_frame_destroy(vpt_frame* f)
{
    if (!f) return VPT_OK;
    DeviceGuard g(f->device);
    if (f->exec) (void)hipGraphExecDestroy(f->exec);
    if (f->graph) (void)hipGraphDestroy(f->graph);
    if (f->ev_in) (void)hipEventDestroy(f->ev_in);
    if (f->ev_out) (void)hipEventDestroy(f->ev_out);
    if (f->cap) (void)hipStreamDestroy(f->cap);
    (void)hipFree(f->d_fc);
    (void)hipFree(f->d_counters);
    (void)hipFree(f->d_queue);
    (void)hipFree(f->d_gi_t);
    (void)hipFree(f->d_gi_fb);
    (void)hipFree(f->d_gi_vis);
    delete f;
    return VPT_OK;
}

extern "C" VPT_API int vpt_frame_create(vpt_ctx* ctx, int width, int height, vpt_frame** out)
{
    if (!ctx || !out) return set_error(VPT_EINVAL, "vpt_frame_create: null argument");
    if (width <= 0 || height <= 0 || width > 16384 || height > 16384)
        return set_error(VPT_EINVAL, "vpt_frame_create: bad size %dx%d", width, height);
    DeviceGuard g(ctx->device);
    vpt_frame* f = new vpt_frame();
    f->device = ctx->device;
    f->cus = cu_count(ctx->device);
    f->W = width;
    f->H = height;
    size_t px = (size_t)width * height;
    hipError_t e = hipMalloc(&f->d_fc, sizeof(FrameConst));
    if (e == hipSuccess) e = hipMalloc(&f->d_counters, sizeof(uint32_t) * 4);
    if (e == hipSuccess) e = hipMalloc(&f->d_queue, sizeof(vpt_qray) * 2 * px);
    if (e == hipSuccess) e = hipMalloc(&f->d_gi_t, sizeof(float) * px);
    if (e == hipSuccess) e = hipMalloc(&f->d_gi_fb, sizeof(uint16_t) * px);
    if (e == hipSuccess) e = hipMalloc(&f->d_gi_vis, px);
    if (e != hipSuccess) {
        vpt_frame_destroy(f);
        return set_error(VPT_ENOMEM, "vpt_frame_create: hipMalloc failed (%s)", hipGetErrorString(e));
    }
    *out = f;
    return VPT_OK;
}

extern "C" VPT_API int vpt_frame_set_cfg(vpt_frame* f, const vpt_launch_cfg* cfg)
{
    if (!f || !cfg) return set_error(VPT_EINVAL, "vpt_frame_set_cfg: null argument");
    if (!cfg_valid(*cfg)) return set_error(VPT_EINVAL, "vpt_frame_set_cfg: invalid launch shape");
    f->cfg = *cfg;
    if (f->exec) {
        (void)hipGraphExecDestroy(f->exec);
        (void)hipGraphDestroy(f->graph);
        f->exec = nullptr;
        f->graph = nullptr;
    }
    return VPT_OK;
}

extern "C" VPT_API int vpt_frame_get_cfg(const vpt_frame* f, vpt_launch_cfg* cfg)
{
    if (!f || !cfg) return set_error(VPT_EINVAL, "vpt_frame_get_cfg: null argument");
    *cfg = f->cfg;
    return VPT_OK;
}

// Enqueues the frame's kernels (everything after the constants upload).
static int frame_enqueue(vpt_frame* f, const vpt_launch_cfg& cfg, vpt_world* w, const vpt_gbuffer* gb, uint8_t* d_mask,
                         float* d_rgb, vpt_svgf_texel* d_tex, hipStream_t s)
{
    WorldView wv = vpt::make_view(w);
    hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(64), 0, s, f->d_counters, 4);
    HIP_TRY(hipGetLastError());
    Outputs out;
    memset(&out, 0, sizeof(out));
    out.gb_t = gb->d_t;
    out.gb_face = gb->d_face;
    out.gb_block = gb->d_block;
    out.gb_voxel = gb->d_voxel;
    const dim3 tiles((unsigned)(((f->W + 15) / 16) * ((f->H + 15) / 16)));
    if (cfg.flags & 1) {
        hipLaunchKernelGGL(frame_primary_kernel, tiles, dim3(256), 0, s, wv, f->d_fc, out, w->d_counter);
        HIP_TRY(hipGetLastError());
    } else {
        const uint32_t n_primary = (uint32_t)(((f->W + 7) / 8) * ((f->H + 7) / 8) * 64);
        if (int rc = launch_persistent<false, 1>(cfg, f->cus, s, wv, nullptr, 0, nullptr, nullptr, n_primary,
                                                 f->d_counters + 0, out, f->d_fc, w->d_counter))
            return rc;
    }
    hipLaunchKernelGGL(gen_kernel, tiles, dim3(256), 0, s, wv, f->d_fc, gb->d_t, gb->d_face, d_mask, f->d_queue,
                       f->d_counters + 2);
    HIP_TRY(hipGetLastError());
    Outputs o2;
    memset(&o2, 0, sizeof(o2));
    o2.mask = d_mask;
    o2.gi_t = f->d_gi_t;
    o2.gi_fb = f->d_gi_fb;
    o2.gi_vis = f->d_gi_vis;
    if (int rc = launch_persistent<false, 0>(cfg, f->cus, s, wv, nullptr, 0, f->d_queue, f->d_counters + 2, 0,
                                             f->d_counters + 1, o2, f->d_fc, w->d_counter))
        return rc;
    hipLaunchKernelGGL(shade_kernel, tiles, dim3(256), 0, s, f->d_fc, gb->d_t, gb->d_face, f->d_gi_t, f->d_gi_fb,
                       f->d_gi_vis, d_rgb, d_tex);
    HIP_TRY(hipGetLastError());
    return VPT_OK;
}

static int frame_check(vpt_frame* f, vpt_world* w, const vpt_camera* cam, const vpt_gi_params* prm, int primary_iter,
                       int shadow_iter, const vpt_gbuffer* gb, uint8_t* d_mask, float* d_rgb, const char* fn)
{
    if (!f || !w || !prm || !d_mask || !d_rgb) return set_error(VPT_EINVAL, "%s: null argument", fn);
    if (int rc = vpt::check_cam(cam, fn)) return rc;
    if (cam->width != f->W || cam->height != f->H)
        return set_error(VPT_EINVAL, "%s: camera %dx%d does not match the frame %dx%d", fn, cam->width, cam->height,
                         f->W, f->H);
    if (int rc = vpt::check_gbuffer(gb, cam, fn, true)) return rc;
    if (int rc = vpt::check_iter(primary_iter, fn)) return rc;
    if (int rc = vpt::check_iter(shadow_iter, fn)) return rc;
    if (int rc = vpt::check_iter(prm->gi_iter, fn)) return rc;
    if (int rc = vpt::check_iter(prm->shadow_iter, fn)) return rc;
    if (w->device != f->device) return set_error(VPT_EINVAL, "%s: world on another device", fn);
    return VPT_OK;
}

static int frame_upload_consts(vpt_frame* f, vpt_world* w, const vpt_camera* cam, const vpt_gi_params* prm,
                               int primary_iter, int shadow_iter, hipStream_t s)
{
    FrameConst fc;
    fc.cam = vpt::make_cam(cam);
    fc.gi = vpt::make_gi(prm);
    fc.primary_iter = primary_iter;
    fc.shadow_iter = shadow_iter;
    fc.albedo = vpt::take_albedo(w->ctx, prm);
    // pageable source: the copy is staged before the call returns
    HIP_TRY(hipMemcpyAsync(f->d_fc, &fc, sizeof(fc), hipMemcpyHostToDevice, s));
    return VPT_OK;
}

extern "C" VPT_API int vpt_frame_trace(vpt_frame* f, vpt_world* w, const vpt_camera* cam, const vpt_gi_params* prm,
                                       int primary_iter, int shadow_iter, const vpt_gbuffer* gb, uint8_t* d_mask,
                                       float* d_rgb, vpt_svgf_texel* d_svgf_in, int use_graph, void* stream)
{
    if (int rc = frame_check(f, w, cam, prm, primary_iter, shadow_iter, gb, d_mask, d_rgb, "vpt_frame_trace"))
        return rc;
    DeviceGuard g(f->device);
    hipStream_t s = (hipStream_t)stream;
    if (int rc = frame_upload_consts(f, w, cam, prm, primary_iter, shadow_iter, s)) return rc;
    if (!use_graph) return frame_enqueue(f, f->cfg, w, gb, d_mask, d_rgb, d_svgf_in, s);

    const uintptr_t key[10] = {(uintptr_t)w,          (uintptr_t)w->d_df,    (uintptr_t)w->d_counter,
                               (uintptr_t)gb->d_t,    (uintptr_t)gb->d_face, (uintptr_t)gb->d_block,
                               (uintptr_t)gb->d_voxel, (uintptr_t)d_mask,    (uintptr_t)d_rgb,
                               (uintptr_t)d_svgf_in};
    if (!f->cap) {
        HIP_TRY(hipStreamCreateWithFlags(&f->cap, hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&f->ev_in, hipEventDisableTiming));
        HIP_TRY(hipEventCreateWithFlags(&f->ev_out, hipEventDisableTiming));
    }
    HIP_TRY(hipEventRecord(f->ev_in, s));
    HIP_TRY(hipStreamWaitEvent(f->cap, f->ev_in, 0));
    if (!f->exec || memcmp(key, f->key, sizeof(key)) != 0) {
        if (f->exec) {
            (void)hipGraphExecDestroy(f->exec);
            (void)hipGraphDestroy(f->graph);
            f->exec = nullptr;
            f->graph = nullptr;
        }
        HIP_TRY(hipStreamBeginCapture(f->cap, hipStreamCaptureModeThreadLocal));
        int rc = frame_enqueue(f, f->cfg, w, gb, d_mask, d_rgb, d_svgf_in, f->cap);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(f->cap, &graph);
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) return set_error(VPT_EDEVICE, "vpt_frame_trace: capture failed: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&f->exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(graph);
            return set_error(VPT_EDEVICE, "vpt_frame_trace: graph instantiation failed: %s", hipGetErrorString(e));
        }
        f->graph = graph;
        memcpy(f->key, key, sizeof(key));
    }
    HIP_TRY(hipGraphLaunch(f->exec, f->cap));
    HIP_TRY(hipEventRecord(f->ev_out, f->cap));
    HIP_TRY(hipStreamWaitEvent(s, f->ev_out, 0));
    return VPT_OK;
}

extern "C" VPT_API int vpt_frame_autotune(vpt_frame* f, vpt_world* w, const vpt_camera* cam, const vpt_gi_params* prm,
                                          int primary_iter, int shadow_iter, const vpt_gbuffer* gb, uint8_t* d_mask,
                                          float* d_rgb, vpt_svgf_texel* d_svgf_in, float* best_ms, void* stream)
{
    if (int rc = frame_check(f, w, cam, prm, primary_iter, shadow_iter, gb, d_mask, d_rgb, "vpt_frame_autotune"))
        return rc;
    DeviceGuard g(f->device);
    hipStream_t s = (hipStream_t)stream;
    if (int rc = frame_upload_consts(f, w, cam, prm, primary_iter, shadow_iter, s)) return rc;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    float best = 1e30f;
    vpt_launch_cfg best_cfg = f->cfg;
    for (const vpt_launch_cfg& c : candidates(true)) {
        if (int rc = frame_enqueue(f, c, w, gb, d_mask, d_rgb, d_svgf_in, s)) return rc;
        HIP_TRY(hipEventRecord(e0, s));
        for (int r = 0; r < 3; r++)
            if (int rc = frame_enqueue(f, c, w, gb, d_mask, d_rgb, d_svgf_in, s)) return rc;
        HIP_TRY(hipEventRecord(e1, s));
        HIP_TRY(hipEventSynchronize(e1));
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        ms /= 3.0f;
        if (ms < best) {
            best = ms;
            best_cfg = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (int rc = vpt_frame_set_cfg(f, &best_cfg)) return rc;
    if (best_ms) *best_ms = best;
    return VPT_OK;
}
