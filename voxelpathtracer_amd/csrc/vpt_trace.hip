// Ray kernels: primary (G-buffer), sun shadow, diffuse GI, ray queries.
//
// One lane per ray.  A 256-lane workgroup covers a 16x16 pixel tile made of
// four 8x8 wavefront squares (square footprints keep a wavefront's rays on
// neighbouring bricks), and tiles are dealt to XCDs in contiguous screen
// bands (xcd_remap) so each XCD's L2 holds one region of the world.
#include <hip/hip_runtime.h>
#include <string.h>

#include "vpt_device.h"
#include "vpt_internal.h"

using vpt::set_error;
using vpt::Vec3;
using vpt::WorldView;

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return set_error(VPT_EDEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));      \
    } while (0)

namespace {

struct CamView {
    float iv[16];
    float ip[16];
    int W, H;
};

struct GiConst {
    float sun[3];
    float light[3];
    float sky[3];
    float sky_pixel[3];
    uint32_t frame;
    int gi_iter;
    int shadow_iter;
};

WorldView make_view(const vpt_world* w)
{
    WorldView v;
    v.grid = w->d_grid;
    v.df = w->d_df;
    v.nx = w->nx;
    v.ny = w->ny;
    v.nz = w->nz;
    v.row_stride = (uint32_t)w->nbx * vpt::kBrickBytes;
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
                                                      unsigned long long* __restrict__ counter)
{
    int px, py;
    tile_pixel(cam.W, cam.H, px, py);
    int samples = 0, rays = 0;
    if (px < cam.W && py < cam.H) {
        Vec3 dir = vpt::normalize3(vpt::ray_dir_at(cam.iv, cam.ip, vpt::texcoord(px, cam.W), vpt::texcoord(py, cam.H)));
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

// G-buffer world position: GetPositionAt (ShadowRayTraceFrag.glsl:299-303)
// at the pixel centre.
__device__ inline Vec3 gbuffer_position(const CamView& cam, int px, int py, float dist)
{
    Vec3 dir = vpt::normalize3(vpt::ray_dir_at(cam.iv, cam.ip, vpt::texcoord(px, cam.W), vpt::texcoord(py, cam.H)));
    return vpt::mk3(cam.iv[12] + dir.x * dist, cam.iv[13] + dir.y * dist, cam.iv[14] + dir.z * dist);
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
        float dist = gb_t[i];
        uint8_t m;
        if (dist < 0.0f) {
            m = 0;
        } else {
            Vec3 P = gbuffer_position(cam, px, py, dist);
            Vec3 N = vpt::face_normal(gb_face[i], 1.0f);
            Vec3 L = vpt::mk3(lx, ly, lz);
            float ndotl = vpt::dot3(N, L);
            if (ndotl <= 0.01f) {
                m = 1;
            } else {
                Vec3 o = vpt::mk3(P.x + N.x * 0.06f, P.y + N.y * 0.06f, P.z + N.z * 0.06f);
                int block_at = vpt::voxel_at(w, o);
                float T = -1.0f;
                if (dist > 0.0f) {
                    vpt::Hit h = vpt::traverse(w, o, L, max_iter);
                    T = h.t;
                    samples = h.samples;
                    rays = 1;
                }
                m = (T > 0.0f || block_at > 0) ? 1 : 0;
            }
        }
        mask[i] = m;
    }
    count_samples(counter, samples, rays);
}

// IrridianceToSH (DiffuseRayTraceFrag.glsl:766-784).
__device__ inline void irradiance_to_sh(Vec3 r, Vec3 dir, float sh[6])
{
    float co = r.x - r.z;
    float t = r.z + co * 0.5f;
    float cg = r.y - t;
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

__device__ inline float fclamp(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }

__device__ inline void store_texel(vpt_svgf_texel* dst, const float sh[6], float luma, float ao, float sky)
{
    uint16_t h[8];
    for (int k = 0; k < 6; k++) h[k] = (uint16_t)vpt::hw_f2h(sh[k]);
    h[6] = (uint16_t)vpt::hw_f2h(luma);
    h[7] = (uint16_t)(vpt_unorm8_encode(ao) | ((uint16_t)vpt_unorm8_encode(sky) << 8));
    uint4 v = make_uint4((uint32_t)h[0] | ((uint32_t)h[1] << 16), (uint32_t)h[2] | ((uint32_t)h[3] << 16),
                         (uint32_t)h[4] | ((uint32_t)h[5] << 16), (uint32_t)h[6] | ((uint32_t)h[7] << 16));
    *reinterpret_cast<uint4*>(dst) = v;
}

// DIFFUSE_TRACE main + CalculateDiffuse (DiffuseRayTraceFrag.glsl:535-664,
// 822-935), 1 spp, 1 bounce; Lambert sun term, PCG randoms, per-id albedo,
// constant sky.
__global__ void __launch_bounds__(256) gi_kernel(WorldView w, CamView cam, const float* __restrict__ gb_t,
                                                 const uint8_t* __restrict__ gb_face, GiConst g,
                                                 const float* __restrict__ albedo, float* __restrict__ rgb,
                                                 vpt_svgf_texel* __restrict__ texels,
                                                 unsigned long long* __restrict__ counter)
{
    int px, py;
    tile_pixel(cam.W, cam.H, px, py);
    int samples = 0, rays = 0;
    if (px < cam.W && py < cam.H) {
        const size_t i = (size_t)py * cam.W + px;
        const float dist = gb_t[i];
        const Vec3 N = vpt::face_normal(gb_face[i], 0.5f);
        float sh[6];
        Vec3 x;
        float luma, ao_out, sky_out;
        if (dist < 0.0f) {
            x = vpt::mk3(g.sky_pixel[0], g.sky_pixel[1], g.sky_pixel[2]);
            irradiance_to_sh(x, N, sh);
            luma = 0.0f;
            ao_out = 1.0f;
            sky_out = 0.0f;
        } else {
            Vec3 P = gbuffer_position(cam, px, py, dist);
            uint32_t seed = (uint32_t)px + (uint32_t)cam.W * (uint32_t)py + (uint32_t)cam.W * (uint32_t)cam.H * g.frame;
            uint32_t h0 = vpt_pcg_hash(seed);
            uint32_t h1 = vpt_pcg_hash(h0);
            float r0 = vpt_u32_to_unit(h0), r1 = vpt_u32_to_unit(h1);
            const float bias = 0.06f;
            Vec3 o = vpt::mk3(P.x + N.x * bias, P.y + N.y * bias, P.z + N.z * bias);
            // cosWeightedRandomHemisphereDirection (:945-967)
            Vec3 uu = vpt::normalize3(vpt::cross3(N, vpt::mk3(0.0f, 1.0f, 1.0f)));
            Vec3 vv = vpt::cross3(uu, N);
            float ra = sqrtf(r1);
            float sn, cs;
            vpt_sincos_2pi(r0, &sn, &cs);
            float rx = ra * cs, ry = ra * sn;
            float rz = sqrtf(1.0f - r1);
            Vec3 dir = vpt::normalize3(vpt::mk3((rx * uu.x + ry * vv.x) + rz * N.x, (rx * uu.y + ry * vv.y) + rz * N.y,
                                                (rx * uu.z + ry * vv.z) + rz * N.z));
            float ao = 1.0f;
            bool skyhit = false;
            Vec3 contrib;
            vpt::Hit h = vpt::traverse(w, o, dir, g.gi_iter);
            samples = h.samples;
            rays = 1;
            float T = h.t;
            if (T > 0.0f && h.block > 0) {
                Vec3 ip = vpt::mk3(o.x + dir.x * T, o.y + dir.y * T, o.z + dir.z * T);
                Vec3 hn = vpt::face_normal(h.face, 0.0f);
                int tex_ref = min(max(h.block, 0), 127);
                Vec3 L = vpt::mk3(g.sun[0], g.sun[1], g.sun[2]);
                float ndotl = fmaxf(vpt::dot3(hn, L), 0.0f);
                float shadow_at = 0.0f;
                if (!(ndotl < 0.001f)) {
                    Vec3 so = vpt::mk3(ip.x + hn.x * 0.045f, ip.y + hn.y * 0.045f, ip.z + hn.z * 0.045f);
                    vpt::Hit sh_hit = vpt::traverse(w, so, L, g.shadow_iter);
                    samples += sh_hit.samples;
                    rays += 1;
                    shadow_at = sh_hit.t > 0.0f ? 1.0f : 0.0f;
                }
                const float PI = 3.14159265359f;
                float brdf = ndotl * (1.0f / PI);
                float vis = 1.0f - shadow_at;
                contrib.x = ((albedo[3 * tex_ref] * brdf) * (g.light[0] * 3.5f)) * vis * PI;
                contrib.y = ((albedo[3 * tex_ref + 1] * brdf) * (g.light[1] * 3.5f)) * vis * PI;
                contrib.z = ((albedo[3 * tex_ref + 2] * brdf) * (g.light[2] * 3.5f)) * vis * PI;
            } else {
                contrib = vpt::mk3(g.sky[0], g.sky[1], g.sky[2]);
                skyhit = true;
            }
            if (T < 2.0f && T > 0.0f) ao = fmaxf(T / 2.0f, 0.0f);
            x = vpt::mk3(fclamp(contrib.x, 0.0f, 8.0f), fclamp(contrib.y, 0.0f, 8.0f), fclamp(contrib.z, 0.0f, 8.0f));
            irradiance_to_sh(x, dir, sh);
            for (int k = 0; k < 6; k++) sh[k] = fclamp(sh[k], -100.0f, 100.0f);
            luma = fclamp(fmaxf((x.x * 0.299f + x.y * 0.587f) + x.z * 0.114f, 0.01f), 0.001f, 64.0f);
            ao_out = fclamp(ao, 0.0f, 1.0f);
            sky_out = skyhit ? 1.0f : 0.0f;
        }
        rgb[3 * i] = x.x;
        rgb[3 * i + 1] = x.y;
        rgb[3 * i + 2] = x.z;
        if (texels) store_texel(texels + i, sh, luma, ao_out, sky_out);
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
    if (max_iter < 0 || max_iter > 100000) return set_error(VPT_EINVAL, "%s: max_iter %d out of range", fn, max_iter);
    return VPT_OK;
}

dim3 tile_grid(int W, int H) { return dim3((unsigned)(((W + 15) / 16) * ((H + 15) / 16))); }

}  // namespace

extern "C" VPT_API int vpt_trace_primary(vpt_world* w, const vpt_camera* cam, int max_iter, const vpt_gbuffer* out,
                                         void* stream)
{
    if (!w) return set_error(VPT_EINVAL, "vpt_trace_primary: null world");
    if (int rc = check_cam(cam, "vpt_trace_primary")) return rc;
    if (int rc = check_gbuffer(out, cam, "vpt_trace_primary", true)) return rc;
    if (int rc = check_iter(max_iter, "vpt_trace_primary")) return rc;
    DeviceGuard g(w->device);
    hipLaunchKernelGGL(primary_kernel, tile_grid(cam->width, cam->height), dim3(256), 0, (hipStream_t)stream,
                       make_view(w), make_cam(cam), max_iter, out->d_t, out->d_face, out->d_block, out->d_voxel,
                       w->d_counter);
    HIP_TRY(hipGetLastError());
    return VPT_OK;
}

extern "C" VPT_API int vpt_trace_shadow(vpt_world* w, const vpt_camera* cam, const vpt_gbuffer* gb,
                                        const float sun_dir[3], int max_iter, uint8_t* d_mask, void* stream)
{
    if (!w || !sun_dir || !d_mask) return set_error(VPT_EINVAL, "vpt_trace_shadow: null argument");
    if (int rc = check_cam(cam, "vpt_trace_shadow")) return rc;
    if (int rc = check_gbuffer(gb, cam, "vpt_trace_shadow", false)) return rc;
    if (int rc = check_iter(max_iter, "vpt_trace_shadow")) return rc;
    DeviceGuard g(w->device);
    hipLaunchKernelGGL(shadow_kernel, tile_grid(cam->width, cam->height), dim3(256), 0, (hipStream_t)stream,
                       make_view(w), make_cam(cam), gb->d_t, gb->d_face, sun_dir[0], sun_dir[1], sun_dir[2], max_iter,
                       d_mask, w->d_counter);
    HIP_TRY(hipGetLastError());
    return VPT_OK;
}

extern "C" VPT_API int vpt_trace_gi(vpt_world* w, const vpt_camera* cam, const vpt_gbuffer* gb,
                                    const vpt_gi_params* prm, float* d_rgb, vpt_svgf_texel* d_svgf_in, void* stream)
{
    if (!w || !prm || !d_rgb) return set_error(VPT_EINVAL, "vpt_trace_gi: null argument");
    if (int rc = check_cam(cam, "vpt_trace_gi")) return rc;
    if (int rc = check_gbuffer(gb, cam, "vpt_trace_gi", false)) return rc;
    if (int rc = check_iter(prm->gi_iter, "vpt_trace_gi")) return rc;
    if (int rc = check_iter(prm->shadow_iter, "vpt_trace_gi")) return rc;
    DeviceGuard g(w->device);
    hipStream_t s = (hipStream_t)stream;
    // The albedo table lives with the context and is re-uploaded only when
    // it changes.
    vpt_ctx* ctx = w->ctx;
    if (!ctx->d_albedo) HIP_TRY(hipMalloc(&ctx->d_albedo, sizeof(ctx->albedo_host)));
    if (!ctx->albedo_valid || memcmp(ctx->albedo_host, prm->albedo, sizeof(ctx->albedo_host)) != 0) {
        memcpy(ctx->albedo_host, prm->albedo, sizeof(ctx->albedo_host));
        HIP_TRY(hipMemcpyAsync(ctx->d_albedo, ctx->albedo_host, sizeof(ctx->albedo_host), hipMemcpyHostToDevice, s));
        HIP_TRY(hipStreamSynchronize(s));
        ctx->albedo_valid = true;
    }
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
    hipLaunchKernelGGL(gi_kernel, tile_grid(cam->width, cam->height), dim3(256), 0, s, make_view(w), make_cam(cam),
                       gb->d_t, gb->d_face, c, ctx->d_albedo, d_rgb, d_svgf_in, w->d_counter);
    HIP_TRY(hipGetLastError());
    return VPT_OK;
}

extern "C" VPT_API int vpt_trace_rays(vpt_world* w, const vpt_ray* d_rays, int n, int max_iter, vpt_hit* d_hits,
                                      void* stream)
{
    if (!w || (n > 0 && (!d_rays || !d_hits))) return set_error(VPT_EINVAL, "vpt_trace_rays: null argument");
    if (n < 0) return set_error(VPT_EINVAL, "vpt_trace_rays: negative ray count");
    if (int rc = check_iter(max_iter, "vpt_trace_rays")) return rc;
    if (n == 0) return VPT_OK;
    DeviceGuard g(w->device);
    hipLaunchKernelGGL(rays_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       make_view(w), d_rays, n, max_iter, d_hits, w->d_counter);
    HIP_TRY(hipGetLastError());
    return VPT_OK;
}
