// Per-pixel pieces of the primary / shadow / diffuse-GI passes, shared by
// the one-launch-per-pass kernels (vpt_trace.hip) and the persistent frame
// path (vpt_batch.hip) so that both run the same operation sequence.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vpt/vpt.h"
#include "vpt_device.h"

namespace vpt {

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

__device__ inline float fclamp(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }

// Primary ray of pixel (px, py): GetRayStuff + normalize (InitialRayTraceFrag.glsl:398-427).
__device__ inline Vec3 primary_dir(const CamView& cam, int px, int py)
{
    return normalize3(ray_dir_at(cam.iv, cam.ip, texcoord(px, cam.W), texcoord(py, cam.H)));
}

// G-buffer world position: GetPositionAt (ShadowRayTraceFrag.glsl:299-303)
// at the pixel centre.
__device__ inline Vec3 gbuffer_position(const CamView& cam, int px, int py, float dist)
{
    Vec3 dir = primary_dir(cam, px, py);
    return mk3(cam.iv[12] + dir.x * dist, cam.iv[13] + dir.y * dist, cam.iv[14] + dir.z * dist);
}

// Shadow ray set-up of ShadowRayTraceFrag.glsl:480-508.  Returns 0: mask
// known without tracing (*mask set), 1: trace `o` toward L and the mask is
// (T > 0).  A ray whose start voxel is solid has mask 1 whatever T is, so it
// is not traced at all.
__device__ inline int shadow_setup(const WorldView& w, const CamView& cam, int px, int py, float dist, int face,
                                   Vec3 L, Vec3& o, uint8_t& mask)
{
    if (dist < 0.0f) {
        mask = 0;
        return 0;
    }
    Vec3 P = gbuffer_position(cam, px, py, dist);
    Vec3 N = face_normal(face, 1.0f);
    float ndotl = dot3(N, L);
    if (ndotl <= 0.01f) {
        mask = 1;
        return 0;
    }
    o = mk3(P.x + N.x * 0.06f, P.y + N.y * 0.06f, P.z + N.z * 0.06f);
    if (voxel_at(w, o) > 0 || !(dist > 0.0f)) {
        mask = (uint8_t)(voxel_at(w, o) > 0 ? 1 : 0);
        return 0;
    }
    return 1;
}

// The GI ray of a hit pixel: origin P + N * 0.06 and the cosine-weighted
// direction (DiffuseRayTraceFrag.glsl:541-549, 945-967) with PCG randoms
// seeded by pcg_hash(px + W*py + W*H*frame).
__device__ inline void gi_ray(const CamView& cam, int px, int py, float dist, Vec3 N, uint32_t frame, Vec3& o,
                              Vec3& dir)
{
    Vec3 P = gbuffer_position(cam, px, py, dist);
    uint32_t seed = (uint32_t)px + (uint32_t)cam.W * (uint32_t)py + (uint32_t)cam.W * (uint32_t)cam.H * frame;
    uint32_t h0 = vpt_pcg_hash(seed);
    uint32_t h1 = vpt_pcg_hash(h0);
    float r0 = vpt_u32_to_unit(h0), r1 = vpt_u32_to_unit(h1);
    const float bias = 0.06f;
    o = mk3(P.x + N.x * bias, P.y + N.y * bias, P.z + N.z * bias);
    Vec3 uu = normalize3(cross3(N, mk3(0.0f, 1.0f, 1.0f)));
    Vec3 vv = cross3(uu, N);
    float ra = sqrtf(r1);
    float sn, cs;
    vpt_sincos_2pi(r0, &sn, &cs);
    float rx = ra * cs, ry = ra * sn;
    float rz = sqrtf(1.0f - r1);
    dir = normalize3(mk3((rx * uu.x + ry * vv.x) + rz * N.x, (rx * uu.y + ry * vv.y) + rz * N.y,
                         (rx * uu.z + ry * vv.z) + rz * N.z));
}

// GI hit -> shadow ray toward the sun from ip + hn * 0.045
// (DiffuseRayTraceFrag.glsl:583-597, GetShadowAt :1202-1222).  Returns
// false when N.L < 0.001 (no shadow ray; ShadowAt = 0).
__device__ inline bool gi_shadow_origin(Vec3 o, Vec3 dir, float T, int hit_face, Vec3 L, Vec3& so, float& ndotl)
{
    Vec3 ip = mk3(o.x + dir.x * T, o.y + dir.y * T, o.z + dir.z * T);
    Vec3 hn = face_normal(hit_face, 0.0f);
    ndotl = fmaxf(dot3(hn, L), 0.0f);
    if (ndotl < 0.001f) return false;
    so = mk3(ip.x + hn.x * 0.045f, ip.y + hn.y * 0.045f, ip.z + hn.z * 0.045f);
    return true;
}

__device__ inline bool gi_shadow_ray(Vec3 o, Vec3 dir, float T, int hit_face, const GiConst& g, Vec3& so,
                                     float& ndotl)
{
    return gi_shadow_origin(o, dir, T, hit_face, mk3(g.sun[0], g.sun[1], g.sun[2]), so, ndotl);
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

__device__ inline void store_texel(vpt_svgf_texel* dst, const float sh[6], float luma, float ao, float sky)
{
    uint32_t h[8];
    for (int k = 0; k < 6; k++) h[k] = hw_f2h(sh[k]);
    h[6] = hw_f2h(luma);
    h[7] = (uint32_t)vpt_unorm8_encode(ao) | ((uint32_t)vpt_unorm8_encode(sky) << 8);
    *reinterpret_cast<uint4*>(dst) = make_uint4(h[0] | (h[1] << 16), h[2] | (h[3] << 16), h[4] | (h[5] << 16),
                                                h[6] | (h[7] << 16));
}

// Writes the outputs of a primary-miss pixel (DiffuseRayTraceFrag.glsl:856-862).
__device__ inline void gi_write_sky_pixel(const GiConst& g, Vec3 N, float* rgb, vpt_svgf_texel* texel)
{
    Vec3 x = mk3(g.sky_pixel[0], g.sky_pixel[1], g.sky_pixel[2]);
    float sh[6];
    irradiance_to_sh(x, N, sh);
    rgb[0] = x.x;
    rgb[1] = x.y;
    rgb[2] = x.z;
    if (texel) store_texel(texel, sh, 0.0f, 1.0f, 0.0f);
}

// Shading of one GI sample (CalculateDiffuse :563-661 with 1 bounce and the
// Lambert sun term) and the per-pixel outputs of main (:895-934).
// hit: the GI ray hit a block (T > 0); shadow_at: 1 if the sun shadow ray
// was traced and hit.
__device__ inline void gi_write_hit_pixel(const GiConst& g, const float* albedo, Vec3 dir, float T, bool hit,
                                          int hit_face, int hit_block, float shadow_at, float* rgb,
                                          vpt_svgf_texel* texel)
{
    Vec3 contrib;
    bool skyhit = false;
    if (hit) {
        Vec3 hn = face_normal(hit_face, 0.0f);
        Vec3 L = mk3(g.sun[0], g.sun[1], g.sun[2]);
        float ndotl = fmaxf(dot3(hn, L), 0.0f);
        int tex_ref = min(max(hit_block, 0), 127);
        const float PI = 3.14159265359f;
        float brdf = ndotl * (1.0f / PI);
        float vis = 1.0f - shadow_at;
        contrib.x = ((albedo[3 * tex_ref] * brdf) * (g.light[0] * 3.5f)) * vis * PI;
        contrib.y = ((albedo[3 * tex_ref + 1] * brdf) * (g.light[1] * 3.5f)) * vis * PI;
        contrib.z = ((albedo[3 * tex_ref + 2] * brdf) * (g.light[2] * 3.5f)) * vis * PI;
    } else {
        contrib = mk3(g.sky[0], g.sky[1], g.sky[2]);
        skyhit = true;
    }
    float ao = 1.0f;
    if (T < 2.0f && T > 0.0f) ao = fmaxf(T / 2.0f, 0.0f);
    Vec3 x = mk3(fclamp(contrib.x, 0.0f, 8.0f), fclamp(contrib.y, 0.0f, 8.0f), fclamp(contrib.z, 0.0f, 8.0f));
    float sh[6];
    irradiance_to_sh(x, dir, sh);
    for (int k = 0; k < 6; k++) sh[k] = fclamp(sh[k], -100.0f, 100.0f);
    float luma = fclamp(fmaxf((x.x * 0.299f + x.y * 0.587f) + x.z * 0.114f, 0.01f), 0.001f, 64.0f);
    rgb[0] = x.x;
    rgb[1] = x.y;
    rgb[2] = x.z;
    if (texel) store_texel(texel, sh, luma, fclamp(ao, 0.0f, 1.0f), skyhit ? 1.0f : 0.0f);
}

// The GI albedo table (128 block ids x RGB).  Kernels receive it by value,
// so every launch carries the table it was enqueued with, in stream order.
struct AlbedoTable {
    float v[128 * 3];
};

// Host helpers shared by the launchers (vpt_trace.hip).
WorldView make_view(const vpt_world* w);
CamView make_cam(const vpt_camera* c);
GiConst make_gi(const vpt_gi_params* prm);
// The call's albedo table; also records it as the context's last table.
AlbedoTable take_albedo(vpt_ctx* ctx, const vpt_gi_params* prm);
int check_cam(const vpt_camera* cam, const char* fn);
int check_gbuffer(const vpt_gbuffer* gb, const vpt_camera* cam, const char* fn, bool need_block);
int check_iter(int max_iter, const char* fn);

}  // namespace vpt
