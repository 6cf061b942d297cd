// The colour pass's composite (ColorPassFrag.glsl main, :773-1010 with
// u_UseDFG on, Pipeline.cpp:3683) for flat dielectric materials: one lane
// per pixel over the G-buffer, the denoised diffuse texels, the denoised
// sun shadow and the reflection colour.  Every expression keeps the
// shader's operation order (no contraction) so the result is bit-exact
// with the CPU restatement of the same pass.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include "vpt_gi.h"
#include "vpt_internal.h"

using vpt::set_error;
using vpt::Vec3;

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return set_error(VPT_EDEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));      \
    } while (0)

namespace {

struct CompositeConst {
    float ls[3], lm[3];    // sun / moon directions
    float rs[3], rm[3];    // sun / moon radiances
    float sky[3];          // primary-miss pixels
    float r;               // roughness
    float spec_ao_exp;     // exp2(-16 r - 1), ComputeSpecularAO (:766-769)
    int Wr, Hr;            // reflection colour plane
    vpt::AlbedoTable albedo;
};

// FresnelSchlickRoughness(Eye, N, F0, r) (ColorPassFrag.glsl:1205-1209),
// scalar F0.
__device__ inline float fresnel_schlick_roughness(Vec3 eye, Vec3 n, float f0, float r)
{
    const float c = vpt::fclamp(vpt::dot3(eye, n), 0.00001f, 1.0f);
    return f0 + (fmaxf(1.0f - r, f0) - f0) * vpt_powi(1.0f - c, 5);
}

// CalculateDirectionalLight of the colour pass (:419-451): Cook-Torrance
// with the GGX specular term, metalness 0.
__device__ inline Vec3 color_direct(Vec3 P, Vec3 eye, Vec3 Li, const float rad[3], Vec3 alb, Vec3 N, float r,
                                    float shadow)
{
    const float PI = 3.14159265359f;
    const float sh = fminf(shadow, 1.0f);
    const Vec3 Lo = vpt::normalize3(vpt::mk3(eye.x - P.x, eye.y - P.y, eye.z - P.z));
    const float cos_lo = fmaxf(0.0f, vpt::dot3(N, Lo));
    const Vec3 Lh = vpt::normalize3(vpt::mk3(Li.x + Lo.x, Li.y + Lo.y, Li.z + Lo.z));
    const float cos_li = fmaxf(0.0f, vpt::dot3(N, Li));
    const float cos_lh = fmaxf(0.0f, vpt::dot3(N, Lh));
    const float F = fresnel_schlick_roughness(Lo, N, 0.04f, r);
    const float alpha = r * r;
    const float alpha2 = alpha * alpha;
    const float denom = (cos_lh * cos_lh) * (alpha2 - 1.0f) + 1.0f;
    const float D = alpha2 / ((PI * denom) * denom);
    const float rk = r + 1.0f;
    const float k = (rk * rk) / 8.0f;
    const float G = (cos_li / (cos_li * (1.0f - k) + k)) * (cos_lo / (cos_lo * (1.0f - k) + k));
    const float kd = 1.0f - F;
    const float spec = vpt::fclamp(((F * D) * G) / fmaxf(0.00001f, (4.0f * cos_li) * cos_lo), 0.0f, 2.0f);
    const float vis = vpt::fclamp(1.0f - sh, 0.0f, 1.0f);
    const float a[3] = {alb.x, alb.y, alb.z};
    float o[3];
    for (int c = 0; c < 3; c++)
        o[c] = vpt::fclamp(((kd * a[c]) * rad[c]) * cos_li + (spec * rad[c]) * cos_li, 0.0f, 2.5f) * vis;
    return vpt::mk3(o[0], o[1], o[2]);
}

// texture(LINEAR, REPEAT) of the RGB of an RGBA binary16 plane at (s, t).
__device__ inline Vec3 rgba16_linear(const uint2* __restrict__ p, int W, int H, float s, float t)
{
    const float u = s * (float)W - 0.5f, v = t * (float)H - 0.5f;
    const float fu = floorf(u), fv = floorf(v);
    const float a = u - fu, b = v - fv;
    auto wrap = [](int i, int n) { const int r = i % n; return r < 0 ? r + n : r; };
    const int x0 = wrap((int)fu, W), x1 = wrap((int)fu + 1, W);
    const size_t r0 = (size_t)wrap((int)fv, H) * W, r1 = (size_t)wrap((int)fv + 1, H) * W;
    const uint2 q00 = p[r0 + x0], q10 = p[r0 + x1], q01 = p[r1 + x0], q11 = p[r1 + x1];
    auto lin = [&](uint32_t w00, uint32_t w10, uint32_t w01, uint32_t w11) {
        const float top = vpt::hw_h2f(w00) * (1.0f - a) + vpt::hw_h2f(w10) * a;
        const float bot = vpt::hw_h2f(w01) * (1.0f - a) + vpt::hw_h2f(w11) * a;
        return top * (1.0f - b) + bot * b;
    };
    return vpt::mk3(lin(q00.x & 0xffffu, q10.x & 0xffffu, q01.x & 0xffffu, q11.x & 0xffffu),
                    lin(q00.x >> 16, q10.x >> 16, q01.x >> 16, q11.x >> 16),
                    lin(q00.y & 0xffffu, q10.y & 0xffffu, q01.y & 0xffffu, q11.y & 0xffffu));
}

__global__ void __launch_bounds__(256) composite_kernel(vpt::CamView cam, const float* __restrict__ gb_t,
                                                        const uint8_t* __restrict__ gb_face,
                                                        const uint8_t* __restrict__ gb_block,
                                                        const vpt_svgf_texel* __restrict__ diffuse,
                                                        const uint8_t* __restrict__ shadow,
                                                        const uint2* __restrict__ refl, const CompositeConst k,
                                                        float* __restrict__ out)
{
    const int W = cam.W, H = cam.H;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)W * H) return;
    const int px = (int)(i % (size_t)W), py = (int)(i / (size_t)W);
    const float u = vpt::texcoord(px, W), v = vpt::texcoord(py, H);
    const float dist = 1.0f / (1.0f / gb_t[i]);   // SamplePositionAt reads the 1 / t plane
    float* o = out + 3 * i;
    if (!(dist > 0.0f)) {
        o[0] = k.sky[0];
        o[1] = k.sky[1];
        o[2] = k.sky[2];
        return;
    }
    const Vec3 eye = vpt::mk3(cam.iv[12], cam.iv[13], cam.iv[14]);
    const Vec3 dir = vpt::normalize3(vpt::ray_dir_at(cam.iv, cam.ip, u, v));
    const Vec3 P = vpt::mk3(eye.x + dir.x * dist, eye.y + dir.y * dist, eye.z + dir.z * dist);
    const Vec3 N = vpt::face_normal(gb_face[i], 1.0f);
    const int ref = gb_block[i] & 127;
    const Vec3 a0 = vpt::mk3(k.albedo.v[3 * ref], k.albedo.v[3 * ref + 1], k.albedo.v[3 * ref + 2]);
    // BasicSaturation(albedo, 1 - 0.1) (:812)
    const float adj = 1.0f - 0.1f;
    const float lum = (a0.x * 0.2125f + a0.y * 0.7154f) + a0.z * 0.0721f;
    const Vec3 A = vpt::mk3(lum * (1.0f - adj) + a0.x * adj, lum * (1.0f - adj) + a0.y * adj,
                            lum * (1.0f - adj) + a0.z * adj);
    const float shd = vpt::fclamp(
        vpt::fclamp(vpt::fclamp(vpt_unorm8_decode(shadow[i]), 0.0f, 1.0f) + 0.0f, 0.0f, 1.0f), 0.0f, 1.0f);
    // SHToIrridiance(SHy, CoCg, N) (:356-368)
    const uint4 q = *reinterpret_cast<const uint4*>(diffuse + i);
    const float s0 = vpt::hw_h2f(q.x & 0xffffu), s1 = vpt::hw_h2f(q.x >> 16);
    const float s2 = vpt::hw_h2f(q.y & 0xffffu), s3 = vpt::hw_h2f(q.y >> 16);
    const float c0 = vpt::hw_h2f(q.z & 0xffffu), c1 = vpt::hw_h2f(q.z >> 16);
    const float x = (s0 * N.x + s1 * N.y) + s2 * N.z;
    float Y = 2.0f * (1.023326f * x + 0.886226f * s3);
    Y = fmaxf(Y, 0.0f);
    const float kk = (Y * 0.282095f) / (s3 + 1e-6f);
    const float co = c0 * kk, cg = c1 * kk;
    const float T = Y - cg * 0.5f;
    const float G = cg + T;
    const float B = T - co * 0.5f;
    const float Rr = B + co;
    Vec3 ed = vpt::mk3(fmaxf(Rr, 0.0f), fmaxf(G, 0.0f), fmaxf(B, 0.0f));
    const float vb = 0.00125f;
    if (u > vb && u < 1.0f - vb && v > vb && v < 1.0f - vb) {   // VXAO (:848-876)
        float aov = vpt_unorm8_decode((uint8_t)((q.w >> 16) & 0xffu));
        const float dx = P.x - eye.x, dz = P.z - eye.z;
        const float fade = sqrtf(dx * dx + dz * dz) < 256.0f ? 0.0f : 1.0f;
        aov = aov * (1.0f - fade) + 1.0f * fade;
        aov = vpt_powf(aov, 2.2f);
        const float m = vpt::fclamp(aov, 0.05f, 1.0f);
        ed = vpt::mk3(ed.x * m, ed.y * m, ed.z * m);
    }
    float sun_vis = vpt::fclamp(k.ls[1] + 0.05f, 0.0f, 0.1f) * 12.0f;
    sun_vis = 1.0f - sun_vis;
    const Vec3 Ls = vpt::mk3(k.ls[0], k.ls[1], k.ls[2]), Lm = vpt::mk3(k.lm[0], k.lm[1], k.lm[2]);
    const Vec3 sd = color_direct(P, eye, Ls, k.rs, A, N, k.r, shd);
    const Vec3 md = color_direct(P, eye, Lm, k.rm, A, N, k.r, shd);
    Vec3 dl = vpt::mk3(sd.x * (1.0f - sun_vis) + md.x * sun_vis, sd.y * (1.0f - sun_vis) + md.y * sun_vis,
                       sd.z * (1.0f - sun_vis) + md.z * sun_vis);
    dl = vpt::mk3(fmaxf(dl.x, 0.000001f), fmaxf(dl.y, 0.000001f), fmaxf(dl.z, 0.000001f));
    const Vec3 di = vpt::mk3(ed.x * A.x, ed.y * A.y, ed.z * A.z);
    Vec3 si = vpt::mk3(0.0f, 0.0f, 0.0f);
    const Vec3 I = vpt::normalize3(vpt::mk3(P.x - eye.x, P.y - eye.y, P.z - eye.z));
    const float kr = 2.0f * vpt::dot3(N, I);
    const Vec3 R = vpt::mk3(I.x - kr * N.x, I.y - kr * N.y, I.z - kr * N.z);
    const float sb = 0.0035f;
    if (u > sb && u < 1.0f - sb && v > sb && v < 1.0f - sb) {
        const float att = vpt::fclamp(vpt_powi(fmaxf(vpt::dot3(N, N), 0.00000001f), 2) * 1.5f, 0.1f, 1.0f);
        const Vec3 rc = rgba16_linear(refl, k.Wr, k.Hr, u, v);
        si = vpt::mk3(rc.x * att, rc.y * att, rc.z * att);
    }
    si = vpt::mk3(fmaxf(0.0001f, si.x), fmaxf(0.0001f, si.y), fmaxf(0.0001f, si.z));
    const Vec3 Lo = vpt::normalize3(vpt::mk3(eye.x - P.x, eye.y - P.y, eye.z - P.z));
    const float nov = vpt::fclamp(vpt::dot3(N, Lo), 0.0000000001f, 1.0f);
    float sao = vpt::fclamp((vpt_powf(nov + 1.0f, k.spec_ao_exp) - 1.0f) + 1.0f, 0.0f, 1.0f);
    sao = vpt_powf(sao, 1.5f);
    float hz = fminf(1.0f + vpt::dot3(R, N), 1.0f);
    hz = hz * (hz * hz);
    si = vpt::mk3((si.x * sao) * hz, (si.y * sao) * hz, (si.z * sao) * hz);
    const float F = vpt::fclamp(fresnel_schlick_roughness(Lo, N, 0.04f, k.r), 0.0f, 1.0f);
    const float kd = (1.0f - F) * (1.0f - 0.0f);
    // KarisEnvBRDFApprox (:385-392)
    const float ndv = vpt::fclamp(fmaxf(-vpt::dot3(I, N), 0.000001f), 0.0f, 1.0f);
    const float rr = vpt::fclamp(k.r, 0.0f, 1.0f);
    const float qx = rr * -1.0f + 1.0f, qy = rr * -0.0275f + 0.0425f, qz = rr * -0.572f + 1.040f,
                qw = rr * 0.022f + -0.040f;
    const float a004 = fminf(qx * qx, vpt_exp2f(-9.28f * ndv)) * qx + qy;
    const float ex = -1.04f * a004 + qz, ey = 1.04f * a004 + qw;
    const float fs = F * ex + ey;
    const float ao_d = vpt::fclamp(1.0f, 0.2f, 1.01f);
    o[0] = dl.x + (kd * (di.x * ao_d) + si.x * fs);
    o[1] = dl.y + (kd * (di.y * ao_d) + si.y * fs);
    o[2] = dl.z + (kd * (di.z * ao_d) + si.z * fs);
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

}  // namespace

extern "C" VPT_API int vpt_composite_params_from_gi(const vpt_gi_params* gi, vpt_composite_params* out)
{
    if (!gi || !out) return set_error(VPT_EINVAL, "vpt_composite_params_from_gi: null argument");
    vpt_composite_params p{};
    memcpy(p.sun_dir, gi->sun_dir, sizeof(p.sun_dir));
    const float mx = -gi->sun_dir[0], my = -gi->sun_dir[1], mz = gi->sun_dir[2];
    const float len = sqrtf((mx * mx + my * my) + mz * mz);
    p.moon_dir[0] = mx / len;
    p.moon_dir[1] = my / len;
    p.moon_dir[2] = mz / len;
    memcpy(p.sun_radiance, gi->light_color, sizeof(p.sun_radiance));
    p.moon_radiance[0] = 0.1f;
    p.moon_radiance[1] = 0.1f;
    p.moon_radiance[2] = 1.0f;
    memcpy(p.sky_pixel_radiance, gi->sky_pixel_radiance, sizeof(p.sky_pixel_radiance));
    p.roughness = 0.5f;
    memcpy(p.albedo, gi->albedo, sizeof(p.albedo));
    *out = p;
    return VPT_OK;
}

extern "C" VPT_API int vpt_composite(vpt_ctx* ctx, const vpt_camera* cam, const vpt_gbuffer* gb,
                                     const vpt_composite_params* prm, const vpt_svgf_texel* d_diffuse,
                                     const uint8_t* d_shadow, const uint16_t* d_reflect, int reflect_div, float* d_rgb,
                                     void* stream)
{
    if (!ctx || !prm || !d_diffuse || !d_shadow || !d_reflect || !d_rgb)
        return set_error(VPT_EINVAL, "vpt_composite: null argument");
    if (int rc = vpt::check_cam(cam, "vpt_composite")) return rc;
    if (int rc = vpt::check_gbuffer(gb, cam, "vpt_composite", true)) return rc;
    if (reflect_div < 1 || reflect_div > 16 || cam->width / reflect_div < 1 || cam->height / reflect_div < 1)
        return set_error(VPT_EINVAL, "vpt_composite: reflect_div must be 1..16 and leave at least one pixel");
    if (!(prm->roughness >= 0.0f && prm->roughness <= 1.0f))
        return set_error(VPT_EINVAL, "vpt_composite: roughness must be in [0, 1]");
    CompositeConst k{};
    for (int c = 0; c < 3; c++) {
        k.ls[c] = prm->sun_dir[c];
        k.lm[c] = prm->moon_dir[c];
        k.rs[c] = prm->sun_radiance[c];
        k.rm[c] = prm->moon_radiance[c];
        k.sky[c] = prm->sky_pixel_radiance[c];
    }
    k.r = prm->roughness;
    k.spec_ao_exp = (float)exp2((double)(-16.0f * prm->roughness - 1.0f));
    k.Wr = cam->width / reflect_div;
    k.Hr = cam->height / reflect_div;
    static_assert(sizeof(k.albedo.v) == sizeof(prm->albedo), "albedo table size");
    memcpy(k.albedo.v, prm->albedo, sizeof(k.albedo.v));
    DeviceGuard g(ctx->device);
    const vpt::CamView cv = vpt::make_cam(cam);
    const size_t n = (size_t)cam->width * cam->height;
    hipLaunchKernelGGL(composite_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, cv,
                       gb->d_t, gb->d_face, gb->d_block, d_diffuse, d_shadow,
                       reinterpret_cast<const uint2*>(d_reflect), k, d_rgb);
    HIP_TRY(hipGetLastError());
    return VPT_OK;
}
