// SVGF chain on gfx950: pre-temporal 3x3 spatial, temporal reprojection,
// variance estimate, a-trous x5, SH decode (Core/Pipeline.cpp:2284-2611).
//
// Storage follows the reference's render targets (Pipeline.cpp:1102-1108):
// SH RGBA16F + CoCg RG16F + one R16F channel + AO/sky RG8, packed as one
// 16-byte texel per pixel [SH0..3, Co, Cg, X, (sky<<8)|ao] so that a tap is a
// single 128-bit load; the temporal utility (spp, moment, luma) is a second
// 8-byte plane.  Every pass computes in fp32 with the same IEEE operation
// sequence as the shader it cites and rounds to binary16 / unorm8 on store.
// Fractional taps (temporal reprojection, a-trous at step * 1.35 + jitter)
// are GL LINEAR + REPEAT bilinear fetches; taps at integer pixel offsets
// read the texel directly.
#include <hip/hip_runtime.h>
#include <string.h>

#include "vpt_device.h"
#include "vpt_internal.h"

using vpt::set_error;
using vpt::Vec3;

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return set_error(VPT_EDEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));      \
    } while (0)

struct vpt_svgf {
    int device = 0;
    int W = 0, H = 0;
    vpt_svgf_params prm{};
    vpt_svgf_texel* d_pre = nullptr;        // pre-spatial output
    vpt_svgf_texel* d_temporal[2] = {};     // temporal output, ping-pong (history)
    uint2* d_util[2] = {};                  // temporal utility (spp, moment, luma, 0)
    vpt_svgf_texel* d_var = nullptr;        // variance-pass output
    vpt_svgf_texel* d_atrous[2] = {};       // a-trous ping-pong
    int cur = 0;                            // index of the newest temporal plane
    int last_final = 0;                     // a-trous plane holding the last output
    bool have_history = false;
};

struct vpt_shadow_denoise {
    int device = 0;
    int W = 0, H = 0;
    uint8_t* d_color[2] = {};     // ShadowTemporalFBO colour (GL_RED unorm8), ping-pong history
    uint16_t* d_frames[2] = {};   // ShadowTemporalFBO frame count (R16F bits), ping-pong
    int cur = 0;                  // index of the newest temporal planes
};

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

struct CamView {
    float iv[16];
    float ip[16];
};

struct Mat16 {
    float m[16];
};

struct Tx {
    float sh[4];
    float co, cg;
    float x;
    float ao, sky;
};

__device__ inline float h2f(uint32_t h) { return vpt::hw_h2f(h); }
__device__ inline uint32_t f2h(float x) { return vpt::hw_f2h(x); }

__device__ inline Tx load_tx_word(uint4 v)
{
    Tx t;
    t.sh[0] = h2f(v.x & 0xffffu);
    t.sh[1] = h2f(v.x >> 16);
    t.sh[2] = h2f(v.y & 0xffffu);
    t.sh[3] = h2f(v.y >> 16);
    t.co = h2f(v.z & 0xffffu);
    t.cg = h2f(v.z >> 16);
    t.x = h2f(v.w & 0xffffu);
    t.ao = vpt_unorm8_decode((uint8_t)((v.w >> 16) & 0xffu));
    t.sky = vpt_unorm8_decode((uint8_t)(v.w >> 24));
    return t;
}

__device__ inline Tx load_tx(const vpt_svgf_texel* p, size_t i) { return load_tx_word(*reinterpret_cast<const uint4*>(p + i)); }

// The same decode with the AO / sky bytes read from a 256-entry table of
// vpt_unorm8_decode in LDS (one LDS read in place of the three VALU
// operations of the exact decode; the table holds the same bits).
__device__ inline Tx load_tx_word_lut(uint4 v, const float* unorm8)
{
    Tx t;
    t.sh[0] = h2f(v.x & 0xffffu);
    t.sh[1] = h2f(v.x >> 16);
    t.sh[2] = h2f(v.y & 0xffffu);
    t.sh[3] = h2f(v.y >> 16);
    t.co = h2f(v.z & 0xffffu);
    t.cg = h2f(v.z >> 16);
    t.x = h2f(v.w & 0xffffu);
    typedef const __attribute__((address_space(3))) float LdsFloat;
    t.ao = ((LdsFloat*)unorm8)[(v.w >> 16) & 0xffu];   // address-space cast: the table is in LDS
    t.sky = ((LdsFloat*)unorm8)[v.w >> 24];
    return t;
}

// Fills the table (a block of at least 256 threads; the caller then waits
// at a barrier).
__device__ inline void fill_unorm8_table(float* unorm8)
{
    if (threadIdx.x < 256) unorm8[threadIdx.x] = vpt_unorm8_decode((uint8_t)threadIdx.x);
}

__device__ inline void store_tx(vpt_svgf_texel* p, size_t i, const float sh[4], float co, float cg, float x, float ao,
                                float sky)
{
    uint32_t w0 = f2h(sh[0]) | (f2h(sh[1]) << 16);
    uint32_t w1 = f2h(sh[2]) | (f2h(sh[3]) << 16);
    uint32_t w2 = f2h(co) | (f2h(cg) << 16);
    uint32_t w3 = f2h(x) | ((uint32_t)vpt_unorm8_encode(ao) << 16) | ((uint32_t)vpt_unorm8_encode(sky) << 24);
    *reinterpret_cast<uint4*>(p + i) = make_uint4(w0, w1, w2, w3);
}

__device__ inline float fclamp(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
__device__ inline float sh_to_y(float w) { return fmaxf(0.0f, 3.544905f * w); }
__device__ inline bool in_screen(float x, float y) { return x > 0.0f && x < 1.0f && y > 0.0f && y < 1.0f; }



// GL LINEAR + REPEAT footprint at normalised (s, t); valid for coordinates
// inside (0, 1), where the taps lie in [-1, n].
struct Bil {
    size_t i00, i10, i01, i11;
    float a, b;
};

__device__ inline int wrap1(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

__device__ inline Bil bil_setup(float s, float t, int W, int H)
{
    float u = s * (float)W - 0.5f;
    float v = t * (float)H - 0.5f;
    float fu = floorf(u), fv = floorf(v);
    Bil b;
    b.a = u - fu;
    b.b = v - fv;
    int x0 = wrap1((int)fu, W), x1 = wrap1((int)fu + 1, W);
    int y0 = wrap1((int)fv, H), y1 = wrap1((int)fv + 1, H);
    b.i00 = (size_t)y0 * W + x0;
    b.i10 = (size_t)y0 * W + x1;
    b.i01 = (size_t)y1 * W + x0;
    b.i11 = (size_t)y1 * W + x1;
    return b;
}

// vpt_expf for arguments x <= 0, which the sun-shadow weights pass: the
// same operations, so the same results, without the NaN and overflow tests
// (they cannot fire) and with the underflow test (x below -87.34 gives 0) as
// a select instead of a branch.  (In the SVGF passes the branch form was as
// fast or faster.)
__device__ inline float expf_nonpos(float x)
{
    const float LOG2E = 1.44269502162933349609375f;
    const float LN2_HI = 0.693145751953125f;
    const float LN2_LO = 1.428606765330187045e-06f;
    const float kf = floorf(x * LOG2E + 0.5f);
    float r = x - kf * LN2_HI;
    r = r - kf * LN2_LO;
    float p = 0.0013829420786350965f;
    p = p * r + 0.008374771103262901f;
    p = p * r + 0.0416683591902256f;
    p = p * r + 0.1666642129421234f;
    p = p * r + 0.49999991059303284f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    const int k = (int)fmaxf(kf, -300.0f);   // kf >= -126 wherever the result is used
    const int k1 = k / 2;
    const int k2 = k - k1;
    const float s1 = vpt_u2f((uint32_t)(k1 + 127) << 23);
    const float s2 = vpt_u2f((uint32_t)(k2 + 127) << 23);
    const float v = (p * s1) * s2;
    return x < -87.33654022216796875f ? 0.0f : v;
}

__device__ inline float bil(float t00, float t10, float t01, float t11, float a, float b)
{
    float top = t00 * (1.0f - a) + t10 * a;
    float bot = t01 * (1.0f - a) + t11 * a;
    return top * (1.0f - b) + bot * b;
}

__device__ inline Tx bil_of(const Tx& q00, const Tx& q10, const Tx& q01, const Tx& q11, const Bil& f)
{
    Tx r;
    for (int c = 0; c < 4; c++) r.sh[c] = bil(q00.sh[c], q10.sh[c], q01.sh[c], q11.sh[c], f.a, f.b);
    r.co = bil(q00.co, q10.co, q01.co, q11.co, f.a, f.b);
    r.cg = bil(q00.cg, q10.cg, q01.cg, q11.cg, f.a, f.b);
    r.x = bil(q00.x, q10.x, q01.x, q11.x, f.a, f.b);
    r.ao = bil(q00.ao, q10.ao, q01.ao, q11.ao, f.a, f.b);
    r.sky = bil(q00.sky, q10.sky, q01.sky, q11.sky, f.a, f.b);
    return r;
}

__device__ inline Tx bil_tx(const vpt_svgf_texel* p, const Bil& f)
{
    return bil_of(load_tx(p, f.i00), load_tx(p, f.i10), load_tx(p, f.i01), load_tx(p, f.i11), f);
}

__device__ inline Tx bil_tx_lut(const vpt_svgf_texel* p, const Bil& f, const float* unorm8)
{
    const uint4* q = reinterpret_cast<const uint4*>(p);
    return bil_of(load_tx_word_lut(q[f.i00], unorm8), load_tx_word_lut(q[f.i10], unorm8),
                  load_tx_word_lut(q[f.i01], unorm8), load_tx_word_lut(q[f.i11], unorm8), f);
}

// GL NEAREST + REPEAT texel index for coordinates inside (0, 1), where
// floor(s * W) lies in [0, W] (W itself when s * W rounds up to W).
__device__ inline int nearest_idx(float s, float t, int W, int H)
{
    int i = wrap1((int)floorf(s * (float)W), W), j = wrap1((int)floorf(t * (float)H), H);
    return j * W + i;
}

// The X channel (binary16 in the low half of the texel's last word).
__device__ inline float load_x(const vpt_svgf_texel* p, size_t i)
{
    return h2f(reinterpret_cast<const uint32_t*>(p + i)[3] & 0xffffu);
}

// Pixel of this lane: 32x8 tiles (a wavefront covers 32x2 pixels, i.e.
// 512-byte texel rows), XCD-banded.
__device__ inline bool pixel_of(int W, int H, int& px, int& py)
{
    const uint32_t tiles_x = (uint32_t)(W + 31) / 32u;
    const uint32_t tiles_y = (uint32_t)(H + 7) / 8u;
    const uint32_t tile = vpt::xcd_remap(blockIdx.x, tiles_x * tiles_y);
    px = (int)(tile % tiles_x) * 32 + (int)(threadIdx.x & 31);
    py = (int)(tile / tiles_x) * 8 + (int)(threadIdx.x >> 5);
    return px < W && py < H;
}

dim3 pass_grid(int W, int H) { return dim3((unsigned)(((W + 31) / 32) * ((H + 7) / 8))); }

__device__ inline float texc(int p, int n) { return ((float)p + 0.5f) / (float)n; }

// a / b given y = 1 / b (the correctly rounded reciprocal), for quotients
// that are zero or normal with an exactly representable remainder, which is
// the case for every weighted mean and weight ratio below: q0 = a y lies
// within an ulp of a / b, the remainder a - b q0 is exact in an fma, and
// q0 + (a - b q0) y rounded once is the correctly rounded quotient
// (Markstein, IBM J. Res. Dev. 34(1), 1990).  A zero q0 is returned as is,
// keeping the sign that the correction's addition would drop.  Three VALU
// operations and a select per quotient, against about ten for a division,
// where several quotients share a divisor.
__device__ inline float div_rcp(float a, float b, float y)
{
    const float q0 = a * y;
    const float q = fmaf(fmaf(-b, q0, a), y, q0);
    return q0 == 0.0f ? q0 : q;
}

// vpt::normalize3 of the camera ray through (sx, sy), its three components
// divided by the length through one reciprocal (div_rcp), then the point at
// `dist` along it.
__device__ inline Vec3 position_from(const CamView& c, float sx, float sy, float dist)
{
    const Vec3 d = vpt::ray_dir_at(c.iv, c.ip, sx, sy);
    const float len = sqrtf((d.x * d.x + d.y * d.y) + d.z * d.z);
    const float rl = 1.0f / len;
    const Vec3 dir = vpt::mk3(div_rcp(d.x, len, rl), div_rcp(d.y, len, rl), div_rcp(d.z, len, rl));
    return vpt::mk3(c.iv[12] + dir.x * dist, c.iv[13] + dir.y * dist, c.iv[14] + dir.z * dist);
}

// ---- Spatial3x3Initial.main (Spatial3x3Initial.glsl:96-176) --------------
__global__ void __launch_bounds__(256) prespatial_kernel(CamView cam, int W, int H, const float* __restrict__ gb_t,
                                                         const uint8_t* __restrict__ gb_face,
                                                         const vpt_svgf_texel* __restrict__ in,
                                                         vpt_svgf_texel* __restrict__ out)
{
    __shared__ float s_unorm8[256];   // AO / sky bytes decoded through a table
    fill_unorm8_table(s_unorm8);
    __syncthreads();
    int px, py;
    if (!pixel_of(W, H, px, py)) return;
    const size_t i = (size_t)py * W + px;
    const float atrous[3] = {1.0f, 2.0f / 3.0f, 1.0f / 6.0f};
    const float tex_w = 1.0f / (float)W, tex_h = 1.0f / (float)H;
    float uvx = texc(px, W), uvy = texc(py, H);
    Vec3 base_pos = position_from(cam, uvx, uvy, gb_t[i]);
    Vec3 base_n = vpt::face_normal(gb_face[i], 1.0f);
    const uint4* in4 = reinterpret_cast<const uint4*>(in);
    Tx base = load_tx_word_lut(in4[i], s_unorm8);
    float base_luma = sh_to_y(base.sh[3]);
    float tsh[4] = {base.sh[0], base.sh[1], base.sh[2], base.sh[3]};
    float tco = base.co, tcg = base.cg, tw = 1.0f;
    float tao = base.ao, tsky = base.sky, taow = 1.0f;
    const float phi_color = 4.0f;
    for (int x = -1; x <= 1; x++) {
        for (int y = -1; y <= 1; y++) {
            if (x == 0 && y == 0) continue;
            float sx = uvx + ((float)x * 1.0f) * tex_w;
            float sy = uvy + ((float)y * 1.0f) * tex_h;
            if (!in_screen(sx, sy)) continue;
            size_t j = (size_t)(py + y) * W + (px + x);
            Vec3 sp = position_from(cam, sx, sy, gb_t[j]);
            Vec3 pd = vpt::mk3(fabsf(sp.x - base_pos.x), fabsf(sp.y - base_pos.y), fabsf(sp.z - base_pos.z));
            float dist_sqr = vpt::dot3(pd, pd);
            if (dist_sqr < 1.0f) {
                Tx s = load_tx_word_lut(in4[j], s_unorm8);
                Vec3 sn = vpt::face_normal(gb_face[j], 1.0f);
                float sluma = sh_to_y(s.sh[3]);
                float nw = vpt_powi(fmaxf(vpt::dot3(base_n, sn), 0.0f), 16);
                float lw = fabsf(sluma - base_luma) / phi_color;
                float w = vpt_expf(-lw - nw);
                w = fmaxf(w, 0.01f);
                w = (atrous[abs(x)] * atrous[abs(y)]) * w;
                w = fmaxf(w, 0.01f);
                w = fclamp(w, 0.0f, 1.0f);
                for (int k = 0; k < 4; k++) tsh[k] = tsh[k] + s.sh[k] * w;
                tco = tco + s.co * w;
                tcg = tcg + s.cg * w;
                tw = tw + w;
                tao = tao + s.ao * w;
                tsky = tsky + s.sky * w;
                taow = taow + w;
            }
        }
    }
    tw = fmaxf(tw, 0.01f);
    float osh[4];
    const float rtw = 1.0f / tw;
    for (int k = 0; k < 4; k++) osh[k] = div_rcp(tsh[k], tw, rtw);
    float aw = fmaxf(taow, 0.01f);
    const float raw = 1.0f / aw;
    store_tx(out, i, osh, div_rcp(tco, tw, rtw), div_rcp(tcg, tw, rtw), base.x, div_rcp(tao, aw, raw),
             div_rcp(tsky, aw, raw));
}

// ---- TemporalFilter.main (SVGF/TemporalFilter.glsl:132-360) --------------
__global__ void __launch_bounds__(256) temporal_kernel(CamView cam, Mat16 prev_vp, int W, int H,
                                                       const float* __restrict__ cur_t,
                                                       const uint8_t* __restrict__ cur_face,
                                                       const uint8_t* __restrict__ cur_block,
                                                       const float* __restrict__ prev_t,
                                                       const uint8_t* __restrict__ prev_face,
                                                       const uint8_t* __restrict__ prev_block,
                                                       const vpt_svgf_texel* __restrict__ cur,
                                                       const vpt_svgf_texel* __restrict__ hist,
                                                       const uint2* __restrict__ hist_util,
                                                       vpt_svgf_texel* __restrict__ out, uint2* __restrict__ out_util)
{
    __shared__ float s_unorm8[256];   // AO / sky bytes decoded through a table
    fill_unorm8_table(s_unorm8);
    __syncthreads();
    int px, py;
    if (!pixel_of(W, H, px, py)) return;
    const size_t i = (size_t)py * W + px;
    const float weights[5] = {3.0f / 32.0f, 3.0f / 32.0f, 9.0f / 64.0f, 3.0f / 32.0f, 3.0f / 32.0f};
    const float offx[5] = {1.0f, 0.0f, 0.0f, -1.0f, 0.0f};
    const float offy[5] = {0.0f, 1.0f, 0.0f, 0.0f, -1.0f};
    const float tex_w = 1.0f / (float)W, tex_h = 1.0f / (float)H;
    const float bth = 0.0035f;
    float uvx = texc(px, W), uvy = texc(py, H);
    float base_w = cur_t[i];
    Vec3 base_pos = position_from(cam, uvx, uvy, base_w);
    Vec3 base_n = vpt::face_normal(cur_face[i], 1.0f);
    Tx base = load_tx_word_lut(*reinterpret_cast<const uint4*>(cur + i), s_unorm8);

    float q0 = vpt::mat_row(prev_vp.m, 0, base_pos.x, base_pos.y, base_pos.z, 1.0f);
    float q1 = vpt::mat_row(prev_vp.m, 1, base_pos.x, base_pos.y, base_pos.z, 1.0f);
    float q3 = vpt::mat_row(prev_vp.m, 3, base_pos.x, base_pos.y, base_pos.z, 1.0f);
    float rx = (q0 / q3) * 0.5f + 0.5f;
    float ry = (q1 / q3) * 0.5f + 0.5f;
    float base_lum = base.x;
    int base_block = min(max((int)cur_block[i], 0), 127);

    bool do_block = true, do_normal = true;
    float tol = 0.75f;
    Vec3 dv = vpt::mk3(base_pos.x - cam.iv[12], base_pos.y - cam.iv[13], base_pos.z - cam.iv[14]);
    float dtp = sqrtf(vpt::dot3(dv, dv));
    if (dtp < 4.0f) { tol = 0.3f; }
    else if (dtp < 6.0f) { tol = 0.65f; }
    else if (dtp < 8.0f) { tol = 0.85f; }
    else if (dtp < 16.0f) { tol = 1.414f; }
    else if (dtp < 32.0f) { tol = 2.4f; }
    else if (dtp < 48.0f) { tol = 3.5f; do_block = false; }
    else if (dtp < 64.0f) { tol = 4.2f; do_block = false; }
    else if (dtp < 96.0f) { tol = 6.25f; do_block = false; do_normal = false; }
    else if (dtp < 128.0f) { tol = 9.0f; do_block = false; do_normal = false; }
    else if (dtp < 200.0f) { tol = 14.0f; do_block = false; do_normal = false; }

    float tw = 0.0f;
    float ssh[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    float sco = 0.0f, scg = 0.0f, sao = 0.0f, ssky = 0.0f, slum = 0.0f, sspp = 0.0f, smom = 0.0f;
    int successful = 0;
    for (int k = 0; k < 5; k++) {
        float sx = rx + (offx[k] + 0.0f) * tex_w;
        float sy = ry + (offy[k] + 0.0f) * tex_h;
        if (!(sx < 1.0f - bth && sx > bth && sy < 1.0f - bth && sy > bth)) continue;
        Bil f = bil_setup(sx, sy, W, H);
        float pw = bil(prev_t[f.i00], prev_t[f.i10], prev_t[f.i01], prev_t[f.i11], f.a, f.b);
        Vec3 pp = position_from(cam, sx, sy, pw);
        int ni = nearest_idx(sx, sy, W, H);
        Vec3 pn = vpt::face_normal(prev_face[ni], 1.0f);
        Vec3 pd = vpt::mk3(fabsf(base_pos.x - pp.x), fabsf(base_pos.y - pp.y), fabsf(base_pos.z - pp.z));
        float perr = vpt::dot3(pd, pd);
        int sblock = min(max((int)prev_block[ni], 0), 127);
        bool valid = false;
        if (perr < tol && ((pw < 0.0f) == (base_w < 0.0f))) {
            valid = true;
            if (do_normal && !(pn.x == base_n.x && pn.y == base_n.y && pn.z == base_n.z)) valid = false;
            if (do_block && base_block != sblock) valid = false;
        }
        if (valid) {
            float cw = weights[k];
            Tx h = bil_tx_lut(hist, f, s_unorm8);
            uint2 u00 = hist_util[f.i00], u10 = hist_util[f.i10], u01 = hist_util[f.i01], u11 = hist_util[f.i11];
            float spp = bil(h2f(u00.x & 0xffffu), h2f(u10.x & 0xffffu), h2f(u01.x & 0xffffu), h2f(u11.x & 0xffffu),
                            f.a, f.b);
            float mom = bil(h2f(u00.x >> 16), h2f(u10.x >> 16), h2f(u01.x >> 16), h2f(u11.x >> 16), f.a, f.b);
            float lum = bil(h2f(u00.y & 0xffffu), h2f(u10.y & 0xffffu), h2f(u01.y & 0xffffu), h2f(u11.y & 0xffffu),
                            f.a, f.b);
            for (int c = 0; c < 4; c++) ssh[c] = ssh[c] + h.sh[c] * cw;
            sco = sco + h.co * cw;
            scg = scg + h.cg * cw;
            sspp = sspp + spp * cw;
            smom = smom + mom * cw;
            slum = slum + lum * cw;
            sao = sao + h.ao * cw;
            ssky = ssky + h.sky * cw;
            tw = tw + cw;
            successful++;
        }
    }
    if (tw > 0.001f) {
        const float rtw = 1.0f / tw;
        for (int c = 0; c < 4; c++) ssh[c] = div_rcp(ssh[c], tw, rtw);
        sco = div_rcp(sco, tw, rtw);
        scg = div_rcp(scg, tw, rtw);
        smom = div_rcp(smom, tw, rtw);
        sspp = div_rcp(sspp, tw, rtw);
        slum = div_rcp(slum, tw, rtw);
        sao = div_rcp(sao, tw, rtw);
        ssky = div_rcp(ssky, tw, rtw);
    } else {
        successful = 0;
    }
    float spp_inc = sspp + 1.0f;
    if (successful <= 0) spp_inc = 0.01f;
    float blend = fmaxf(1.0f / spp_inc, 0.05f);
    float moment_factor = fmaxf(1.0f / spp_inc, 0.05f);
    float util_spp = successful <= 0 ? 0.0f : spp_inc;
    float util_moment = (1.0f - moment_factor) * smom + moment_factor * (base_lum * base_lum);
    float store_luma = slum * (1.0f - blend) + base_lum * blend;
    float osh[4];
    float oco, ocg, oao, osky;
    if (successful <= 0) {
        for (int c = 0; c < 4; c++) osh[c] = base.sh[c];
        oco = base.co;
        ocg = base.cg;
        oao = base.ao;
        osky = base.sky;
    } else {
        for (int c = 0; c < 4; c++) osh[c] = ssh[c] * (1.0f - blend) + base.sh[c] * blend;
        oco = sco * (1.0f - blend) + base.co * blend;
        ocg = scg * (1.0f - blend) + base.cg * blend;
        oao = sao * (1.0f - blend) + base.ao * blend;
        osky = ssky * (1.0f - blend) + base.sky * blend;
    }
    for (int c = 0; c < 4; c++) osh[c] = fclamp(osh[c], -100.0f, 100.0f);
    store_tx(out, i, osh, fclamp(oco, -10.0f, 100.0f), fclamp(ocg, -10.0f, 100.0f), 0.0f, fclamp(oao, 0.0f, 1.0f),
             fclamp(osky, 0.0f, 1.0f));
    uint32_t u0 = f2h(fclamp(util_spp, -150.0f, 150.0f)) | (f2h(fclamp(util_moment, -150.0f, 150.0f)) << 16);
    uint32_t u1 = f2h(fclamp(store_luma, -150.0f, 150.0f));
    out_util[i] = make_uint2(u0, u1);
}

// ---- VarianceEstimate.main (SVGF/VarianceEstimate.glsl:73-192) -----------
// The output texel also carries the temporal AO/sky, which the first a-trous
// pass reads (Pipeline.cpp:2531-2535).
// The variance pass stages its neighbourhood: a 32x8 block plus the 4-pixel
// apron of the 9x9 spatial estimate, decoded once into LDS (depth, normal
// index, SH, CoCg, luminance, moment) instead of once per tap.  A block in
// which no pixel has a short history skips the staging.  Taps are visited
// in the reference's order and combined with the same operations.
//
// Uniform blocks: when every staged on-screen cell holds the same bits
// (depth, normal index, SH, CoCg, luminance, moment; depth and luminance
// finite) -- the interior of the sky, whose pixels never accumulate
// history and take the 9x9 estimate every frame -- every tap has depth and
// luminance differences 0, so its weight is the normal weight of the cell's
// class against itself (w = w2 = that weight, >= 1, above both floors) and
// it adds the same increment to each accumulator.  The sums then depend only
// on the number of on-screen taps, kx * ky (the screen test is separable),
// and are read from a table of the sequential sums, k = 0..81, built with
// the same float additions.  The moment does not enter the weights: when
// only it differs between cells, the moment sum alone walks the taps.
constexpr int kVarR = 4;
constexpr int kVarSW = 32 + 2 * kVarR, kVarSH = 8 + 2 * kVarR, kVarCells = kVarSW * kVarSH;

__device__ inline int normal_index(uint32_t face) { return face <= 5u ? (int)face : 6; }

__global__ void __launch_bounds__(256) variance_kernel(int W, int H, const float* __restrict__ gb_t,
                                                       const uint8_t* __restrict__ gb_face,
                                                       const vpt_svgf_texel* __restrict__ in,
                                                       const uint2* __restrict__ util,
                                                       vpt_svgf_texel* __restrict__ out)
{
    __shared__ float4 s_sh[kVarCells];
    __shared__ float4 s_aux[kVarCells];   // co, cg, luminance, moment
    __shared__ float s_t[kVarCells];
    __shared__ uint8_t s_n[kVarCells];
    __shared__ float s_nw[7 * 8];         // normal weight by (base, sample) normal index
    // uniform blocks: tw, tmom, tsh[4], tco, tcg, tlum after k taps, in the
    // SH cells' storage (built after the last read of s_sh; a uniform block
    // reads no SH cell afterwards), so that six blocks fit in a CU's LDS
    static_assert(sizeof(float) * 9 * 82 <= sizeof(float4) * kVarCells, "s_tab fits in s_sh");
    float (*const s_tab)[82] = reinterpret_cast<float (*)[82]>(s_sh);
    const uint32_t tiles_x = (uint32_t)(W + 31) / 32u;
    const uint32_t tile = vpt::xcd_remap(blockIdx.x, tiles_x * ((uint32_t)(H + 7) / 8u));
    const int x0 = (int)(tile % tiles_x) * 32, y0 = (int)(tile / tiles_x) * 8;
    const int lx = (int)(threadIdx.x & 31), ly = (int)(threadIdx.x >> 5);
    const int px = x0 + lx, py = y0 + ly;
    const bool valid = px < W && py < H;
    const float thresh = 4.0f + 4.0f + 4.0f;
    const size_t i = valid ? (size_t)py * W + px : 0;
    float frames = 0.0f, base_moment = 0.0f;
    if (valid) {
        const uint2 bu = util[i];
        frames = h2f(bu.x & 0xffffu);
        base_moment = h2f(bu.x >> 16);
    }
    const bool spatial = valid && frames < thresh;
    bool uniform = false, uniform_m = false;
    if (__syncthreads_or(spatial)) {
        // the block's cells, three per thread (256 threads), read together:
        // an off-screen cell reads the nearest on-screen one and is never
        // read back from LDS
        constexpr int kPer = (kVarCells + 255) / 256;
        const int tid = (int)threadIdx.x;
        __builtin_assume(tid >= 0 && tid < 256);
        uint4 tw4[kPer];
        uint32_t um[kPer], fc[kPer];
        float tt[kPer];
#pragma unroll
        for (int k = 0; k < kPer; k++) {
            const int c = tid + 256 * k;
            if (c < kVarCells) {
                const int gx = min(max(x0 - kVarR + c % kVarSW, 0), W - 1);
                const int gy = min(max(y0 - kVarR + c / kVarSW, 0), H - 1);
                const size_t j = (size_t)gy * W + gx;
                tw4[k] = *reinterpret_cast<const uint4*>(in + j);
                um[k] = util[j].x;
                tt[k] = gb_t[j];
                fc[k] = gb_face[j];
            }
        }
#pragma unroll
        for (int k = 0; k < kPer; k++) {
            const int c = tid + 256 * k;
            if (c < kVarCells) {
                const Tx t = load_tx_word(tw4[k]);
                s_sh[c] = make_float4(t.sh[0], t.sh[1], t.sh[2], t.sh[3]);
                s_aux[c] = make_float4(t.co, t.cg, sh_to_y(t.sh[3]), h2f(um[k] >> 16));
                s_t[c] = tt[k];
                s_n[c] = (uint8_t)normal_index(fc[k]);
            }
        }
        if (threadIdx.x < 49) {
            const int a = (int)threadIdx.x / 7, b = (int)threadIdx.x % 7;
            const Vec3 na = vpt::face_normal(a == 6 ? 255 : a, 1.0f), nb = vpt::face_normal(b == 6 ? 255 : b, 1.0f);
            s_nw[a * 8 + b] = vpt_powi(fmaxf(vpt::dot3(na, nb), 0.0f), 16);
        }
        __syncthreads();
        // uniform block: every staged on-screen cell equals the block's
        // first pixel (always on screen) bit for bit
        const int ref = kVarR * kVarSW + kVarR;
        const float4 rs = s_sh[ref], ra = s_aux[ref];
        const float rt = s_t[ref];
        const uint8_t rn = s_n[ref];
        bool eq = isfinite(rt) && isfinite(ra.z), eq_m = true;
        for (int c = (int)threadIdx.x; c < kVarCells; c += (int)blockDim.x) {
            const int gx = x0 - kVarR + c % kVarSW, gy = y0 - kVarR + c / kVarSW;
            if (gx < 0 || gx >= W || gy < 0 || gy >= H) continue;
            const float4 cs = s_sh[c], ca = s_aux[c];
            eq = eq && __float_as_uint(s_t[c]) == __float_as_uint(rt) && s_n[c] == rn &&
                 __float_as_uint(cs.x) == __float_as_uint(rs.x) && __float_as_uint(cs.y) == __float_as_uint(rs.y) &&
                 __float_as_uint(cs.z) == __float_as_uint(rs.z) && __float_as_uint(cs.w) == __float_as_uint(rs.w) &&
                 __float_as_uint(ca.x) == __float_as_uint(ra.x) && __float_as_uint(ca.y) == __float_as_uint(ra.y) &&
                 __float_as_uint(ca.z) == __float_as_uint(ra.z);
            eq_m = eq_m && __float_as_uint(ca.w) == __float_as_uint(ra.w);
        }
        uniform = __syncthreads_and(eq);
        uniform_m = __syncthreads_and(eq_m);
        if (uniform) {
            if (threadIdx.x < 9) {
                const float nw0 = s_nw[rn * 8 + rn];
                const float src[9] = {1.0f, ra.w, rs.x, rs.y, rs.z, rs.w, ra.x, ra.y, ra.z};
                const int a = (int)threadIdx.x;
                const float inc = a == 0 ? nw0 : src[a] * nw0;
                float acc = 0.0f;
                s_tab[a][0] = acc;
                for (int k = 1; k <= 81; k++) {
                    acc = acc + inc;
                    s_tab[a][k] = acc;
                }
            }
            __syncthreads();
        }
    }
    if (!valid) return;
    const float tex_w = 1.0f / (float)W, tex_h = 1.0f / (float)H;
    float uvx = texc(px, W), uvy = texc(py, H);
    Tx base = load_tx(in, i);
    float base_lum = sh_to_y(base.sh[3]);
    float osh[4], oco, ocg;
    float variance;
    if (spatial) {
        const float color_phi = 5.0f;
        const float base_w = gb_t[i];
        const float* nw_row = s_nw + normal_index(gb_face[i]) * 8;
        float tw = 0.0f, tmom = 0.0f, tlum = 0.0f, tw2 = 0.0f;
        float tsh[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        float tco = 0.0f, tcg = 0.0f;
        if (uniform) {
            int kx = 0, ky = 0;
            for (int x = -4; x <= 4; x++) {
                const float sx = uvx + (float)x * tex_w;
                kx += (sx > 0.0f && sx < 1.0f) ? 1 : 0;
            }
            for (int y = -4; y <= 4; y++) {
                const float sy = uvy + (float)y * tex_h;
                ky += (sy > 0.0f && sy < 1.0f) ? 1 : 0;
            }
            const int k = kx * ky;
            tw = s_tab[0][k];
            tw2 = tw;
            if (uniform_m) {
                tmom = s_tab[1][k];
            } else {
                const float nw0 = nw_row[s_n[kVarR * kVarSW + kVarR]];
                for (int x = -4; x <= 4; x++) {
                    for (int y = -4; y <= 4; y++) {
                        float sx = uvx + (float)x * tex_w;
                        float sy = uvy + (float)y * tex_h;
                        if (!in_screen(sx, sy)) continue;
                        tmom = tmom + s_aux[(ly + kVarR + y) * kVarSW + (lx + kVarR + x)].w * nw0;
                    }
                }
            }
            for (int c = 0; c < 4; c++) tsh[c] = s_tab[2 + c][k];
            tco = s_tab[6][k];
            tcg = s_tab[7][k];
            tlum = s_tab[8][k];
        } else
        for (int x = -4; x <= 4; x++) {
            for (int y = -4; y <= 4; y++) {
                float sx = uvx + (float)x * tex_w;
                float sy = uvy + (float)y * tex_h;
                if (!in_screen(sx, sy)) continue;   // exactly the taps inside the image
                const int c = (ly + kVarR + y) * kVarSW + (lx + kVarR + x);
                const float sw = s_t[c];
                const float nw = nw_row[s_n[c]];
                const float4 a = s_aux[c];
                const float4 sh = s_sh[c];
                const float slum = a.z, smom = a.w;
                // vpt_expf of a non-positive argument lies in [0, 1] (checked
                // for every float), so a zero normal weight makes the product
                // exactly 0.  At a depth gap of 25 or more the product is below
                // exp(-50) * 3^16 < 1.5e-9 (3^16 is the largest normal weight,
                // sky against sky), under both clamps below, so it is taken as
                // 0 without evaluating the exps.  vpt_expf(-0) is exactly 1.
                float w = 0.0f;
                const float dd = fabsf(sw - base_w);
                if (nw != 0.0f && dd < 25.0f) {
                    float e = 1.0f;
                    if (dd != 0.0f) e = vpt_expf(-dd);
                    float dw = e * e;
                    float lw = fabsf(slum - base_lum) / color_phi;
                    float el = 1.0f;
                    if (lw != 0.0f) el = vpt_expf(-lw);
                    w = (el * nw) * dw;
                }
                float w2 = w;
                w = fmaxf(w, 0.000000015f);
                w2 = fmaxf(w2, 0.0000000015f);
                tw = tw + w;
                tmom = tmom + smom * w2;
                tsh[0] = tsh[0] + sh.x * w;
                tsh[1] = tsh[1] + sh.y * w;
                tsh[2] = tsh[2] + sh.z * w;
                tsh[3] = tsh[3] + sh.w * w;
                tco = tco + a.x * w;
                tcg = tcg + a.y * w;
                tlum = tlum + slum * w2;
                tw2 = tw2 + w2;
            }
        }
        if (tw > 0.0f) {
            const float rtw = 1.0f / tw, rtw2 = 1.0f / tw2;
            tmom = div_rcp(tmom, tw2, rtw2);
            tlum = div_rcp(tlum, tw2, rtw2);
            tco = div_rcp(tco, tw, rtw);
            tcg = div_rcp(tcg, tw, rtw);
            for (int k = 0; k < 4; k++) tsh[k] = div_rcp(tsh[k], tw, rtw);
        }
        for (int k = 0; k < 4; k++) osh[k] = tsh[k];
        oco = tco;
        ocg = tcg;
        float acc = tlum * tlum;
        variance = tmom - acc;
        variance = variance * 3.0f;
    } else {
        for (int k = 0; k < 4; k++) osh[k] = base.sh[k];
        oco = base.co;
        ocg = base.cg;
        variance = base_moment - base_lum * base_lum;
    }
    variance = variance * (thresh / frames);
    for (int k = 0; k < 4; k++) osh[k] = fclamp(osh[k], -100.0f, 100.0f);
    store_tx(out, i, osh, fclamp(oco, -10.0f, 100.0f), fclamp(ocg, -10.0f, 100.0f), fclamp(variance, -1.0f, 50.0f),
             base.ao, base.sky);
}

// GradientNoise (SpatialFilter.glsl:160-165).
__device__ inline float gradient_noise(int px, int py, float time)
{
    float t = time * 100.493850275f;
    float m = t - 500.0f * floorf(t / 500.0f);
    float cx = ((float)px + 0.5f) + m;
    float cy = ((float)py + 0.5f) + m;
    float inner = 0.06711056f * cx + 0.00583715f * cy;
    inner = inner - floorf(inner);
    float n = 52.9829189f * inner;
    return n - floorf(n);
}

// ---- SpatialFilter.main (SVGF/SpatialFilter.glsl:175-355), one pass ------
//
// The pass reads texels, depth and face at virtual texel coordinates
// (before GL REPEAT wrapping) through this source, addressed with 32-bit
// byte offsets from the planes' bases (vpt_svgf_create keeps a texel plane
// below 4 GiB), so that the gathers take a scalar base and one 32-bit
// offset register instead of 64-bit address arithmetic per tap.  (A variant
// that staged each tile's texel region in LDS gave identical results but ran
// 10-25% slower: the pass is bound by its per-tap VALU work, not by the
// gathers.)
struct GlobalSrc {
    const vpt_svgf_texel* in;
    const float* t;
    const uint8_t* face;
    int W, H;
    __device__ uint32_t idx(int vx, int vy) const { return (uint32_t)wrap1(vy, H) * (uint32_t)W + (uint32_t)wrap1(vx, W); }
    __device__ Tx tex_at(uint32_t j) const
    {
        return load_tx_word(*reinterpret_cast<const uint4*>(reinterpret_cast<const char*>(in) + j * 16u));
    }
    __device__ float depth_at(uint32_t j) const
    {
        return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(t) + j * 4u);
    }
    __device__ Tx tex(int vx, int vy) const { return tex_at(idx(vx, vy)); }
    __device__ float depth(int vx, int vy) const { return depth_at(idx(vx, vy)); }
    __device__ int fc(int vx, int vy) const { return face[idx(vx, vy)]; }
    __device__ float xch(int vx, int vy) const
    {
        return h2f(*reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(in) + idx(vx, vy) * 16u + 12u) &
                   0xffffu);
    }
};

// The same source with the AO / sky bytes decoded through the LDS table.
struct GlobalSrcLut : GlobalSrc {
    const float* unorm8;   // LDS
    __device__ Tx tex_at(uint32_t j) const
    {
        return load_tx_word_lut(*reinterpret_cast<const uint4*>(reinterpret_cast<const char*>(in) + j * 16u), unorm8);
    }
};

// Normal-weight class of a face byte: the six faces, or 6 for every other
// value (GetNormalFromID's (1, 1, 1)).
__device__ inline uint32_t face_class(int f) { return f <= 5 ? (uint32_t)f : 6u; }

// clamp(pow(max(dot(n_a, n_b), 0), 32), 0.001, 1) of two face normals is 1
// where the dot is positive (the same face; (1, 1, 1) against itself or
// against +x, +y, +z) and 0.001 elsewhere (dots 0 and -1): bit a * 8 + b of
// this mask.
constexpr uint64_t kNw32Mask = (1ull << (0 * 8 + 0)) | (1ull << (1 * 8 + 1)) | (1ull << (2 * 8 + 2)) |
                               (1ull << (3 * 8 + 3)) | (1ull << (4 * 8 + 4)) | (1ull << (5 * 8 + 5)) |
                               (1ull << (0 * 8 + 6)) | (1ull << (2 * 8 + 6)) | (1ull << (5 * 8 + 6)) |
                               (1ull << (6 * 8 + 0)) | (1ull << (6 * 8 + 2)) | (1ull << (6 * 8 + 5)) |
                               (1ull << (6 * 8 + 6));

__device__ inline float nw32_of(uint32_t base_class, uint32_t sample_class)
{
    return ((kNw32Mask >> (base_class * 8u + sample_class)) & 1ull) ? 1.0f : 0.001f;
}

// One pixel of the pass, reading through `src`.  kMode is the pass's AO /
// sky filtering as a compile-time value (0: both, step <= 4; 1: sky only,
// step <= 6; 2: neither), so that the passes that filter neither skip those
// two channels' taps entirely.  The 3 x 3 tap loops are unrolled, so the
// a-trous weights and offsets are constants.
template <class Src, int kMode>
__device__ __attribute__((always_inline)) inline void atrous_px(const Src& src, int W, int H, int px, int py,
                                                                const uint2* __restrict__ tutil, int step,
                                                                float scale, float color_phi_bias, float time,
                                                                vpt_svgf_texel* __restrict__ out)
{
    const uint32_t i = (uint32_t)py * (uint32_t)W + (uint32_t)px;
    const float atrous[3] = {1.0f, 2.0f / 3.0f, 1.0f / 6.0f};
    const float gauss[2] = {0.60283f, 0.198585f};
    const float tex_w = 1.0f / (float)W, tex_h = 1.0f / (float)H;
    constexpr bool filter_ao = kMode == 0;
    constexpr bool filter_sky = kMode <= 1;
    float uvx = texc(px, W), uvy = texc(py, H);
    int jit = (int)((gradient_noise(px, py, time) - 0.5f) * ((float)step * 0.8f));
    float base_depth = src.depth_at(i);
    const uint32_t base_class = face_class(src.face[i]);
    Tx base = src.tex_at(i);
    float base_lum = sh_to_y(base.sh[3]);

    // GaussianVariance (:94-129) over the previous pass's variance.  The
    // eight neighbours are read first (their wrapped indices are always in
    // the image), so the reads are in flight together; the sums then take
    // the in-screen ones, in the same order.
    float nv[9];
#pragma unroll
    for (int x = -1; x <= 1; x++)
#pragma unroll
        for (int y = -1; y <= 1; y++) nv[(x + 1) * 3 + y + 1] = (x == 0 && y == 0) ? base.x : src.xch(px + x, py + y);
    float base_var = 0.0f, vsum = 0.0f, ksum = 0.0f;
#pragma unroll
    for (int x = -1; x <= 1; x++) {
#pragma unroll
        for (int y = -1; y <= 1; y++) {
            const bool in = in_screen(uvx + (float)x * tex_w, uvy + (float)y * tex_h);
            const float kv = gauss[abs(x)] * gauss[abs(y)];
            const float v = nv[(x + 1) * 3 + y + 1];
            if (x == 0 && y == 0) base_var = in ? v : base_var;
            vsum = in ? vsum + v * kv : vsum;
            ksum = in ? ksum + kv : ksum;
        }
    }
    float var_est = vsum / fmaxf(ksum, 0.01f);

    float tsh[4] = {base.sh[0], base.sh[1], base.sh[2], base.sh[3]};
    float tco = base.co, tcg = base.cg, tw = 1.0f, tvar = base_var;
    float tao = base.ao, tsky = base.sky, taow = 1.0f;
    float frames = h2f(tutil[i].x & 0xffffu);
    bool strong = frames <= 8.0f && step <= 8;
    float curve = 0.0f;
    if (var_est < 0.01f) curve = 128.0f;
    else if (var_est < 0.025f) curve = 112.0f;
    else if (var_est < 0.05f) curve = 96.0f;
    else if (var_est < 0.075f) curve = 84.0f;
    else if (var_est < 0.1f) curve = 70.0f;
    float tweaked;
    if (var_est < 0.1f) {
        float fv = fclamp(var_est, 0.0f, 1.0f);
        tweaked = fv * vpt_powi(1.0f - fv, (int)(curve + 6.0f));
    } else {
        tweaked = var_est;
    }
    float phi = sqrtf(fmaxf(0.0f, 0.000001f + tweaked));
    phi = phi / fmaxf(color_phi_bias, 0.1f);
    const float rphi = 1.0f / phi;

    if (!(base_depth < 0.0f)) {
        const float jh = (float)jit * 0.5f;
#pragma unroll
        for (int x = -1; x <= 1; x++) {
#pragma unroll
            for (int y = -1; y <= 1; y++) {
                if (x == 0 && y == 0) continue;
                float ox = ((float)x * (float)step) * scale + jh;
                float oy = ((float)y * (float)step) * scale + jh;
                float sx = uvx + ox * tex_w;
                float sy = uvy + oy * tex_h;
                if (!in_screen(sx, sy)) continue;
                // GL LINEAR footprint (as bil_setup), at virtual coordinates
                const float u = sx * (float)W - 0.5f, v = sy * (float)H - 0.5f;
                const float fu = floorf(u), fv = floorf(v);
                const float a = u - fu, b = v - fv;
                const int x0 = (int)fu, y0 = (int)fv;
                const uint32_t c0 = (uint32_t)wrap1(x0, W), c1 = (uint32_t)wrap1(x0 + 1, W);
                const uint32_t r0 = (uint32_t)wrap1(y0, H) * (uint32_t)W, r1 = (uint32_t)wrap1(y0 + 1, H) * (uint32_t)W;
                const uint32_t j00 = r0 + c0, j10 = r0 + c1, j01 = r1 + c0, j11 = r1 + c1;
                float sdepth = bil(src.depth_at(j00), src.depth_at(j10), src.depth_at(j01), src.depth_at(j11), a, b);
                float ddiff = fabsf(sdepth - base_depth);
                const float nw = nw32_of(base_class,
                                         face_class(src.fc((int)floorf(sx * (float)W), (int)floorf(sy * (float)H))));
                const Tx q00 = src.tex_at(j00), q10 = src.tex_at(j10), q01 = src.tex_at(j01), q11 = src.tex_at(j11);
                Tx s;
                for (int c = 0; c < 4; c++) s.sh[c] = bil(q00.sh[c], q10.sh[c], q01.sh[c], q11.sh[c], a, b);
                s.co = bil(q00.co, q10.co, q01.co, q11.co, a, b);
                s.cg = bil(q00.cg, q10.cg, q01.cg, q11.cg, a, b);
                s.x = bil(q00.x, q10.x, q01.x, q11.x, a, b);
                s.ao = bil(q00.ao, q10.ao, q01.ao, q11.ao, a, b);
                s.sky = bil(q00.sky, q10.sky, q01.sky, q11.sky, a, b);
                float slum = sh_to_y(s.sh[3]);
                float lw = div_rcp(fabsf(slum - base_lum), phi, rphi);
                // exp(-d)^2 < 1e-4 for every float d >= 4.7, so the depth
                // weight sits on its lower clamp there without the exp
                float dw = 0.0001f;
                if (ddiff < 4.7f) {
                    float e = vpt_expf(-fmaxf(ddiff, 0.00001f));
                    dw = fclamp(e * e, 0.0001f, 1.0f);
                }
                // vpt_expf(-lw) <= 1, so (exp * nw) * dw <= nw * dw: once that
                // is at most 0.001 the weight sits on its lower clamp
                float w = nw * dw;
                if (!strong && w > 0.001f) w = (vpt_expf(-lw) * nw) * dw;
                w = fclamp(w, 0.001f, 1.0f);
                float xw = atrous[abs(x)], yw = atrous[abs(y)];
                w = (xw * yw) * w;
                w = fmaxf(w, 0.00000001f);
                for (int c = 0; c < 4; c++) tsh[c] = tsh[c] + s.sh[c] * w;
                tco = tco + s.co * w;
                tcg = tcg + s.cg * w;
                tvar = tvar + (w * w) * s.x;
                tw = tw + w;
                if (filter_sky || filter_ao) {
                    float aow = fclamp(((xw * yw) * nw) * dw, 0.000001f, 1.0f);
                    tao = tao + s.ao * aow;
                    tsky = tsky + s.sky * aow;
                    taow = taow + aow;
                }
            }
        }
    }
    float osh[4];
    const float rtw = 1.0f / tw, rtaow = 1.0f / taow;
    for (int c = 0; c < 4; c++) osh[c] = fclamp(div_rcp(tsh[c], tw, rtw), -100.0f, 100.0f);
    float oao = div_rcp(tao, taow, rtaow), osky = div_rcp(tsky, taow, rtaow);
    if (!filter_ao) oao = base.ao;
    store_tx(out, i, osh, fclamp(div_rcp(tco, tw, rtw), -10.0f, 100.0f), fclamp(div_rcp(tcg, tw, rtw), -10.0f, 100.0f),
             fclamp(tvar / (tw * tw), -1.0f, 50.0f), fclamp(oao, 0.0f, 1.0f), fclamp(osky, 0.0f, 1.0f));
}

template <int kMode>
__global__ void __launch_bounds__(256) atrous_kernel(int W, int H, const float* __restrict__ gb_t,
                                                     const uint8_t* __restrict__ gb_face,
                                                     const vpt_svgf_texel* __restrict__ in,
                                                     const uint2* __restrict__ tutil, int step, float scale,
                                                     float color_phi_bias, float time,
                                                     vpt_svgf_texel* __restrict__ out)
{
    int px, py;
    if constexpr (kMode == 2) {
        if (!pixel_of(W, H, px, py)) return;
        const GlobalSrc src{in, gb_t, gb_face, W, H};
        atrous_px<GlobalSrc, kMode>(src, W, H, px, py, tutil, step, scale, color_phi_bias, time, out);
    } else {
        // the passes that filter AO or sky decode their bytes through a table
        __shared__ float s_unorm8[256];
        fill_unorm8_table(s_unorm8);
        __syncthreads();
        if (!pixel_of(W, H, px, py)) return;
        GlobalSrcLut src;
        static_cast<GlobalSrc&>(src) = GlobalSrc{in, gb_t, gb_face, W, H};
        src.unorm8 = s_unorm8;
        atrous_px<GlobalSrcLut, kMode>(src, W, H, px, py, tutil, step, scale, color_phi_bias, time, out);
    }
}

// ---- SHToIrridiance(shY, CoCg, v) (ColorPassFrag.glsl:356-368) ------------
__global__ void __launch_bounds__(256) decode_kernel(int W, int H, const uint8_t* __restrict__ gb_face,
                                                     const vpt_svgf_texel* __restrict__ in, float* __restrict__ rgb)
{
    int px, py;
    if (!pixel_of(W, H, px, py)) return;
    const size_t i = (size_t)py * W + px;
    Tx t = load_tx(in, i);
    Vec3 v = vpt::face_normal(gb_face[i], 1.0f);
    float x = (t.sh[0] * v.x + t.sh[1] * v.y) + t.sh[2] * v.z;
    float Y = 2.0f * (1.023326f * x + 0.886226f * t.sh[3]);
    Y = fmaxf(Y, 0.0f);
    float k = (Y * 0.282095f) / (t.sh[3] + 1e-6f);
    float co = t.co * k, cg = t.cg * k;
    float T = Y - cg * 0.5f;
    float G = cg + T;
    float B = T - co * 0.5f;
    float R = B + co;
    rgb[3 * i] = fmaxf(R, 0.0f);
    rgb[3 * i + 1] = fmaxf(G, 0.0f);
    rgb[3 * i + 2] = fmaxf(B, 0.0f);
}

__global__ void fill_u32_kernel(uint32_t* p, size_t n, uint32_t v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---- Sun-shadow denoise: ShadowTemporalFilter + ShadowFilter ---------------
// Planes: raw mask u8 0/1 (o_Shadow), transversal binary16
// (o_IntersectionTransversal), temporal colour u8 unorm and frame count
// binary16 (Pipeline.cpp:1152-1154).  texture() at the pixel's own centre or
// at whole-texel offsets from it reads the texel; other coordinates are GL
// LINEAR + REPEAT; normal ids are NEAREST.  GLSL pow with an integer-valued
// literal exponent is vpt_powi, with a run-time exponent vpt_powf.

__device__ inline int wrapn(int i, int n)
{
    if (i >= 0 && i < n) return i;
    if (i < 0 && i >= -n) return i + n;
    if (i >= n && i < 2 * n) return i - n;
    const int r = i % n;   // taps further out than one image (tiny images, large scales)
    return r < 0 ? r + n : r;
}

// GL LINEAR + REPEAT footprint for any coordinate (the shadow filter's taps
// reach several texels past the image).
__device__ inline Bil bil_setup_any(float s, float t, int W, int H)
{
    float u = s * (float)W - 0.5f;
    float v = t * (float)H - 0.5f;
    float fu = floorf(u), fv = floorf(v);
    Bil b;
    b.a = u - fu;
    b.b = v - fv;
    const int x0 = wrapn((int)fu, W), x1 = wrapn((int)fu + 1, W);
    const int y0 = wrapn((int)fv, H), y1 = wrapn((int)fv + 1, H);
    const uint32_t r0 = (uint32_t)y0 * (uint32_t)W, r1 = (uint32_t)y1 * (uint32_t)W;   // W * H < 2^28
    b.i00 = r0 + (uint32_t)x0;
    b.i10 = r0 + (uint32_t)x1;
    b.i01 = r1 + (uint32_t)x0;
    b.i11 = r1 + (uint32_t)x1;
    return b;
}

__device__ inline float dec8(uint8_t c) { return vpt_unorm8_decode(c); }

// x / 3 for 0 <= x <= 2: q0 = x fl(1/3), then one correction with the exact
// remainder x - 3 q0 (an fma).  Equal to the IEEE division for every float
// in [0, 2], subnormals included (tests/test_exact_division_cpu.py checks
// all of them); three VALU operations in place of the division's ten.
__device__ inline float div3_unit(float x)
{
    const float y = 1.0f / 3.0f;
    const float q0 = x * y;
    return fmaf(fmaf(-3.0f, q0, x), y, q0);
}

// (m - 1) / (m + 1) for the m that vpt_logf's reduction leaves, [sqrt(2) / 2,
// sqrt(2)]: v_rcp_f32, one Newton step for the reciprocal, then one
// correction with the exact remainder.  Equal to the IEEE division for
// every float m in that range (vpt_check_log_division compares all
// 8,388,609 of them on the device), in six VALU operations instead of ten.
__device__ inline float log_ratio(float m)
{
    const float a = m - 1.0f, b = m + 1.0f;
    const float y0 = __builtin_amdgcn_rcpf(b);
    const float y = fmaf(fmaf(-b, y0, 1.0f), y0, y0);
    const float q0 = a * y;
    return fmaf(fmaf(-b, q0, a), y, q0);
}

// vpt_logf and vpt_powf with log_ratio in place of their one division:
// the same operations otherwise, so the same results.
__device__ inline float logf_rcp(float x)
{
    if (!(x == x)) return x;
    if (x < 0.0f) return vpt_u2f(0x7fc00000u);
    if (x == 0.0f) return vpt_u2f(0xff800000u);
    if (x == vpt_u2f(0x7f800000u)) return x;
    uint32_t u = vpt_f2u(x);
    int e = (int)((u >> 23) & 0xffu) - 127;
    if (e == -127) {
        u = vpt_f2u(x * 8388608.0f);
        e = (int)((u >> 23) & 0xffu) - 127 - 23;
    }
    float m = vpt_u2f((u & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421353816986083984375f) {
        m = m * 0.5f;
        e = e + 1;
    }
    const float LN2_HI = 0.693145751953125f;
    const float LN2_LO = 1.428606765330187045e-06f;
    float s = log_ratio(m);
    float s2 = s * s;
    float p = 0.111111111111f;
    p = p * s2 + 0.142857142857f;
    p = p * s2 + 0.2f;
    p = p * s2 + 0.333333333333f;
    float lm = (s + s) + (s + s) * (s2 * p);
    float fe = (float)e;
    return fe * LN2_HI + (fe * LN2_LO + lm);
}

__device__ inline float powf_rcp(float x, float y)
{
    if (y == 0.0f) return 1.0f;
    if (x == 0.0f) return y > 0.0f ? 0.0f : vpt_u2f(0x7f800000u);
    if (x == 1.0f) return 1.0f;
    return expf_nonpos(y * logf_rcp(x));
}

// Every float m in [sqrt(2) / 2, sqrt(2)]: log_ratio(m) against the division.
__global__ void log_division_check_kernel(uint32_t u0, uint32_t n, unsigned long long* __restrict__ bad)
{
    const uint32_t k = blockIdx.x * 256u + threadIdx.x;
    if (k >= n) return;
    const float m = __uint_as_float(u0 + k);
    if (__float_as_uint(log_ratio(m)) != __float_as_uint((m - 1.0f) / (m + 1.0f))) atomicAdd(bad, 1ull);
}

// GetShadowSpatial (ShadowTemporalFilter.glsl:98-150) over the raw mask.
__device__ inline float shadow_spatial_pre(int px, int py, int W, int H, float tva, float base_depth,
                                           const float* __restrict__ cur_t, const uint8_t* __restrict__ cur_face,
                                           const uint8_t* __restrict__ mask)
{
    const float unit_diagonal = sqrtf(2.0f);
    if (tva <= unit_diagonal * 2.0f) return 1.0f;
    const size_t i = (size_t)py * W + px;
    const float tex_w = 1.0f / (float)W, tex_h = 1.0f / (float)H;
    const float uvx = texc(px, W), uvy = texc(py, H);
    float total = (float)mask[i];
    const float base = total;
    float weight = 1.0f;
    const Vec3 bn = vpt::face_normal(cur_face[i], 1.0f);
    const float b = 0.03f;
    // the eight neighbours are read first (clamped into the image; a tap off
    // the screen margin is skipped below), so the reads are in flight together
    float nd[9];
    uint32_t nf[9], nm[9];
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const int x = k / 3 - 1, y = k % 3 - 1;
        const size_t j = (size_t)min(max(py + y, 0), H - 1) * W + min(max(px + x, 0), W - 1);
        nd[k] = cur_t[j];
        nf[k] = cur_face[j];
        nm[k] = mask[j];
    }
#pragma unroll
    for (int x = -1; x <= 1; x++) {
#pragma unroll
        for (int y = -1; y <= 1; y++) {
            if (x == 0 && y == 0) continue;
            const float sx = uvx + (float)x * tex_w, sy = uvy + (float)y * tex_h;
            if (!(sx > b && sx < 1.0f - b && sy > b && sy < 1.0f - b)) continue;
            const int k = (x + 1) * 3 + y + 1;
            const float sd = nd[k];
            const Vec3 sn = vpt::face_normal((uint8_t)nf[k], 1.0f);
            if (sn.x == bn.x && sn.y == bn.y && sn.z == bn.z && fabsf(sd - base_depth) < 1.0f) {
                const float smp = (float)nm[k];
                float wa = fclamp(1.0f - fclamp(fabsf(smp - base) / 3.0f, 0.0f, 1.0f), 0.0f, 1.0f);
                wa = fclamp(vpt_powi(wa, 7), 0.000001f, 1.0f);
                total = total + smp * wa;
                weight = weight + wa;
            }
        }
    }
    return total / weight;
}

// ShadowTemporalFilter.main (ShadowTemporalFilter.glsl:187-267), u_ShadowTemporal true.
__global__ void __launch_bounds__(256) shadow_temporal_kernel(CamView cam, Mat16 prev_vp, int W, int H,
                                                              const float* __restrict__ cur_t,
                                                              const uint8_t* __restrict__ cur_face,
                                                              const float* __restrict__ prev_t,
                                                              const uint8_t* __restrict__ mask,
                                                              const uint16_t* __restrict__ tv,
                                                              const uint8_t* __restrict__ hist_color,
                                                              const uint16_t* __restrict__ hist_frames,
                                                              uint8_t* __restrict__ out_color,
                                                              uint16_t* __restrict__ out_frames)
{
    int px, py;
    if (!pixel_of(W, H, px, py)) return;
    const size_t i = (size_t)py * W + px;
    const float uvx = texc(px, W), uvy = texc(py, H);
    const float dist = cur_t[i];
    float oc, of;
    if (dist > 0.0f) {
        const Vec3 pos = position_from(cam, uvx, uvy, dist);
        const float q0 = vpt::mat_row(prev_vp.m, 0, pos.x, pos.y, pos.z, 1.0f);
        const float q1 = vpt::mat_row(prev_vp.m, 1, pos.x, pos.y, pos.z, 1.0f);
        const float q3 = vpt::mat_row(prev_vp.m, 3, pos.x, pos.y, pos.z, 1.0f);
        const float rx = (q0 / q3) * 0.5f + 0.5f, ry = (q1 / q3) * 0.5f + 0.5f;
        const float tva = h2f(tv[i]) * 100.0f;
        float cur = shadow_spatial_pre(px, py, W, H, tva, dist, cur_t, cur_face, mask);
        const float bias = 0.005f;
        const bool rejected = !(rx > 0.0f + bias && rx < 1.0f - bias && ry > 0.0f + bias && ry < 1.0f - bias);
        if (!rejected) {
            const Bil f = bil_setup(rx, ry, W, H);
            const float prev_orig =
                bil(dec8(hist_color[f.i00]), dec8(hist_color[f.i10]), dec8(hist_color[f.i01]), dec8(hist_color[f.i11]),
                    f.a, f.b);
            float prev = prev_orig;
            if (tva < 1.414f * 3.0f) {   // ClipShadow (:166-185) and clipAABB (:156-164)
                const int ox[5] = {-1, 1, 0, 0, 0}, oy[5] = {0, 0, 0, -1, 1};
                float mn = 100.0f, mx = -100.0f;
#pragma unroll
                for (int k = 0; k < 5; k++) {
                    const float smp = (float)mask[(size_t)wrap1(py + oy[k], H) * W + wrap1(px + ox[k], W)];
                    mn = fminf(smp, mn);
                    mx = fmaxf(smp, mx);
                }
                const float lo = mn - 0.025f, hi = mx + 0.025f;
                const float pcl = 0.5f * (hi + lo), ecl = 0.5f * (hi - lo);
                const float vcl = prev_orig - pcl;
                const float denom = fabsf(vcl / ecl);
                prev = denom > 1.0f ? pcl + vcl / denom : prev_orig;
            }
            const float pw = bil(prev_t[f.i00], prev_t[f.i10], prev_t[f.i01], prev_t[f.i11], f.a, f.b);
            const Vec3 pp = position_from(cam, rx, ry, pw);
            const Vec3 dd = vpt::mk3(pp.x - pos.x, pp.y - pos.y, pp.z - pos.z);
            const float d = sqrtf(vpt::dot3(dd, dd));
            cur = fclamp(cur, 0.0f, 1.0f);
            prev = fclamp(prev, 0.0f, 1.0f);
            const float vx = (uvx - rx) * (float)W, vy = (uvy - ry) * (float)H;
            const float clip_err = fabsf(prev_orig - prev);
            const float inc = clip_err < 0.2f ? 1.0f : 0.6f;
            const float fc = bil(h2f(hist_frames[f.i00]), h2f(hist_frames[f.i10]), h2f(hist_frames[f.i01]),
                                 h2f(hist_frames[f.i11]), f.a, f.b);
            const float fi = fc + inc;
            float bf = fclamp((1.0f - (1.0f / fi)) * 1.2f, 0.01f, 0.97f);
            const float vr = fclamp(expf_nonpos(-sqrtf(vx * vx + vy * vy)) * 0.8f + 0.6f, 0.00000001f, 1.0f);
            bf = bf * vr;
            float dr = 1.0f;
            if (d > 0.4f) {
                dr = vpt_powi(expf_nonpos(-d), 48);
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
    out_color[i] = vpt_unorm8_encode(oc);
    out_frames[i] = (uint16_t)f2h(of);
}

// ShadowFilter.main / ShadowSpatial (ShadowFilter.glsl:69-165).
//
// A block filters a 32x8 tile.  Its neighbourhood -- the tile plus an
// 11-pixel apron, which holds every bilinear footprint and nearest tap of
// the 7x7 kernel at filter scales up to 1 (offsets up to 3 * 1.2 * 2.6 =
// 9.36 pixels) -- is staged once in LDS: depth and the decoded colour per
// cell, and the face's normal class.  A tap whose footprint leaves the
// staged cells or wraps at the image edge (REPEAT) reads global memory;
// both paths read the same values.  The normal weight pow(max(dot, 1e-9),
// 32) of two face normals takes one of three values and comes from a 7x7
// table built with the same expression.
constexpr int kSfA = 11;
constexpr int kSfW = 32 + 2 * kSfA, kSfH = 8 + 2 * kSfA, kSfCells = kSfW * kSfH;

__global__ void __launch_bounds__(256) shadow_filter_kernel(int W, int H, const float* __restrict__ gb_t,
                                                            const uint8_t* __restrict__ gb_face,
                                                            const uint8_t* __restrict__ color,
                                                            const uint16_t* __restrict__ tv,
                                                            const uint16_t* __restrict__ frames, float scale,
                                                            uint8_t* __restrict__ out)
{
    __shared__ float2 s_dc[kSfCells];   // depth, decoded colour
    __shared__ uint8_t s_f[kSfCells];   // normal class (face, or 6)
    __shared__ float s_nw[7 * 8];
    const uint32_t tiles_x = (uint32_t)(W + 31) / 32u;
    const uint32_t tile = vpt::xcd_remap(blockIdx.x, tiles_x * ((uint32_t)(H + 7) / 8u));
    const int tx0 = (int)(tile % tiles_x) * 32, ty0 = (int)(tile / tiles_x) * 8;
    const int gx0 = tx0 - kSfA, gy0 = ty0 - kSfA;   // image position of staged cell (0, 0)
    {
        // the staged cells, seven per thread (256 threads), read together: a
        // cell off the image reads the nearest one on it and is never read
        // back from LDS
        constexpr int kPer = (kSfCells + 255) / 256;
        const int tid = (int)threadIdx.x;
        __builtin_assume(tid >= 0 && tid < 256);
        float st[kPer];
        uint32_t sc[kPer], sf[kPer];
#pragma unroll
        for (int k = 0; k < kPer; k++) {
            const int c = tid + 256 * k;
            if (c < kSfCells) {
                const int gx = min(max(gx0 + c % kSfW, 0), W - 1), gy = min(max(gy0 + c / kSfW, 0), H - 1);
                const uint32_t j = (uint32_t)gy * (uint32_t)W + (uint32_t)gx;
                st[k] = gb_t[j];
                sc[k] = color[j];
                sf[k] = gb_face[j];
            }
        }
#pragma unroll
        for (int k = 0; k < kPer; k++) {
            const int c = tid + 256 * k;
            if (c < kSfCells) {
                s_dc[c] = make_float2(st[k], dec8((uint8_t)sc[k]));
                s_f[c] = (uint8_t)face_class((int)sf[k]);
            }
        }
    }
    if (threadIdx.x < 49) {
        const int a = (int)threadIdx.x / 7, b = (int)threadIdx.x % 7;
        const Vec3 na = vpt::face_normal(a == 6 ? 255 : a, 1.0f), nb = vpt::face_normal(b == 6 ? 255 : b, 1.0f);
        s_nw[a * 8 + b] = vpt_powi(fmaxf(vpt::dot3(na, nb), 0.000000001f), 32);
    }
    __syncthreads();
    const int px = tx0 + (int)(threadIdx.x & 31), py = ty0 + (int)(threadIdx.x >> 5);
    if (px >= W || py >= H) return;
    const uint32_t i = (uint32_t)py * (uint32_t)W + (uint32_t)px;
    const float uvx = texc(px, W), uvy = texc(py, H);
    const float tex_w = 1.0f / (float)W, tex_h = 1.0f / (float)H;
    const float fr = h2f(frames[i]);
    const bool apply_luma = fr > 7.5f;
    const float dist = gb_t[i];
    const bool sky = dist < 0.0f;
    const float* nw_row = s_nw + face_class(gb_face[i]) * 8;
    const float cs = dec8(color[i]);
    const float T = h2f(tv[i]) * 100.0f;
    const float cutoff = sqrtf(2.0f);
    if ((T > 0.0f && T < cutoff) || sky) {
        out[i] = vpt_unorm8_encode(cs);
        return;
    }
    const int size = T < cutoff * 1.414f ? 1 : 3;
    float sc = 1.0f;
    if (T > 6.0f) sc = 2.0f;
    if (T > 16.0f) sc = 2.4f;
    if (T > 32.0f) sc = 2.6f;
    const float ct = fclamp(T, 0.0f, 10.0f);
    const float am = ct / 10.0f;
    float ve = (20.0f * (1.0f - am) + 6.0f * am) + (T < 6.0f ? 5.0f : 2.0f);
    ve = fclamp(ve - 1.75f, 0.0000001f, 64.0f);
    float lm = 1.0f;
    if (!apply_luma) {
        const float fn = fr / 7.5f;
        lm = 0.1f * (1.0f - fn) + 0.5f * fn;
    }
    const float expo = (ve * lm) * 0.9f;
    float tw = 0.0f, ts = 0.0f;
    for (int x = -size; x <= size; x++) {
        for (int y = -size; y <= size; y++) {
            const float sx = uvx + (((float)x * tex_w) * 1.2f * sc) * scale;
            const float sy = uvy + (((float)y * tex_h) * 1.2f * sc) * scale;
            // GL LINEAR footprint (as bil_setup_any) and NEAREST texel
            const float u = sx * (float)W - 0.5f, v = sy * (float)H - 0.5f;
            const float fu = floorf(u), fv = floorf(v);
            const float fa = u - fu, fb = v - fv;
            const int x0 = (int)fu, y0 = (int)fv;
            const int nxi = (int)floorf(sx * (float)W), nyi = (int)floorf(sy * (float)H);
            const int lx = x0 - gx0, ly = y0 - gy0, nlx = nxi - gx0, nly = nyi - gy0;
            const bool staged = x0 >= 0 && x0 + 1 < W && y0 >= 0 && y0 + 1 < H && lx >= 0 && lx + 1 < kSfW &&
                                ly >= 0 && ly + 1 < kSfH && nxi >= 0 && nxi < W && nyi >= 0 && nyi < H &&
                                nlx >= 0 && nlx < kSfW && nly >= 0 && nly < kSfH;
            float d00, d10, d01, d11, c00, c10, c01, c11, nw;
            if (staged) {
                const int c = ly * kSfW + lx;
                const float2 q00 = s_dc[c], q10 = s_dc[c + 1], q01 = s_dc[c + kSfW], q11 = s_dc[c + kSfW + 1];
                d00 = q00.x; d10 = q10.x; d01 = q01.x; d11 = q11.x;
                c00 = q00.y; c10 = q10.y; c01 = q01.y; c11 = q11.y;
                nw = nw_row[s_f[nly * kSfW + nlx]];
            } else {
                const uint32_t wx0 = (uint32_t)wrapn(x0, W), wx1 = (uint32_t)wrapn(x0 + 1, W);
                const uint32_t r0 = (uint32_t)wrapn(y0, H) * (uint32_t)W, r1 = (uint32_t)wrapn(y0 + 1, H) * (uint32_t)W;
                d00 = gb_t[r0 + wx0]; d10 = gb_t[r0 + wx1]; d01 = gb_t[r1 + wx0]; d11 = gb_t[r1 + wx1];
                c00 = dec8(color[r0 + wx0]); c10 = dec8(color[r0 + wx1]);
                c01 = dec8(color[r1 + wx0]); c11 = dec8(color[r1 + wx1]);
                const uint32_t ni = (uint32_t)wrapn(nyi, H) * (uint32_t)W + (uint32_t)wrapn(nxi, W);
                nw = nw_row[face_class(gb_face[ni])];
            }
            const float sa = bil(c00, c10, c01, c11, fa, fb);
            // The weight is ((clamp(pow) * dw) * nw) with pow <= 1 (vpt_expf of a
            // non-positive argument, checked for every float), so it sits on
            // its 1e-9 clamp once dw * nw < 1e-9: nw is exactly 0 or 1 here
            // (axis normals, or (1,1,1) against an axis normal), and
            // exp(-d)^3 < 1e-9 for every float d >= 21.  Those taps skip the
            // exp and the pow.
            float w = 0.000000001f;
            if (nw != 0.0f) {
                const float sd = bil(d00, d10, d01, d11, fa, fb);
                const float dd = fabsf(dist - sd);
                if (dd < 21.0f) {
                    const float dw = vpt_powi(expf_nonpos(-dd), 3);
                    if (dw * nw >= 0.000000001f) {
                        const float le = fclamp(1.0f - fclamp(div3_unit(fabsf(sa - cs)), 0.0f, 1.0f), 0.0f, 1.0f);
                        w = 1.0f;
                        w = w * fclamp(powf_rcp(le, expo), 0.0f, 1.0f);
                        w = w * dw;
                        w = w * nw;
                        w = fclamp(w, 0.000000001f, 1.0f);
                    }
                }
            }
            ts = ts + sa * w;
            tw = tw + w;
        }
    }
    out[i] = vpt_unorm8_encode(ts / fmaxf(tw, 0.01f));
}

int fill_zero(void* p, size_t bytes, hipStream_t s)
{
    size_t n = bytes / 4;
    hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (uint32_t*)p, n, 0u);
    HIP_TRY(hipGetLastError());
    return VPT_OK;
}

CamView make_cam(const vpt_camera* c)
{
    CamView v;
    for (int i = 0; i < 16; i++) {
        v.iv[i] = c->inv_view[i];
        v.ip[i] = c->inv_proj[i];
    }
    return v;
}

}  // namespace

extern "C" VPT_API int vpt_svgf_create(vpt_ctx* ctx, int width, int height, const vpt_svgf_params* prm, vpt_svgf** out)
{
    if (!ctx || !out) return set_error(VPT_EINVAL, "vpt_svgf_create: null argument");
    if (width <= 1 || height <= 1 || width > 16384 || height > 16384 ||
        (uint64_t)width * (uint64_t)height * sizeof(vpt_svgf_texel) > 0xFFFFFFFFull)
        return set_error(VPT_EINVAL, "vpt_svgf_create: bad size %dx%d (at most 16384 per side, a texel plane "
                                     "below 4 GiB)", width, height);
    DeviceGuard g(ctx->device);
    vpt_svgf* s = new vpt_svgf();
    s->device = ctx->device;
    s->W = width;
    s->H = height;
    if (prm) s->prm = *prm;
    else vpt_svgf_params_default(&s->prm);
    size_t n = (size_t)width * height;
    hipError_t e = hipSuccess;
    void** planes[] = {(void**)&s->d_pre, (void**)&s->d_temporal[0], (void**)&s->d_temporal[1], (void**)&s->d_var,
                       (void**)&s->d_atrous[0], (void**)&s->d_atrous[1]};
    for (void** p : planes)
        if (e == hipSuccess) e = hipMalloc(p, n * sizeof(vpt_svgf_texel));
    for (int k = 0; k < 2; k++)
        if (e == hipSuccess) e = hipMalloc(&s->d_util[k], n * sizeof(uint2));
    if (e != hipSuccess) {
        vpt_svgf_destroy(s);
        return set_error(VPT_ENOMEM, "vpt_svgf_create: hipMalloc failed (%s)", hipGetErrorString(e));
    }
    if (int rc = vpt_svgf_reset(s, nullptr)) {
        vpt_svgf_destroy(s);
        return rc;
    }
    HIP_TRY(hipDeviceSynchronize());
    *out = s;
    return VPT_OK;
}

extern "C" VPT_API int vpt_svgf_destroy(vpt_svgf* s)
{
    if (!s) return VPT_OK;
    DeviceGuard g(s->device);
    (void)hipFree(s->d_pre);
    (void)hipFree(s->d_var);
    for (int k = 0; k < 2; k++) {
        (void)hipFree(s->d_temporal[k]);
        (void)hipFree(s->d_util[k]);
        (void)hipFree(s->d_atrous[k]);
    }
    delete s;
    return VPT_OK;
}

extern "C" VPT_API int vpt_svgf_reset(vpt_svgf* s, void* stream)
{
    if (!s) return set_error(VPT_EINVAL, "vpt_svgf_reset: null object");
    DeviceGuard g(s->device);
    hipStream_t st = (hipStream_t)stream;
    size_t n = (size_t)s->W * s->H;
    for (int k = 0; k < 2; k++) {
        if (int rc = fill_zero(s->d_temporal[k], n * sizeof(vpt_svgf_texel), st)) return rc;
        if (int rc = fill_zero(s->d_util[k], n * sizeof(uint2), st)) return rc;
    }
    s->cur = 0;
    s->have_history = false;
    return VPT_OK;
}

extern "C" VPT_API int vpt_svgf_frame(vpt_svgf* s, const vpt_camera* cam, const float prev_viewproj[16],
                                      const vpt_gbuffer* cur, const vpt_gbuffer* prev, const vpt_svgf_texel* d_in,
                                      float time, float* d_rgb_out, vpt_svgf_texel* d_texel_out, void* stream)
{
    if (!s || !cam || !prev_viewproj || !cur || !prev || !d_in || !d_rgb_out)
        return set_error(VPT_EINVAL, "vpt_svgf_frame: null argument");
    const int W = s->W, H = s->H;
    if (cam->width != W || cam->height != H || cur->width != W || cur->height != H || prev->width != W ||
        prev->height != H)
        return set_error(VPT_EINVAL, "vpt_svgf_frame: size mismatch (denoiser %dx%d)", W, H);
    if (!cur->d_t || !cur->d_face || !cur->d_block || !prev->d_t || !prev->d_face || !prev->d_block)
        return set_error(VPT_EINVAL, "vpt_svgf_frame: G-buffer planes missing");
    DeviceGuard g(s->device);
    hipStream_t st = (hipStream_t)stream;
    CamView cv = make_cam(cam);
    Mat16 vp;
    for (int k = 0; k < 16; k++) vp.m[k] = prev_viewproj[k];
    dim3 grid = pass_grid(W, H), block(256);
    const vpt_svgf_texel* temporal_in = d_in;
    if (s->prm.pre_spatial) {
        hipLaunchKernelGGL(prespatial_kernel, grid, block, 0, st, cv, W, H, cur->d_t, cur->d_face, d_in, s->d_pre);
        HIP_TRY(hipGetLastError());
        temporal_in = s->d_pre;
    }
    const int hist = s->cur, next = s->cur ^ 1;
    hipLaunchKernelGGL(temporal_kernel, grid, block, 0, st, cv, vp, W, H, cur->d_t, cur->d_face, cur->d_block,
                       prev->d_t, prev->d_face, prev->d_block, temporal_in, s->d_temporal[hist], s->d_util[hist],
                       s->d_temporal[next], s->d_util[next]);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(variance_kernel, grid, block, 0, st, W, H, cur->d_t, cur->d_face, s->d_temporal[next],
                       s->d_util[next], s->d_var);
    HIP_TRY(hipGetLastError());
    const float scale = 1.0f + (2.4f - 1.0f) * s->prm.resolution_scale;   // mix(1, 2.4, u_ResolutionScale)
    const vpt_svgf_texel* src = s->d_var;
    int dst_idx = 0;
    for (int p = 0; p < 5; p++) {
        dst_idx = p % 2;   // DiffuseDenoiseFBO / DiffuseDenoisedFBO2 ping-pong (Pipeline.cpp:2501-2503)
        // FilterAO = step <= 4, FilterSky = step <= 6 || FilterAO (SpatialFilter.glsl:183-185)
        const int step = s->prm.steps[p];
        auto kern = step <= 4 ? atrous_kernel<0> : (step <= 6 ? atrous_kernel<1> : atrous_kernel<2>);
        hipLaunchKernelGGL(kern, grid, block, 0, st, W, H, cur->d_t, cur->d_face, src, s->d_util[next],
                           step, scale, s->prm.color_phi_bias, time, s->d_atrous[dst_idx]);
        HIP_TRY(hipGetLastError());
        src = s->d_atrous[dst_idx];
    }
    hipLaunchKernelGGL(decode_kernel, grid, block, 0, st, W, H, cur->d_face, src, d_rgb_out);
    HIP_TRY(hipGetLastError());
    if (d_texel_out)
        HIP_TRY(hipMemcpyAsync(d_texel_out, src, (size_t)W * H * sizeof(vpt_svgf_texel), hipMemcpyDeviceToDevice, st));
    s->cur = next;
    s->last_final = dst_idx;
    s->have_history = true;
    return VPT_OK;
}

extern "C" VPT_API int vpt_svgf_read_utility(vpt_svgf* s, uint16_t* d_dst, void* stream)
{
    if (!s || !d_dst) return set_error(VPT_EINVAL, "vpt_svgf_read_utility: null argument");
    DeviceGuard g(s->device);
    HIP_TRY(hipMemcpyAsync(d_dst, s->d_util[s->cur], (size_t)s->W * s->H * sizeof(uint2), hipMemcpyDeviceToDevice,
                           (hipStream_t)stream));
    return VPT_OK;
}

extern "C" VPT_API int vpt_svgf_final(const vpt_svgf* s, const vpt_svgf_texel** out)
{
    if (!s || !out) return set_error(VPT_EINVAL, "vpt_svgf_final: null argument");
    *out = s->d_atrous[s->last_final];
    return VPT_OK;
}

extern "C" VPT_API int vpt_svgf_read_plane(vpt_svgf* s, int which, vpt_svgf_texel* d_dst, void* stream)
{
    if (!s || !d_dst) return set_error(VPT_EINVAL, "vpt_svgf_read_plane: null argument");
    const vpt_svgf_texel* src = nullptr;
    switch (which) {
    case 0: src = s->d_pre; break;
    case 1: src = s->d_temporal[s->cur]; break;
    case 2: src = s->d_var; break;
    case 3: src = s->d_atrous[s->last_final]; break;
    default: return set_error(VPT_EINVAL, "vpt_svgf_read_plane: plane %d unknown", which);
    }
    DeviceGuard g(s->device);
    HIP_TRY(hipMemcpyAsync(d_dst, src, (size_t)s->W * s->H * sizeof(vpt_svgf_texel), hipMemcpyDeviceToDevice,
                           (hipStream_t)stream));
    return VPT_OK;
}

// ---- sun-shadow denoise object ----------------------------------------------

extern "C" VPT_API int vpt_shadow_denoise_create(vpt_ctx* ctx, int width, int height, vpt_shadow_denoise** out)
{
    if (!ctx || !out) return set_error(VPT_EINVAL, "vpt_shadow_denoise_create: null argument");
    if (width <= 1 || height <= 1 || width > 16384 || height > 16384)
        return set_error(VPT_EINVAL, "vpt_shadow_denoise_create: bad size %dx%d", width, height);
    DeviceGuard g(ctx->device);
    vpt_shadow_denoise* d = new vpt_shadow_denoise();
    d->device = ctx->device;
    d->W = width;
    d->H = height;
    const size_t n = (size_t)width * height;
    hipError_t e = hipSuccess;
    for (int k = 0; k < 2; k++) {
        if (e == hipSuccess) e = hipMalloc(&d->d_color[k], n);
        if (e == hipSuccess) e = hipMalloc(&d->d_frames[k], n * sizeof(uint16_t));
    }
    if (e != hipSuccess) {
        vpt_shadow_denoise_destroy(d);
        return set_error(VPT_ENOMEM, "vpt_shadow_denoise_create: hipMalloc failed (%s)", hipGetErrorString(e));
    }
    if (int rc = vpt_shadow_denoise_reset(d, nullptr)) {
        vpt_shadow_denoise_destroy(d);
        return rc;
    }
    HIP_TRY(hipDeviceSynchronize());
    *out = d;
    return VPT_OK;
}

extern "C" VPT_API int vpt_shadow_denoise_destroy(vpt_shadow_denoise* d)
{
    if (!d) return VPT_OK;
    DeviceGuard g(d->device);
    for (int k = 0; k < 2; k++) {
        (void)hipFree(d->d_color[k]);
        (void)hipFree(d->d_frames[k]);
    }
    delete d;
    return VPT_OK;
}

extern "C" VPT_API int vpt_shadow_denoise_reset(vpt_shadow_denoise* d, void* stream)
{
    if (!d) return set_error(VPT_EINVAL, "vpt_shadow_denoise_reset: null object");
    DeviceGuard g(d->device);
    const size_t n = (size_t)d->W * d->H;
    for (int k = 0; k < 2; k++) {
        if (int rc = fill_zero(d->d_color[k], n, (hipStream_t)stream)) return rc;
        if (int rc = fill_zero(d->d_frames[k], n * sizeof(uint16_t), (hipStream_t)stream)) return rc;
    }
    d->cur = 0;
    return VPT_OK;
}

extern "C" VPT_API int vpt_shadow_denoise_frame(vpt_shadow_denoise* d, const vpt_camera* cam,
                                                const float prev_viewproj[16], const vpt_gbuffer* cur,
                                                const vpt_gbuffer* prev, const uint8_t* d_mask,
                                                const uint16_t* d_transversal, float filter_scale, uint8_t* d_out,
                                                void* stream)
{
    if (!d || !cam || !prev_viewproj || !cur || !prev || !d_mask || !d_transversal || !d_out)
        return set_error(VPT_EINVAL, "vpt_shadow_denoise_frame: null argument");
    const int W = d->W, H = d->H;
    if (cam->width != W || cam->height != H || cur->width != W || cur->height != H || prev->width != W ||
        prev->height != H)
        return set_error(VPT_EINVAL, "vpt_shadow_denoise_frame: size mismatch (denoiser %dx%d)", W, H);
    if (!cur->d_t || !cur->d_face || !prev->d_t)
        return set_error(VPT_EINVAL, "vpt_shadow_denoise_frame: G-buffer planes missing");
    if (!(filter_scale > 0.0f && filter_scale <= 16.0f))
        return set_error(VPT_EINVAL, "vpt_shadow_denoise_frame: filter scale %g outside (0, 16]", filter_scale);
    DeviceGuard g(d->device);
    hipStream_t st = (hipStream_t)stream;
    const CamView cv = make_cam(cam);
    Mat16 vp;
    for (int k = 0; k < 16; k++) vp.m[k] = prev_viewproj[k];
    const dim3 grid = pass_grid(W, H), block(256);
    const int hist = d->cur, next = d->cur ^ 1;
    hipLaunchKernelGGL(shadow_temporal_kernel, grid, block, 0, st, cv, vp, W, H, cur->d_t, cur->d_face, prev->d_t,
                       d_mask, d_transversal, d->d_color[hist], d->d_frames[hist], d->d_color[next],
                       d->d_frames[next]);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(shadow_filter_kernel, grid, block, 0, st, W, H, cur->d_t, cur->d_face, d->d_color[next],
                       d_transversal, d->d_frames[next], filter_scale, d_out);
    HIP_TRY(hipGetLastError());
    d->cur = next;
    return VPT_OK;
}

extern "C" VPT_API int vpt_shadow_denoise_read(vpt_shadow_denoise* d, int which, void* d_dst, void* stream)
{
    if (!d || !d_dst) return set_error(VPT_EINVAL, "vpt_shadow_denoise_read: null argument");
    const size_t n = (size_t)d->W * d->H;
    const void* src = nullptr;
    size_t bytes = 0;
    switch (which) {
    case 0: src = d->d_color[d->cur]; bytes = n; break;
    case 1: src = d->d_frames[d->cur]; bytes = n * sizeof(uint16_t); break;
    default: return set_error(VPT_EINVAL, "vpt_shadow_denoise_read: plane %d unknown", which);
    }
    DeviceGuard g(d->device);
    HIP_TRY(hipMemcpyAsync(d_dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return VPT_OK;
}

extern "C" VPT_API int vpt_check_log_division(unsigned long long* mismatches, void* stream)
{
    if (!mismatches) return set_error(VPT_EINVAL, "vpt_check_log_division: null output");
    hipStream_t st = (hipStream_t)stream;
    const uint32_t u0 = __builtin_bit_cast(uint32_t, 0.70710677f), u1 = __builtin_bit_cast(uint32_t, 1.41421354f);
    const uint32_t n = u1 - u0 + 1u;
    unsigned long long* d = nullptr;
    HIP_TRY(hipMallocAsync(reinterpret_cast<void**>(&d), sizeof(unsigned long long), st));
    hipError_t e = hipMemsetAsync(d, 0, sizeof(unsigned long long), st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(log_division_check_kernel, dim3((n + 255u) / 256u), dim3(256), 0, st, u0, n, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(mismatches, d, sizeof(unsigned long long), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFreeAsync(d, st);
    HIP_TRY(e);
    return VPT_OK;
}
