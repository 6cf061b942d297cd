// Host-side utilities of the C ABI: synthetic terrain, camera matrices,
// GI frame constants, error strings.  No device code here.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "vpt/vpt.h"
#include "vpt/vpt_numerics.h"
#include "vpt_internal.h"

namespace vpt {

static thread_local char g_last_error[512] = "";

int set_error(int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return code;
}

}  // namespace vpt

using vpt::set_error;

extern "C" VPT_API const char* vpt_last_error(void) { return vpt::g_last_error; }

extern "C" VPT_API const char* vpt_version(void) { return "vpt-mi355x 0.1 (gfx950)"; }

// ---------------------------------------------------------------------------
// Terrain.  Value noise: 4 octaves with lattice periods 64, 32, 16, 8 and
// amplitudes 1, 1/2, 1/4, 1/8; lattice values from pcg_hash; smoothstep
// interpolation.  Heights map [0, 1) onto [h_min, h_max].  Layering follows
// WorldGenerator.cpp:28-47 (top voxel grass, next 4 dirt, rest stone) with
// the block ids of blockdb.txt order (Grass = 1, Dirt = 2, Stone = 3).
// ---------------------------------------------------------------------------

static float lattice_value(uint32_t seed, int ix, int iz, int octave)
{
    uint32_t h = vpt_pcg_hash((uint32_t)ix * 73856093u ^ (uint32_t)iz * 19349663u ^ (uint32_t)octave * 83492791u);
    return vpt_u32_to_unit(vpt_pcg_hash(h ^ seed));
}

static float smooth01(float t) { return t * t * (3.0f - 2.0f * t); }

extern "C" VPT_API int vpt_generate_terrain(int nx, int ny, int nz, uint32_t seed, uint32_t seed_random, int h_min,
                                            int h_max, int random_permille, uint8_t* host_grid)
{
    if (nx <= 0 || ny <= 0 || nz <= 0 || !host_grid || h_min < 0 || h_max < h_min || random_permille < 0 ||
        random_permille > 1000)
        return set_error(VPT_EINVAL, "vpt_generate_terrain: bad arguments");
    const size_t plane = (size_t)nx * ny;
    memset(host_grid, 0, plane * nz);
    for (int z = 0; z < nz; z++) {
        for (int x = 0; x < nx; x++) {
            float n = 0.0f, amp_sum = 0.0f, amp = 1.0f;
            for (int o = 0; o < 4; o++) {
                int period = 64 >> o;
                float fx = ((float)x + 0.5f) / (float)period;
                float fz = ((float)z + 0.5f) / (float)period;
                int ix = (int)floorf(fx), iz = (int)floorf(fz);
                float tx = smooth01(fx - (float)ix), tz = smooth01(fz - (float)iz);
                float v00 = lattice_value(seed, ix, iz, o), v10 = lattice_value(seed, ix + 1, iz, o);
                float v01 = lattice_value(seed, ix, iz + 1, o), v11 = lattice_value(seed, ix + 1, iz + 1, o);
                float a = v00 + (v10 - v00) * tx;
                float b = v01 + (v11 - v01) * tx;
                n += amp * (a + (b - a) * tz);
                amp_sum += amp;
                amp *= 0.5f;
            }
            n /= amp_sum;
            int h = h_min + (int)floorf(n * (float)(h_max - h_min + 1));
            h = std::min(std::max(h, h_min), h_max);
            int top = std::min(h, ny);
            for (int y = 0; y < top; y++) {
                uint8_t id = (y >= h - 1) ? 1 : (y >= h - 5 ? 2 : 3);
                host_grid[(size_t)x + (size_t)y * nx + (size_t)z * plane] = id;
            }
        }
    }
    if (random_permille > 0) {
        const size_t total = plane * nz;
        for (size_t i = 0; i < total; i++) {
            if (host_grid[i] == 0) continue;
            uint32_t r = vpt_pcg_hash((uint32_t)i ^ vpt_pcg_hash(seed_random ^ (uint32_t)(i >> 32)));
            if (r % 1000u < (uint32_t)random_permille) host_grid[i] = (uint8_t)(1u + vpt_pcg_hash(r) % 100u);
        }
    }
    return VPT_OK;
}

// ---------------------------------------------------------------------------
// Camera.  glm 0.9.8.5 right-handed lookAt / perspective (clip z in [-1,1]),
// evaluated in double and rounded to float once.
// ---------------------------------------------------------------------------

extern "C" VPT_API int vpt_camera_make(const float pos[3], float yaw_deg, float pitch_deg, float fov_deg, int width,
                                       int height, float znear, float zfar, vpt_camera* cam, float view[16],
                                       float proj[16])
{
    if (!pos || width <= 0 || height <= 0 || !(fov_deg > 0.0f && fov_deg < 180.0f) || !(znear > 0.0f) ||
        !(zfar > znear) || !(pitch_deg >= -89.0f && pitch_deg <= 89.0f) || !isfinite(yaw_deg))
        return set_error(VPT_EINVAL, "vpt_camera_make: bad arguments");
    for (int i = 0; i < 3; i++)
        if (!isfinite(pos[i])) return set_error(VPT_EINVAL, "vpt_camera_make: non-finite position");
    const double PI = 3.14159265358979323846;
    double yaw = yaw_deg * PI / 180.0, pitch = pitch_deg * PI / 180.0;
    double f[3] = {cos(pitch) * cos(yaw), sin(pitch), cos(pitch) * sin(yaw)};
    double fl = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (double& v : f) v /= fl;
    double s[3] = {-f[2], 0.0, f[0]};  // cross(f, up) with up = +y
    double sl = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (double& v : s) v /= sl;
    double u[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2], s[0] * f[1] - s[1] * f[0]};
    double e[3] = {pos[0], pos[1], pos[2]};
    double aspect = (double)width / (double)height;
    double th = tan(fov_deg * PI / 360.0);
    double P0 = 1.0 / (aspect * th), P5 = 1.0 / th;
    double P10 = -((double)zfar + znear) / ((double)zfar - znear);
    double P14 = -(2.0 * zfar * znear) / ((double)zfar - znear);
    if (view) {
        double V[16] = {s[0], u[0], -f[0], 0, s[1], u[1], -f[1], 0, s[2], u[2], -f[2], 0,
                        -(s[0] * e[0] + s[1] * e[1] + s[2] * e[2]), -(u[0] * e[0] + u[1] * e[1] + u[2] * e[2]),
                        (f[0] * e[0] + f[1] * e[1] + f[2] * e[2]), 1};
        for (int i = 0; i < 16; i++) view[i] = (float)V[i];
    }
    if (proj) {
        double P[16] = {P0, 0, 0, 0, 0, P5, 0, 0, 0, 0, P10, -1, 0, 0, P14, 0};
        for (int i = 0; i < 16; i++) proj[i] = (float)P[i];
    }
    if (cam) {
        double IV[16] = {s[0], s[1], s[2], 0, u[0], u[1], u[2], 0, -f[0], -f[1], -f[2], 0, e[0], e[1], e[2], 1};
        double IP[16] = {1.0 / P0, 0, 0, 0, 0, 1.0 / P5, 0, 0, 0, 0, 0, 1.0 / P14, 0, 0, -1.0, P10 / P14};
        for (int i = 0; i < 16; i++) {
            cam->inv_view[i] = (float)IV[i];
            cam->inv_proj[i] = (float)IP[i];
        }
        cam->width = width;
        cam->height = height;
    }
    return VPT_OK;
}

// ---------------------------------------------------------------------------
// GI frame constants (Pipeline.cpp:1652-1671, 1828;
// DiffuseRayTraceFrag.glsl:181-186, 829-840).  GISunStrength = 1,
// GISkyStrength = 1.125 (Pipeline.cpp:280-281).  The reference samples an
// atmosphere cubemap for the sky; the workload uses the constant sky colour
// (0.35, 0.55, 0.90).
// ---------------------------------------------------------------------------

extern "C" VPT_API int vpt_reflect_params_default(const float light_dir[3], uint32_t frame, vpt_reflect_params* out)
{
    if (!light_dir || !out) return set_error(VPT_EINVAL, "vpt_reflect_params_default: null argument");
    memset(out, 0, sizeof(*out));
    for (int i = 0; i < 3; i++) out->light_dir[i] = light_dir[i];
    out->div = 4;              // ReflectionTraceResolution 0.25 (Pipeline.cpp:101)
    out->max_iter = 64;        // ReflectionTraceLength (Pipeline.cpp:100)
    out->shadow_iter = 150;    // VoxelTraversalDF's shadow cap (ReflectionTraceFrag.glsl:1098)
    out->spp = 2;              // ReflectionSPP (Pipeline.cpp:103)
    out->rough = 1;            // RoughReflections (Pipeline.cpp:104)
    out->roughness = 0.5f * 0.85f;   // a flat material's 0.5 times RoughnessBias (ReflectionTraceFrag.glsl:814-815)
    out->frame = frame;
    return VPT_OK;
}

extern "C" VPT_API int vpt_gi_params_default(float sun_tick, uint32_t frame, uint32_t albedo_seed, vpt_gi_params* prm)
{
    if (!prm) return set_error(VPT_EINVAL, "vpt_gi_params_default: null output");
    const double PI = 3.14159265358979323846;
    double ang = (double)sun_tick * 2.0 * PI / 180.0;
    double c = cos(ang), s = sin(ang);
    double sun[3] = {c - s, s + c, 1.0};
    double sl = sqrt(sun[0] * sun[0] + sun[1] * sun[1] + sun[2] * sun[2]);
    for (double& v : sun) v /= sl;
    bool sun_stronger = -sun[1] < 0.01;
    if (!sun_stronger)
        return set_error(VPT_EINVAL, "vpt_gi_params_default: sun below the horizon (moon-lit GI is not supported)");
    const double sun_color[3] = {192.0 / 255.0 * 16.0, 216.0 / 255.0 * 16.0, 255.0 / 255.0 * 16.0};
    const double dusk_color[3] = {96.0 / 255.0 * 0.9, 192.0 / 255.0 * 0.9, 255.0 / 255.0 * 0.9};
    double dusk = std::min(std::max(pow(fabs(sun[1] - 1.0), 2.9), 0.0), 1.0);
    double sun_visibility = std::min(std::max(sun[1] + 0.05, 0.0), 0.1) * 12.0;
    double sky_x = std::min(std::max((1.0 + (1.05 - 1.0) * sun_visibility) * 1.125, 0.0), 5.0);
    const double sky_color[3] = {0.35, 0.55, 0.90};
    memset(prm, 0, sizeof(*prm));
    for (int i = 0; i < 3; i++) {
        prm->sun_dir[i] = (float)sun[i];
        prm->light_color[i] = (float)((sun_color[i] + (dusk_color[i] - sun_color[i]) * dusk) * 0.4 * 1.0);
        prm->sky_radiance[i] = (float)(sky_color[i] * sky_x);
        prm->sky_pixel_radiance[i] = (float)(sky_color[i] * 2.66);
    }
    for (int id = 0; id < 128; id++) {
        for (int k = 0; k < 3; k++) {
            uint32_t h = vpt_pcg_hash(albedo_seed ^ vpt_pcg_hash((uint32_t)(id * 3 + k)));
            prm->albedo[3 * id + k] = 0.1f + 0.8f * vpt_u32_to_unit(h);
        }
    }
    prm->frame = frame;
    prm->gi_iter = 48;
    prm->shadow_iter = 128;
    (void)PI;
    return VPT_OK;
}

// ---------------------------------------------------------------------------
// Block picking (World::RaycastDetect, World.cpp:496-546).
// ---------------------------------------------------------------------------

extern "C" VPT_API int vpt_raycast_detect(const uint8_t* host_grid, int nx, int ny, int nz, const float pos[3],
                                          const float dir[3], int32_t out[4])
{
    if (!host_grid || !pos || !dir || !out || nx <= 0 || ny <= 0 || nz <= 0)
        return set_error(VPT_EINVAL, "vpt_raycast_detect: bad arguments");
    out[0] = out[1] = out[2] = out[3] = -1;
    float p[3] = {pos[0], pos[1], pos[2]};
    float sign[3];
    for (int i = 0; i < 3; i++) sign[i] = dir[i] > 0.0f ? 1.0f : 0.0f;
    const size_t plane = (size_t)nx * ny;
    for (int step = 0; step < 48; step++) {
        float tv[3];
        for (int i = 0; i < 3; i++) tv[i] = (floorf(p[i] + sign[i]) - p[i]) / dir[i];
        float t = std::min(tv[0], std::min(tv[1], tv[2]));
        for (int i = 0; i < 3; i++) p[i] = p[i] + dir[i] * (t + 0.001f);
        int fx = (int)floorf(p[0]), fy = (int)floorf(p[1]), fz = (int)floorf(p[2]);
        if (fx >= nx || fy >= ny || fz >= nz || fx <= 0 || fy <= 0 || fz <= 0) continue;
        uint8_t b = host_grid[(size_t)fx + (size_t)fy * nx + (size_t)fz * plane];
        if (b != 0) {
            out[0] = fx;
            out[1] = fy;
            out[2] = fz;
            out[3] = b;
            return VPT_OK;
        }
    }
    return VPT_OK;
}

extern "C" VPT_API int vpt_svgf_params_default(vpt_svgf_params* prm)
{
    if (!prm) return set_error(VPT_EINVAL, "vpt_svgf_params_default: null output");
    prm->color_phi_bias = 3.325f;
    prm->resolution_scale = 0.25f;
    prm->pre_spatial = 1;
    const int steps[5] = {16, 8, 4, 2, 1};
    for (int i = 0; i < 5; i++) prm->steps[i] = steps[i];
    return VPT_OK;
}
