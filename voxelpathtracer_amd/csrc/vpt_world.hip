// Context, world residency and the distance-field build.
//
// DF build (replaces ManhattanDistance{X,Y,Z}.comp dispatched by
// World::GenerateDistanceField, World.cpp:69-113).  Each reference pass is a
// forward relaxation d[i] = min(d[i], d[i-1] + 1) followed by a backward one
// along one axis (the X pass starts from solid ? 0 : 254); composed they give
// min(254, L1 distance to the nearest solid voxel).  The passes keep the
// reference's order and integer semantics (the unorm8 store/load round trip
// is exact for 0..254), so the result is identical byte for byte.
//
//   df_xy_slab_kernel : one workgroup per z-slab when the slab (nx * ny
//                       bytes) fits in 64 KiB of LDS: the slab is staged
//                       with 16-byte loads, one lane per row runs the X pass,
//                       one lane per column the Y pass, and the slab is
//                       written back with 16-byte stores.
//   df_x_kernel / df_y_kernel : the same passes in global memory for larger
//                       slabs (lanes along x for the Y pass, so each step is
//                       a coalesced row access).
//   df_z_kernel       : one lane per (x, y) column, lanes along x; the
//                       backward sweep writes the final byte straight into
//                       the bricked DF.
#include <hip/hip_runtime.h>
#include <string.h>

#include <stdio.h>

#include <algorithm>
#include <vector>

#include "vpt_device.h"
#include "vpt_internal.h"

using vpt::set_error;

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return set_error(VPT_EDEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));      \
    } while (0)

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

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// ---- layout conversion ----------------------------------------------------

// Linear -> bricked copy: one thread per 16-byte group along x.
__global__ void brick_kernel(const uint8_t* __restrict__ lin, uint8_t* __restrict__ brk, int nx, int ny, int nz,
                             uint32_t row_stride, uint32_t slab_stride)
{
    size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t groups = (size_t)(nx / 16) * ny * nz;
    if (gid >= groups) return;
    int gx = (int)(gid % (nx / 16));
    size_t rest = gid / (nx / 16);
    int y = (int)(rest % ny);
    int z = (int)(rest / ny);
    const uint4 v = *reinterpret_cast<const uint4*>(lin + (size_t)z * nx * ny + (size_t)y * nx + (size_t)gx * 16);
    uint32_t words[4] = {v.x, v.y, v.z, v.w};
    for (int k = 0; k < 16; k++) {
        uint8_t b = (uint8_t)(words[k >> 2] >> ((k & 3) * 8));
        brk[vpt::brick_offset(gx * 16 + k, y, z, row_stride, slab_stride)] = b;
    }
}

// Bricked -> linear copy (downloads).
__global__ void unbrick_kernel(const uint8_t* __restrict__ brk, uint8_t* __restrict__ lin, int nx, int ny, int nz,
                               uint32_t row_stride, uint32_t slab_stride)
{
    size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t groups = (size_t)(nx / 16) * ny * nz;
    if (gid >= groups) return;
    int gx = (int)(gid % (nx / 16));
    size_t rest = gid / (nx / 16);
    int y = (int)(rest % ny);
    int z = (int)(rest / ny);
    uint32_t words[4] = {0, 0, 0, 0};
    for (int k = 0; k < 16; k++) {
        uint32_t b = brk[vpt::brick_offset(gx * 16 + k, y, z, row_stride, slab_stride)];
        words[k >> 2] |= b << ((k & 3) * 8);
    }
    *reinterpret_cast<uint4*>(lin + (size_t)z * nx * ny + (size_t)y * nx + (size_t)gx * 16) =
        make_uint4(words[0], words[1], words[2], words[3]);
}

__global__ void set_blocks_kernel(uint8_t* __restrict__ lin, uint8_t* __restrict__ brk, const uint32_t* __restrict__ edits,
                                  const uint8_t* __restrict__ ids, int n, int nx, int ny, int nz, uint32_t row_stride,
                                  uint32_t slab_stride)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t e = edits[i];
    int x = (int)(e & 1023u), y = (int)((e >> 10) & 1023u), z = (int)((e >> 20) & 1023u);
    if (x >= nx || y >= ny || z >= nz) return;
    lin[(size_t)z * nx * ny + (size_t)y * nx + x] = ids[i];
    brk[vpt::brick_offset(x, y, z, row_stride, slab_stride)] = ids[i];
}

// ---- DF passes ---------------------------------------------------------------

constexpr int kDfCap = 254;

// X then Y pass over one z-slab held in LDS (slabs of at most kSlabMax
// bytes; larger slabs use df_x_kernel + df_y_kernel).
//
// X pass: the reference's forward/backward sweep along a row equals
// min(cap, distance to the nearest solid cell of the row), which one
// wavefront computes per row with two scans: each lane owns 4 consecutive
// bytes, a wave prefix-max gives the last solid position at or before every
// byte and a wave suffix-min the next solid position at or after it.
// Y pass: one lane per column, forward and backward relaxation with the
// running value in a register (LDS reads independent of it).
constexpr int kSlabMax = 128 * 1024;
constexpr int kNone = 1 << 20;

__device__ inline int wave_incl_max(int v)
{
    const int lane = threadIdx.x & 63;
    for (int d = 1; d < 64; d <<= 1) {
        int t = __shfl_up(v, d);
        if (lane >= d) v = max(v, t);
    }
    return v;
}

__device__ inline int wave_incl_min_rev(int v)
{
    const int lane = threadIdx.x & 63;
    for (int d = 1; d < 64; d <<= 1) {
        int t = __shfl_down(v, d);
        if (lane + d < 64) v = min(v, t);
    }
    return v;
}

__global__ void __launch_bounds__(512) df_xy_slab_kernel(const uint8_t* __restrict__ grid, uint8_t* __restrict__ out,
                                                         int nx, int ny, int nz)
{
    extern __shared__ uint8_t slab[];
    const int z = blockIdx.x;
    const size_t base = (size_t)z * nx * ny;
    const int bytes = nx * ny;
    for (int i = threadIdx.x * 16; i < bytes; i += blockDim.x * 16)
        *reinterpret_cast<uint4*>(slab + i) = *reinterpret_cast<const uint4*>(grid + base + i);
    __syncthreads();
    const int cap = min(kDfCap, nx + ny + nz);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
    const int chunks = (nx + 255) / 256;   // 256 bytes (64 lanes x 4) per chunk, nx <= 1024
    // ---- X pass, one wavefront per row ----------------------------------
    for (int y = wave; y < ny; y += waves) {
        uint8_t* row = slab + y * nx;
        int fwd[4][4];
        int carry = -kNone;   // last solid position left of the chunk
#pragma unroll
        for (int c = 0; c < 4; c++) {
            if (c >= chunks) break;
            const int x0 = c * 256 + lane * 4;
            uint32_t v = x0 < nx ? *reinterpret_cast<const uint32_t*>(row + x0) : 0u;
            int last_local = -kNone;
            int lastpos[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if ((v >> (8 * k)) & 0xffu) last_local = x0 + k;
                lastpos[k] = last_local;
            }
            int incl = wave_incl_max(last_local);
            int excl = __shfl_up(incl, 1);
            if (lane == 0) excl = -kNone;
            excl = max(excl, carry);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                int l = max(lastpos[k], excl);
                fwd[c][k] = l < 0 ? kNone : x0 + k - l;
            }
            carry = max(carry, __shfl(incl, 63));
        }
        int carry_r = kNone;  // next solid position right of the chunk
#pragma unroll
        for (int c = 3; c >= 0; c--) {
            if (c >= chunks) continue;
            const int x0 = c * 256 + lane * 4;
            uint32_t v = x0 < nx ? *reinterpret_cast<const uint32_t*>(row + x0) : 0u;
            int next_local = kNone;
            int nextpos[4];
#pragma unroll
            for (int k = 3; k >= 0; k--) {
                if ((v >> (8 * k)) & 0xffu) next_local = x0 + k;
                nextpos[k] = next_local;
            }
            int incl = wave_incl_min_rev(next_local);
            int excl = __shfl_down(incl, 1);
            if (lane == 63) excl = kNone;
            excl = min(excl, carry_r);
            uint32_t outw = 0;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                int nx_pos = min(nextpos[k], excl);
                int bwd = nx_pos >= kNone ? kNone : nx_pos - (x0 + k);
                int d = min(cap, min(fwd[c][k], bwd));
                outw |= (uint32_t)d << (8 * k);
            }
            if (x0 < nx) *reinterpret_cast<uint32_t*>(row + x0) = outw;
            carry_r = min(carry_r, __shfl(incl, 0));
        }
    }
    __syncthreads();
    // ---- Y pass (ManhattanDistanceY.comp:103-118), one lane per column ---
    for (int x = threadIdx.x; x < nx; x += blockDim.x) {
        int prev = slab[x];
#pragma unroll 8
        for (int y = 1; y < ny; y++) {
            int v = min((int)slab[y * nx + x], prev + 1);
            slab[y * nx + x] = (uint8_t)v;
            prev = v;
        }
        int next = prev;
#pragma unroll 8
        for (int y = ny - 2; y >= 0; y--) {
            int v = min((int)slab[y * nx + x], next + 1);
            slab[y * nx + x] = (uint8_t)v;
            next = v;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x * 16; i < bytes; i += blockDim.x * 16)
        *reinterpret_cast<uint4*>(out + base + i) = *reinterpret_cast<const uint4*>(slab + i);
}

// Z pass over a 32 x 8 x nz column tile held in LDS (tiles of at most
// kZTileMax bytes; larger ones use df_z_kernel): one lane per (x, y)
// column, forward and backward relaxation along z (ManhattanDistanceZ.comp:
// 151-165) with the running value in a register, then the tile -- four
// whole 8x8 bricks wide and one brick high -- is written to the bricked DF
// with 16-byte stores.
constexpr int kZTileMax = 128 * 1024;

__global__ void __launch_bounds__(256) df_z_tile_kernel(const uint8_t* __restrict__ in, uint8_t* __restrict__ df_brick,
                                                        int nx, int ny, int nz, uint32_t row_stride,
                                                        uint32_t slab_stride)
{
    extern __shared__ uint8_t tile[];   // [z][y 0..7][x 0..31]
    const int tiles_x = nx / 32;
    const int x0 = (blockIdx.x % tiles_x) * 32, y0 = (blockIdx.x / tiles_x) * 8;
    const size_t plane = (size_t)nx * ny;
    // stage: 8 rows of 32 bytes per z, two 16-byte loads per row
    for (int i = threadIdx.x; i < nz * 16; i += blockDim.x) {
        const int z = i >> 4, r = (i >> 1) & 7, h = i & 1;
        *reinterpret_cast<uint4*>(tile + z * 256 + r * 32 + h * 16) =
            *reinterpret_cast<const uint4*>(in + (size_t)z * plane + (size_t)(y0 + r) * nx + x0 + h * 16);
    }
    __syncthreads();
    {
        const int c = threadIdx.x;   // column: y = c / 32, x = c % 32
        int prev = tile[c];
#pragma unroll 8
        for (int z = 1; z < nz; z++) {
            int v = min((int)tile[z * 256 + c], prev + 1);
            tile[z * 256 + c] = (uint8_t)v;
            prev = v;
        }
        int next = prev;
#pragma unroll 8
        for (int z = nz - 2; z >= 0; z--) {
            int v = min((int)tile[z * 256 + c], next + 1);
            tile[z * 256 + c] = (uint8_t)v;
            next = v;
        }
    }
    __syncthreads();
    // write whole bricks: brick (bx, bz) of this tile, 512 bytes each, one
    // 16-byte group per lane: bytes m..m+15 of the Morton order
    const int nbz = nz / 8;
    for (int i = threadIdx.x; i < 4 * nbz * 32; i += blockDim.x) {
        const int g = i & 31, brick = i >> 5;       // 32 groups of 16 bytes per brick
        const int bxl = brick & 3, bz = brick >> 2;
        uint32_t words[4] = {0u, 0u, 0u, 0u};
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const uint32_t m = (uint32_t)(g * 16 + k);   // Morton index in the brick
            const uint32_t lx = (m & 1u) | ((m >> 2) & 2u) | ((m >> 4) & 4u);
            const uint32_t ly = ((m >> 1) & 1u) | ((m >> 3) & 2u) | ((m >> 5) & 4u);
            const uint32_t lz = ((m >> 2) & 1u) | ((m >> 4) & 2u) | ((m >> 6) & 4u);
            const uint32_t b = tile[(bz * 8 + lz) * 256 + ly * 32 + bxl * 8 + lx];
            words[k >> 2] |= b << (8 * (k & 3));
        }
        const size_t dst = (size_t)((x0 >> 3) + bxl) * vpt::kBrickBytes + (size_t)(y0 >> 3) * row_stride +
                           (size_t)bz * slab_stride + (size_t)g * 16;
        *reinterpret_cast<uint4*>(df_brick + dst) = make_uint4(words[0], words[1], words[2], words[3]);
    }
}

// X pass in global memory (slabs too large for LDS): one lane per row.
__global__ void df_x_kernel(const uint8_t* __restrict__ grid, uint8_t* __restrict__ out, int nx, int ny, int nz)
{
    size_t row_id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row_id >= (size_t)ny * nz) return;
    const int cap = min(kDfCap, nx + ny + nz);
    const uint8_t* g = grid + row_id * nx;
    uint8_t* o = out + row_id * nx;
    int prev = g[0] > 0 ? 0 : cap;
    o[0] = (uint8_t)prev;
    for (int x = 1; x < nx; x++) {
        int v = g[x] > 0 ? 0 : min(cap, 1 + prev);
        o[x] = (uint8_t)v;
        prev = v;
    }
    int next = prev;
    for (int x = nx - 2; x >= 0; x--) {
        int v = o[x];
        if (next < v) v = 1 + next;
        o[x] = (uint8_t)v;
        next = v;
    }
}

// Y pass in global memory: one lane per (x, z) column, lanes along x.
__global__ void df_y_kernel(uint8_t* __restrict__ buf, int nx, int ny, int nz)
{
    size_t col = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= (size_t)nx * nz) return;
    int x = (int)(col % nx), z = (int)(col / nx);
    uint8_t* p = buf + (size_t)z * nx * ny + x;
    int prev = p[0];
    for (int y = 1; y < ny; y++) {
        int v = p[(size_t)y * nx];
        if (prev < v) v = 1 + prev;
        p[(size_t)y * nx] = (uint8_t)v;
        prev = v;
    }
    int next = prev;
    for (int y = ny - 2; y >= 0; y--) {
        int v = p[(size_t)y * nx];
        if (next < v) v = 1 + next;
        p[(size_t)y * nx] = (uint8_t)v;
        next = v;
    }
}

// Z pass (ManhattanDistanceZ.comp:151-165): one lane per (x, y) column.
// The final value of every cell is known during the backward sweep and is
// written straight into the bricked DF.
__global__ void df_z_kernel(uint8_t* __restrict__ buf, uint8_t* __restrict__ df_brick, int nx, int ny, int nz,
                            uint32_t row_stride, uint32_t slab_stride)
{
    size_t col = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= (size_t)nx * ny) return;
    int x = (int)(col % nx), y = (int)(col / nx);
    const size_t plane = (size_t)nx * ny;
    uint8_t* p = buf + (size_t)y * nx + x;
    int prev = p[0];
    for (int z = 1; z < nz; z++) {
        int v = p[(size_t)z * plane];
        if (prev < v) v = 1 + prev;
        p[(size_t)z * plane] = (uint8_t)v;
        prev = v;
    }
    int next = prev;
    df_brick[vpt::brick_offset(x, y, nz - 1, row_stride, slab_stride)] = (uint8_t)next;
    for (int z = nz - 2; z >= 0; z--) {
        int v = p[(size_t)z * plane];
        if (next < v) v = 1 + next;
        df_brick[vpt::brick_offset(x, y, z, row_stride, slab_stride)] = (uint8_t)v;
        next = v;
    }
}

int check_world(const vpt_world* w, const char* fn)
{
    if (!w) return set_error(VPT_EINVAL, "%s: null world", fn);
    return VPT_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" VPT_API int vpt_ctx_create(int device, vpt_ctx** out)
{
    if (!out) return set_error(VPT_EINVAL, "vpt_ctx_create: null output");
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count));
    if (device < 0 || device >= count)
        return set_error(VPT_EINVAL, "vpt_ctx_create: device %d out of range (%d devices)", device, count);
    DeviceGuard g(device);
    HIP_TRY(hipFree(nullptr));  // initialise the device context
    vpt_ctx* c = new vpt_ctx();
    c->device = device;
    *out = c;
    return VPT_OK;
}

extern "C" VPT_API int vpt_ctx_destroy(vpt_ctx* ctx)
{
    if (!ctx) return VPT_OK;
    DeviceGuard g(ctx->device);
    if (ctx->d_albedo) (void)hipFree(ctx->d_albedo);
    delete ctx;
    return VPT_OK;
}

extern "C" VPT_API int vpt_ctx_device(const vpt_ctx* ctx, int* device)
{
    if (!ctx || !device) return set_error(VPT_EINVAL, "vpt_ctx_device: null argument");
    *device = ctx->device;
    return VPT_OK;
}

extern "C" VPT_API int vpt_world_create(vpt_ctx* ctx, int nx, int ny, int nz, vpt_world** out)
{
    if (!ctx || !out) return set_error(VPT_EINVAL, "vpt_world_create: null argument");
    if (nx < 16 || ny < 8 || nz < 8 || nx > 1024 || ny > 1024 || nz > 1024 || nx % 16 || ny % 8 || nz % 8)
        return set_error(VPT_EINVAL,
                         "vpt_world_create: dims %dx%dx%d invalid (nx multiple of 16, ny/nz multiples of 8, <= 1024)",
                         nx, ny, nz);
    DeviceGuard g(ctx->device);
    vpt_world* w = new vpt_world();
    w->ctx = ctx;
    w->device = ctx->device;
    w->nx = nx;
    w->ny = ny;
    w->nz = nz;
    w->nbx = nx / 8;
    w->nby = ny / 8;
    w->nbz = nz / 8;
    size_t v = w->voxels();
    hipError_t e = hipMalloc(&w->d_grid_lin, v);
    if (e == hipSuccess) e = hipMalloc(&w->d_grid, v);
    if (e == hipSuccess) e = hipMalloc(&w->d_df, v);
    if (e == hipSuccess) e = hipMalloc(&w->d_tmp, v);
    if (e != hipSuccess) {
        (void)hipFree(w->d_grid_lin);
        (void)hipFree(w->d_grid);
        (void)hipFree(w->d_df);
        (void)hipFree(w->d_tmp);
        delete w;
        return set_error(VPT_ENOMEM, "vpt_world_create: hipMalloc of %zu bytes failed", 4 * v);
    }
    w->bytes = 4 * v;
    HIP_TRY(hipMemset(w->d_grid_lin, 0, v));
    HIP_TRY(hipMemset(w->d_grid, 0, v));
    HIP_TRY(hipMemset(w->d_df, 0, v));
    *out = w;
    return VPT_OK;
}

extern "C" VPT_API int vpt_world_destroy(vpt_world* w)
{
    if (!w) return VPT_OK;
    DeviceGuard g(w->device);
    (void)hipFree(w->d_grid_lin);
    (void)hipFree(w->d_grid);
    (void)hipFree(w->d_df);
    (void)hipFree(w->d_tmp);
    delete w;
    return VPT_OK;
}

extern "C" VPT_API int vpt_world_dims(const vpt_world* w, int* nx, int* ny, int* nz)
{
    if (int rc = check_world(w, "vpt_world_dims")) return rc;
    if (nx) *nx = w->nx;
    if (ny) *ny = w->ny;
    if (nz) *nz = w->nz;
    return VPT_OK;
}

extern "C" VPT_API int vpt_world_bytes(const vpt_world* w, size_t* bytes)
{
    if (int rc = check_world(w, "vpt_world_bytes")) return rc;
    if (!bytes) return set_error(VPT_EINVAL, "vpt_world_bytes: null output");
    *bytes = w->bytes;
    return VPT_OK;
}

static int rebrick_grid(vpt_world* w, hipStream_t s)
{
    size_t groups = (size_t)(w->nx / 16) * w->ny * w->nz;
    int threads = 256;
    int blocks = (int)((groups + threads - 1) / threads);
    uint32_t rs = (uint32_t)w->nbx * vpt::kBrickBytes, ss = rs * (uint32_t)w->nby;
    hipLaunchKernelGGL(brick_kernel, dim3(blocks), dim3(threads), 0, s, w->d_grid_lin, w->d_grid, w->nx, w->ny,
                       w->nz, rs, ss);
    HIP_TRY(hipGetLastError());
    return VPT_OK;
}

extern "C" VPT_API int vpt_world_upload(vpt_world* w, const uint8_t* host_grid, void* stream)
{
    if (int rc = check_world(w, "vpt_world_upload")) return rc;
    if (!host_grid) return set_error(VPT_EINVAL, "vpt_world_upload: null grid");
    DeviceGuard g(w->device);
    hipStream_t s = as_stream(stream);
    HIP_TRY(hipMemcpyAsync(w->d_grid_lin, host_grid, w->voxels(), hipMemcpyHostToDevice, s));
    return rebrick_grid(w, s);
}

extern "C" VPT_API int vpt_world_upload_device(vpt_world* w, const uint8_t* d_grid, void* stream)
{
    if (int rc = check_world(w, "vpt_world_upload_device")) return rc;
    if (!d_grid) return set_error(VPT_EINVAL, "vpt_world_upload_device: null grid");
    DeviceGuard g(w->device);
    hipStream_t s = as_stream(stream);
    HIP_TRY(hipMemcpyAsync(w->d_grid_lin, d_grid, w->voxels(), hipMemcpyDeviceToDevice, s));
    return rebrick_grid(w, s);
}

extern "C" VPT_API int vpt_world_set_blocks(vpt_world* w, const uint32_t* edits, const uint8_t* ids, int n,
                                            void* stream)
{
    if (int rc = check_world(w, "vpt_world_set_blocks")) return rc;
    if (n < 0 || (n > 0 && (!edits || !ids))) return set_error(VPT_EINVAL, "vpt_world_set_blocks: bad edit list");
    if (n == 0) return VPT_OK;
    for (int i = 0; i < n; i++) {
        uint32_t e = edits[i];
        int x = (int)(e & 1023u), y = (int)((e >> 10) & 1023u), z = (int)((e >> 20) & 1023u);
        if (x >= w->nx || y >= w->ny || z >= w->nz || (e >> 30))
            return set_error(VPT_EINVAL, "vpt_world_set_blocks: edit %d at (%d,%d,%d) outside the world", i, x, y, z);
    }
    DeviceGuard g(w->device);
    hipStream_t s = as_stream(stream);
    uint32_t* d_e = nullptr;
    uint8_t* d_i = nullptr;
    HIP_TRY(hipMallocAsync(reinterpret_cast<void**>(&d_e), sizeof(uint32_t) * n, s));
    HIP_TRY(hipMallocAsync(reinterpret_cast<void**>(&d_i), n, s));
    HIP_TRY(hipMemcpyAsync(d_e, edits, sizeof(uint32_t) * n, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_i, ids, n, hipMemcpyHostToDevice, s));
    uint32_t rs = (uint32_t)w->nbx * vpt::kBrickBytes, ss = rs * (uint32_t)w->nby;
    hipLaunchKernelGGL(set_blocks_kernel, dim3((n + 255) / 256), dim3(256), 0, s, w->d_grid_lin, w->d_grid, d_e, d_i,
                       n, w->nx, w->ny, w->nz, rs, ss);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipFreeAsync(d_e, s));
    HIP_TRY(hipFreeAsync(d_i, s));
    return VPT_OK;
}

extern "C" VPT_API int vpt_df_build(vpt_world* w, void* stream)
{
    if (int rc = check_world(w, "vpt_df_build")) return rc;
    DeviceGuard g(w->device);
    hipStream_t s = as_stream(stream);
    const int nx = w->nx, ny = w->ny, nz = w->nz;
    const size_t slab = (size_t)nx * ny;
    if (slab <= (size_t)kSlabMax) {
        HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(&df_xy_slab_kernel),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, kSlabMax));
        hipLaunchKernelGGL(df_xy_slab_kernel, dim3(nz), dim3(512), slab, s, w->d_grid_lin, w->d_tmp, nx, ny, nz);
        HIP_TRY(hipGetLastError());
    } else {
        size_t rows = (size_t)ny * nz;
        hipLaunchKernelGGL(df_x_kernel, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, s, w->d_grid_lin,
                           w->d_tmp, nx, ny, nz);
        HIP_TRY(hipGetLastError());
        size_t cols = (size_t)nx * nz;
        hipLaunchKernelGGL(df_y_kernel, dim3((unsigned)((cols + 255) / 256)), dim3(256), 0, s, w->d_tmp, nx, ny, nz);
        HIP_TRY(hipGetLastError());
    }
    uint32_t rs = (uint32_t)w->nbx * vpt::kBrickBytes, ss = rs * (uint32_t)w->nby;
    const size_t ztile = (size_t)nz * 256;
    if (nx % 32 == 0 && ztile <= (size_t)kZTileMax) {
        HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(&df_z_tile_kernel),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, kZTileMax));
        hipLaunchKernelGGL(df_z_tile_kernel, dim3((unsigned)((nx / 32) * (ny / 8))), dim3(256), ztile, s, w->d_tmp,
                           w->d_df, nx, ny, nz, rs, ss);
    } else {
        size_t cols = (size_t)nx * ny;
        hipLaunchKernelGGL(df_z_kernel, dim3((unsigned)((cols + 255) / 256)), dim3(256), 0, s, w->d_tmp, w->d_df, nx,
                           ny, nz, rs, ss);
    }
    HIP_TRY(hipGetLastError());
    return VPT_OK;
}

static int download_bricked(vpt_world* w, const uint8_t* d_src, uint8_t* host, hipStream_t s)
{
    size_t groups = (size_t)(w->nx / 16) * w->ny * w->nz;
    uint32_t rs = (uint32_t)w->nbx * vpt::kBrickBytes, ss = rs * (uint32_t)w->nby;
    hipLaunchKernelGGL(unbrick_kernel, dim3((unsigned)((groups + 255) / 256)), dim3(256), 0, s, d_src, w->d_tmp,
                       w->nx, w->ny, w->nz, rs, ss);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(host, w->d_tmp, w->voxels(), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return VPT_OK;
}

extern "C" VPT_API int vpt_df_download(vpt_world* w, uint8_t* host_df, void* stream)
{
    if (int rc = check_world(w, "vpt_df_download")) return rc;
    if (!host_df) return set_error(VPT_EINVAL, "vpt_df_download: null output");
    DeviceGuard g(w->device);
    return download_bricked(w, w->d_df, host_df, as_stream(stream));
}

extern "C" VPT_API int vpt_grid_download(vpt_world* w, uint8_t* host_grid, void* stream)
{
    if (int rc = check_world(w, "vpt_grid_download")) return rc;
    if (!host_grid) return set_error(VPT_EINVAL, "vpt_grid_download: null output");
    DeviceGuard g(w->device);
    // Read the bricked copy back, so the test also covers the layout.
    return download_bricked(w, w->d_grid, host_grid, as_stream(stream));
}

extern "C" VPT_API int vpt_world_set_sample_counter(vpt_world* w, unsigned long long* d_counter)
{
    if (int rc = check_world(w, "vpt_world_set_sample_counter")) return rc;
    w->d_counter = d_counter;
    return VPT_OK;
}

extern "C" VPT_API int vpt_world_save(vpt_world* w, const char* path)
{
    if (int rc = check_world(w, "vpt_world_save")) return rc;
    if (!path) return set_error(VPT_EINVAL, "vpt_world_save: null path");
    std::vector<uint8_t> host(w->voxels());
    DeviceGuard g(w->device);
    HIP_TRY(hipMemcpy(host.data(), w->d_grid_lin, host.size(), hipMemcpyDeviceToHost));
    FILE* f = fopen(path, "wb");
    if (!f) return set_error(VPT_EINVAL, "vpt_world_save: cannot open %s for writing", path);
    const size_t n = fwrite(host.data(), 1, host.size(), f);
    const int closed = fclose(f);
    if (n != host.size() || closed != 0) return set_error(VPT_EINVAL, "vpt_world_save: short write to %s", path);
    return VPT_OK;
}

extern "C" VPT_API int vpt_world_load(vpt_world* w, const char* path, void* stream)
{
    if (int rc = check_world(w, "vpt_world_load")) return rc;
    if (!path) return set_error(VPT_EINVAL, "vpt_world_load: null path");
    FILE* f = fopen(path, "rb");
    if (!f) return set_error(VPT_EINVAL, "vpt_world_load: cannot open %s", path);
    std::vector<uint8_t> host(w->voxels());
    const size_t n = fread(host.data(), 1, host.size(), f);
    const bool extra = fgetc(f) != EOF;
    fclose(f);
    if (n != host.size() || extra)
        return set_error(VPT_EINVAL, "vpt_world_load: %s does not hold exactly %zu bytes (%dx%dx%d)", path,
                         host.size(), w->nx, w->ny, w->nz);
    if (int rc = vpt_world_upload(w, host.data(), stream)) return rc;
    // the host buffer is pageable and goes out of scope: finish the copy first
    DeviceGuard g(w->device);
    HIP_TRY(hipStreamSynchronize(as_stream(stream)));
    return VPT_OK;
}
