// Learned denoiser: a 3-level U-Net over the tracer's 1-spp irradiance.
//
//   x (32 ch: log1p irradiance, albedo, normal, depth, zero padding)
//   e0a = relu(conv3x3(x))            32   full res
//   e0b = relu(conv3x3(e0a))          32          p0 = avgpool2(e0b)
//   e1a = relu(conv3x3(p0))           64   1/2
//   e1b = relu(conv3x3(e1a))          64          p1 = avgpool2(e1b)
//   b0  = relu(conv3x3(p1))          128   1/4
//   b1  = relu(conv3x3(b0))          128
//   d1a = relu(conv3x3(cat(up2(b1), e1b)))   64   1/2
//   d1b = relu(conv3x3(d1a))          64
//   d0a = relu(conv3x3(cat(up2(d1b), e0b)))  32   full
//   d0b = relu(conv3x3(d0a))          32
//   out = conv1x1(d0b) -> 3 (log1p irradiance; fp32)
//
// Activations are bf16 NHWC (channels a multiple of 32); weights are bf16
// [tap][cout][cin] copies of fp32 master weights; every product accumulates
// in fp32 on the MFMA units (v_mfma_f32_16x16x32_bf16).  A 3x3 convolution
// is an implicit GEMM: M = 16 pixels of an output row, N = 16 output
// channels, K = 32 input channels of one tap.  A 16x32 pixel tile and its
// one-pixel halo are staged in LDS per 32-channel slice together with the
// slice's weights for all nine taps; upsampling and concatenation are folded
// into that staging, bias / ReLU / 2x2 average pooling into the epilogue.
// The backward pass reuses the forward kernel with flipped, transposed
// weights (data gradient, ReLU mask in the epilogue) and an MFMA kernel for
// the weight gradient (K = pixels, LDS-transposed tiles, per-slice partials
// reduced in a fixed order, so gradients are deterministic).  Adam updates
// the fp32 master weights and refreshes the bf16 copies.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <random>
#include <vector>

#include "vpt_device.h"
#include "vpt_gi.h"
#include "vpt_internal.h"

using vpt::set_error;

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return set_error(VPT_EDEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));      \
    } while (0)

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

// bf16 <-> fp32 (round to nearest even; NaN stays NaN), identical on host
// and device.
__host__ __device__ inline uint16_t f2bf(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__host__ __device__ inline float bf2f(uint16_t h)
{
    uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// a bf16 value > 0 (the ReLU mask of a stored activation)
__device__ inline bool bf_pos(uint16_t h) { return h != 0 && !(h & 0x8000u); }

// Two bf16 values of v kept where the matching halves of m are > 0.
__device__ inline uint32_t mask_pair(uint32_t v, uint32_t m)
{
    const uint32_t lo = bf_pos((uint16_t)(m & 0xffffu)) ? (v & 0xffffu) : 0u;
    const uint32_t hi = bf_pos((uint16_t)(m >> 16)) ? (v & 0xffff0000u) : 0u;
    return lo | hi;
}

constexpr int kCi = 32;            // input channels per K slice
constexpr int kRow = 40;           // LDS row pitch in halves (32 + 8: conflict-free 16-byte reads)

struct ConvArgs {
    const uint16_t* a;   // input A, NHWC with ca channels (half resolution when a_up)
    const uint16_t* b;   // concatenated input B, NHWC with cb channels (full resolution)
    const uint16_t* w;   // [9][cout][cin] bf16
    const float* bias;   // [cout] or null
    const uint16_t* mask;// [n][h][w][cout]: y = acc where mask > 0, else 0 (no bias / ReLU)
    uint16_t* y;         // [n][h][w][cout]
    uint16_t* pool;      // [n][h/2][w/2][cout] 2x2 average of y, or null
    int ca, cb, a_up, relu;
    int n, h, w_, cout;
};

// Loads 8 channels (16 bytes) of the conv input at (img, gy, gx), channel
// c0 of the concatenated input; zero outside the image.
__device__ inline uint4 load_in8(const ConvArgs& p, int img, int gy, int gx, int c0)
{
    if (gy < 0 || gx < 0 || gy >= p.h || gx >= p.w_) return make_uint4(0u, 0u, 0u, 0u);
    if (c0 < p.ca) {
        if (p.a_up) {
            const int hh = p.h >> 1, ww = p.w_ >> 1;
            const size_t off = (((size_t)img * hh + (gy >> 1)) * ww + (gx >> 1)) * p.ca + c0;
            return *reinterpret_cast<const uint4*>(p.a + off);
        }
        const size_t off = (((size_t)img * p.h + gy) * p.w_ + gx) * p.ca + c0;
        return *reinterpret_cast<const uint4*>(p.a + off);
    }
    const size_t off = (((size_t)img * p.h + gy) * p.w_ + gx) * p.cb + (c0 - p.ca);
    return *reinterpret_cast<const uint4*>(p.b + off);
}

// 3x3 convolution, stride 1, zero padding, 512 threads on 16 x 32 pixel
// tiles (wave v owns rows 4(v%4)..+3, columns 16(v/4)..+15), persistent: a workgroup
// walks (tile, 32-channel slice) steps, tile = blockIdx.x + k * gridDim.x,
// with the next step's global loads in flight (in registers) while the
// current step is multiplied and the finished tile is written out.  The
// epilogue goes through LDS so each lane stores 16 contiguous bytes.
// Per wave and tap: NT weight fragments and 4 pixel-row fragments from LDS
// (conflict-free 16-byte reads at an 80-byte pixel pitch) feed 4 x NT MFMAs.
template <int COUT_T>
__global__ void __launch_bounds__(512) conv3x3_wide_kernel(ConvArgs p)
{
    constexpr int NT = COUT_T / 16;
    constexpr int TW = 32, TH = 16, HW = TW + 2, HH = TH + 2;
    constexpr int A_ELEMS = HH * HW * 4;          // 16-byte pieces of the halo slice
    constexpr int B_ELEMS = 9 * COUT_T * 4;       // 16-byte pieces of the weight slice
    constexpr int A_PER = (A_ELEMS + 511) / 512;
    constexpr int B_PER = (B_ELEMS + 511) / 512;
    constexpr int OROW = COUT_T + 8;              // epilogue pitch (halves) per pixel
    constexpr int A_BYTES = HH * HW * kRow * 2, B_BYTES = 9 * COUT_T * kRow * 2;
    static_assert(8 * 64 * OROW * 2 <= A_BYTES + B_BYTES, "epilogue staging does not fit");
    __shared__ __attribute__((aligned(16))) uint8_t smem[A_BYTES + B_BYTES];
    uint16_t* sA = reinterpret_cast<uint16_t*>(smem);
    uint16_t* sB = reinterpret_cast<uint16_t*>(smem + A_BYTES);

    const int tiles_x = (p.w_ + TW - 1) / TW, tiles_y = (p.h + TH - 1) / TH;
    const int per_img = tiles_x * tiles_y;
    const int ntiles = p.n * per_img;
    const int co0 = blockIdx.y * COUT_T;
    const int cin = p.ca + p.cb;
    const int nch = cin / kCi;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lm = lane & 15, lk = (lane >> 4) * 8;
    const int wr = (wave & 3) * 4, wc = (wave >> 2) * 16;
    if ((int)blockIdx.x >= ntiles) return;
    const int my_tiles = (ntiles - (int)blockIdx.x + (int)gridDim.x - 1) / (int)gridDim.x;
    const int nsteps = my_tiles * nch;

    uint4 ra[A_PER], rb[B_PER];
    // loads of step `st` (tile blockIdx.x + (st / nch) * gridDim.x, slice st % nch)
#define VPT_FETCH(ST)                                                                                               \
    do {                                                                                                            \
        const int t_ = (int)blockIdx.x + ((ST) / nch) * (int)gridDim.x;                                             \
        const int c_ = ((ST) % nch) * kCi;                                                                          \
        const int im_ = t_ / per_img;                                                                               \
        const int y_ = ((t_ % per_img) / tiles_x) * TH, x_ = ((t_ % per_img) % tiles_x) * TW;                       \
        _Pragma("unroll") for (int k = 0; k < A_PER; k++)                                                           \
        {                                                                                                           \
            const int e = tid + k * 512;                                                                            \
            const int pix = e >> 2, q = e & 3;                                                                      \
            ra[k] = e < A_ELEMS ? load_in8(p, im_, y_ + pix / HW - 1, x_ + pix % HW - 1, c_ + q * 8)                \
                                : make_uint4(0u, 0u, 0u, 0u);                                                       \
        }                                                                                                           \
        _Pragma("unroll") for (int k = 0; k < B_PER; k++)                                                           \
        {                                                                                                           \
            const int e = tid + k * 512;                                                                            \
            const int row = e >> 2, q = e & 3;                                                                      \
            const int tap = row / COUT_T, co = row % COUT_T;                                                        \
            rb[k] = e < B_ELEMS ? *reinterpret_cast<const uint4*>(p.w + ((size_t)tap * p.cout + co0 + co) * cin +   \
                                                                  c_ + q * 8)                                       \
                                : make_uint4(0u, 0u, 0u, 0u);                                                       \
        }                                                                                                           \
    } while (0)
#define VPT_STASH()                                                                                                 \
    do {                                                                                                            \
        _Pragma("unroll") for (int k = 0; k < A_PER; k++)                                                           \
        {                                                                                                           \
            const int e = tid + k * 512;                                                                            \
            if (e < A_ELEMS) *reinterpret_cast<uint4*>(&sA[(e >> 2) * kRow + (e & 3) * 8]) = ra[k];                \
        }                                                                                                           \
        _Pragma("unroll") for (int k = 0; k < B_PER; k++)                                                           \
        {                                                                                                           \
            const int e = tid + k * 512;                                                                            \
            if (e < B_ELEMS) *reinterpret_cast<uint4*>(&sB[(e >> 2) * kRow + (e & 3) * 8]) = rb[k];                \
        }                                                                                                           \
    } while (0)

    f32x4 acc[4][NT];
    VPT_FETCH(0);
    VPT_STASH();
    __syncthreads();
    for (int st = 0; st < nsteps; st++) {
        const int chunk = st % nch;
        if (chunk == 0) {
#pragma unroll
            for (int r = 0; r < 4; r++)
#pragma unroll
                for (int j = 0; j < NT; j++) acc[r][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        }
        const bool more = st + 1 < nsteps;
        if (more) VPT_FETCH(st + 1);
#pragma unroll
        for (int tap = 0; tap < 9; tap++) {
            const int dy = tap / 3, dx = tap % 3;
            bf16x8 bfr[NT];
#pragma unroll
            for (int j = 0; j < NT; j++)
                bfr[j] = *reinterpret_cast<const bf16x8*>(&sB[(tap * COUT_T + j * 16 + lm) * kRow + lk]);
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int hy = wr + r + dy, hx = wc + lm + dx;
                const bf16x8 afr = *reinterpret_cast<const bf16x8*>(&sA[(hy * HW + hx) * kRow + lk]);
#pragma unroll
                for (int j = 0; j < NT; j++)
                    acc[r][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr, bfr[j], acc[r][j], 0, 0, 0);
            }
        }
        __syncthreads();
        if (chunk == nch - 1) {
            // the tile is complete: epilogue, part 1: bias / ReLU in
            // registers, bf16 values into this wave's LDS block [64 px][COUT_T],
            // pooled values stored directly
            const int t = (int)blockIdx.x + (st / nch) * (int)gridDim.x;
            const int img = t / per_img;
            const int ty0 = ((t % per_img) / tiles_x) * TH, tx0 = ((t % per_img) % tiles_x) * TW;
            uint16_t* so = reinterpret_cast<uint16_t*>(smem) + wave * 64 * OROW;
            const int xl = 4 * (lane >> 4);
#pragma unroll
            for (int j = 0; j < NT; j++) {
                const int col = j * 16 + lm;
                const float bias = p.bias ? p.bias[co0 + col] : 0.0f;
#pragma unroll
                for (int r = 0; r < 4; r++) {
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        float v = acc[r][j][i];
                        if (!p.mask) {
                            v += bias;
                            if (p.relu) v = fmaxf(v, 0.0f);
                        }
                        const uint16_t hv = f2bf(v);
                        acc[r][j][i] = bf2f(hv);   // the stored value, for the pooled output
                        so[(r * 16 + xl + i) * OROW + col] = hv;
                    }
                }
                if (p.pool) {
                    const int ph = p.h >> 1, pw = p.w_ >> 1;
#pragma unroll
                    for (int r2 = 0; r2 < 2; r2++) {
                        const int gy = ty0 + wr + 2 * r2;
#pragma unroll
                        for (int i2 = 0; i2 < 2; i2++) {
                            const int gx = tx0 + wc + xl + 2 * i2;
                            if (gy < p.h && gx < p.w_) {
                                const f32x4 a = acc[2 * r2][j], b = acc[2 * r2 + 1][j];
                                const float sm = ((a[2 * i2] + a[2 * i2 + 1]) + (b[2 * i2] + b[2 * i2 + 1])) * 0.25f;
                                p.pool[(((size_t)img * ph + (gy >> 1)) * pw + (gx >> 1)) * p.cout + co0 + col] =
                                    f2bf(sm);
                            }
                        }
                    }
                }
            }
            __builtin_amdgcn_s_waitcnt(0);
            __builtin_amdgcn_wave_barrier();
            // part 2: 16-byte stores of 8 channels of one pixel (mask applied here)
            constexpr int CH8 = COUT_T / 8;
#pragma unroll
            for (int k = 0; k < CH8; k++) {
                const int e = lane + k * 64;
                const int px = e / CH8, c8 = e % CH8;
                const int gy = ty0 + wr + (px >> 4), gx = tx0 + wc + (px & 15);
                if (gy >= p.h || gx >= p.w_) continue;
                uint4 v = *reinterpret_cast<const uint4*>(&so[px * OROW + c8 * 8]);
                const size_t off = (((size_t)img * p.h + gy) * p.w_ + gx) * p.cout + co0 + c8 * 8;
                if (p.mask) {
                    const uint4 mk = *reinterpret_cast<const uint4*>(p.mask + off);
                    v.x = mask_pair(v.x, mk.x);
                    v.y = mask_pair(v.y, mk.y);
                    v.z = mask_pair(v.z, mk.z);
                    v.w = mask_pair(v.w, mk.w);
                }
                *reinterpret_cast<uint4*>(p.y + off) = v;
            }
            __syncthreads();
        }
        if (more) {
            VPT_STASH();
            __syncthreads();
        }
    }
#undef VPT_FETCH
#undef VPT_STASH
}

// Weight gradient of a 3x3 convolution:
//   dW[tap][co][ci] = sum_px m[px][co] * X[px + off(tap)][ci],  db[co] = sum_px m[px][co]
// as MFMA products with K = 32 pixels of one output row.  Grid: x = pixel
// slices, y = COUT_T output-channel blocks, z = 32 input-channel blocks.
// Each workgroup reduces its slices' row segments and writes a partial
// [9][COUT_T][32] block (and the bias partial) for the fixed-order reduction.
template <int COUT_T>
__global__ void __launch_bounds__(256) wgrad_kernel(ConvArgs p, const uint16_t* __restrict__ m, int cin,
                                                    float* __restrict__ part, float* __restrict__ part_b)
{
    constexpr int SUB = (COUT_T / 16) * 2;      // 16x16 subtiles of the COUT_T x 32 block
    constexpr int PER_WAVE = SUB / 4;
    __shared__ __attribute__((aligned(16))) uint16_t sM[COUT_T * kRow];
    __shared__ __attribute__((aligned(16))) uint16_t sX[9 * kCi * kRow];   // [dx][dy][ci][px]
    __shared__ float sBias[COUT_T];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lm = lane & 15, lk = (lane >> 4) * 8;
    const int co0 = blockIdx.y * COUT_T, ci0 = blockIdx.z * kCi;
    const int segs_x = (p.w_ + 31) / 32;
    const long total = (long)p.n * p.h * segs_x;

    f32x4 acc[PER_WAVE][9];
#pragma unroll
    for (int s = 0; s < PER_WAVE; s++)
#pragma unroll
        for (int t = 0; t < 9; t++) acc[s][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    float bsum = 0.0f;

    for (long sg = blockIdx.x; sg < total; sg += gridDim.x) {
        const int img = (int)(sg / ((long)p.h * segs_x));
        const int rem = (int)(sg % ((long)p.h * segs_x));
        const int y = rem / segs_x, x0 = (rem % segs_x) * 32;
        __syncthreads();
        // m tile, transposed: sM[co][px]
        for (int e = tid; e < 32 * (COUT_T / 8); e += 256) {
            const int px = e / (COUT_T / 8), q = e % (COUT_T / 8);
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (x0 + px < p.w_)
                v = *reinterpret_cast<const uint4*>(m + (((size_t)img * p.h + y) * p.w_ + x0 + px) * p.cout + co0 +
                                                    q * 8);
            const uint16_t* h = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
            for (int k = 0; k < 8; k++) sM[(q * 8 + k) * kRow + px] = h[k];
        }
        // input rows y-1..y+1, pixels x0-1..x0+32, channels ci0..ci0+31,
        // written as three shifted, transposed copies sX[dx][dy][ci][px]
        for (int e = tid; e < 3 * 34 * 4; e += 256) {
            const int q = e & 3, pos = e >> 2;
            const int dy = pos / 34, xx = pos % 34;   // xx = pixel x0 - 1 + xx
            const uint4 v = load_in8(p, img, y + dy - 1, x0 - 1 + xx, ci0 + q * 8);
            const uint16_t* h = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
            for (int dx = 0; dx < 3; dx++) {
                const int px = xx - dx;   // X[x0 + px + dx - 1] = loaded pixel
                if (px < 0 || px >= 32) continue;
#pragma unroll
                for (int k = 0; k < 8; k++) sX[((dx * 3 + dy) * kCi + q * 8 + k) * kRow + px] = h[k];
            }
        }
        __syncthreads();
        if (blockIdx.z == 0 && tid < COUT_T) {
            float s = 0.0f;
            for (int px = 0; px < 32; px++) s += bf2f(sM[tid * kRow + px]);
            bsum += s;
        }
#pragma unroll
        for (int s = 0; s < PER_WAVE; s++) {
            const int sub = wave * PER_WAVE + s;
            const int cs = sub >> 1, ks = sub & 1;
            const bf16x8 afr = *reinterpret_cast<const bf16x8*>(&sM[(cs * 16 + lm) * kRow + lk]);
#pragma unroll
            for (int t = 0; t < 9; t++) {
                const int dy = t / 3, dx = t % 3;
                const bf16x8 bfr = *reinterpret_cast<const bf16x8*>(&sX[((dx * 3 + dy) * kCi + ks * 16 + lm) * kRow + lk]);
                acc[s][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr, bfr, acc[s][t], 0, 0, 0);
            }
        }
    }
    // partial block: part[slice][tap][co][ci] over this block's (co, ci) range
    const int cout_total = p.cout;
    float* dst = part + (size_t)blockIdx.x * 9 * cout_total * cin;
#pragma unroll
    for (int s = 0; s < PER_WAVE; s++) {
        const int sub = wave * PER_WAVE + s;
        const int cs = sub >> 1, ks = sub & 1;
#pragma unroll
        for (int t = 0; t < 9; t++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int co = co0 + cs * 16 + (lane >> 4) * 4 + i, ci = ci0 + ks * 16 + lm;
                dst[((size_t)t * cout_total + co) * cin + ci] = acc[s][t][i];
            }
    }
    if (blockIdx.z == 0 && tid < COUT_T) part_b[(size_t)blockIdx.x * cout_total + co0 + tid] = bsum;
}

// out[i] = sum_z part[z][i] (z ascending), for n values.
__global__ void reduce_parts_kernel(const float* __restrict__ part, int slices, size_t n, float* __restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int z = 0; z < slices; z++) s += part[(size_t)z * n + i];
    out[i] = s;
}

// wt[tap][ci][co] = w[8 - tap][co][ci]: the data gradient of a convolution
// is the convolution of the output gradient with these weights.
__global__ void flip_transpose_kernel(const uint16_t* __restrict__ w, int cout, int cin, uint16_t* __restrict__ wt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 9 * cout * cin) return;
    const int tap = i / (cout * cin), rem = i % (cout * cin);
    const int co = rem / cin, ci = rem % cin;
    wt[((size_t)(8 - tap) * cin + ci) * cout + co] = w[i];
}

// Output head: out[px][j] = b[j] + sum_c W[j][c] * d0b[px][c] (fp32).
// mode 0: log-space prediction; mode 1: irradiance max(exp(out) - 1, 0).
__global__ void head_kernel(const uint16_t* __restrict__ x, size_t npx, const float* __restrict__ wh,
                            const float* __restrict__ bh, int mode, float* __restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npx) return;
    float v[32];
    const uint4* src = reinterpret_cast<const uint4*>(x + i * 32);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const uint4 u = src[q];
        const uint16_t* h = reinterpret_cast<const uint16_t*>(&u);
#pragma unroll
        for (int k = 0; k < 8; k++) v[q * 8 + k] = bf2f(h[k]);
    }
#pragma unroll
    for (int j = 0; j < 3; j++) {
        float s = bh[j];
        for (int c = 0; c < 32; c++) s += wh[j * 32 + c] * v[c];
        out[i * 3 + j] = mode ? fmaxf(expm1f(s), 0.0f) : s;
    }
}

// Loss and head gradient, one thread per pixel; 256-thread blocks write
// partial sums [block][100]: 96 dW, 3 db, 1 loss.
//   L = mean |pred - target|,  g = sign(pred - target) / (3 npx)
//   m_d0b[px][c] = (sum_j W[j][c] g[px][j]) where d0b > 0
__global__ void __launch_bounds__(256) head_bwd_kernel(const uint16_t* __restrict__ x, size_t npx,
                                                       const float* __restrict__ wh, const float* __restrict__ pred,
                                                       const float* __restrict__ target, uint16_t* __restrict__ m,
                                                       float* __restrict__ part)
{
    __shared__ float red[100][8];
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    float g[3] = {0.0f, 0.0f, 0.0f}, loss = 0.0f;
    float v[32];
    const float inv = 1.0f / (3.0f * (float)npx);
    if (i < npx) {
        const uint4* src = reinterpret_cast<const uint4*>(x + i * 32);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint4 u = src[q];
            const uint16_t* h = reinterpret_cast<const uint16_t*>(&u);
#pragma unroll
            for (int k = 0; k < 8; k++) v[q * 8 + k] = bf2f(h[k]);
        }
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const float d = pred[i * 3 + j] - target[i * 3 + j];
            loss += fabsf(d);
            g[j] = (d > 0.0f ? inv : (d < 0.0f ? -inv : 0.0f));
        }
        uint16_t out[32];
#pragma unroll
        for (int c = 0; c < 32; c++) {
            const float s = (wh[c] * g[0] + wh[32 + c] * g[1]) + wh[64 + c] * g[2];
            out[c] = v[c] > 0.0f ? f2bf(s) : (uint16_t)0;
        }
        uint4* dstm = reinterpret_cast<uint4*>(m + i * 32);
#pragma unroll
        for (int q = 0; q < 4; q++) dstm[q] = *reinterpret_cast<const uint4*>(&out[q * 8]);
    } else {
#pragma unroll
        for (int c = 0; c < 32; c++) v[c] = 0.0f;
    }
    // block reduction of the 100 sums: wave shuffles, then 4 waves in LDS
    for (int k = 0; k < 100; k++) {
        float s;
        if (k < 96) s = g[k / 32] * v[k % 32];
        else if (k < 99) s = g[k - 96];
        else s = loss;
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
        if (lane == 0) red[k][wave] = s;
    }
    __syncthreads();
    if (threadIdx.x < 100) {
        const int k = threadIdx.x;
        part[(size_t)blockIdx.x * 100 + k] = (red[k][0] + red[k][1]) + (red[k][2] + red[k][3]);
    }
}

// Upsample branch of a concatenated-input gradient:
//   out[n][y][x][c] = (sum of g[n][2y+i][2x+j][c_off + c]) where y_low > 0
__global__ void sum2x2_mask_kernel(const uint16_t* __restrict__ g, int cg, int c_off, int c, int n, int h, int w,
                                   const uint16_t* __restrict__ ylow, uint16_t* __restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;   // low-res element
    const int hh = h >> 1, ww = w >> 1;
    if (i >= (size_t)n * hh * ww * c) return;
    const int ch = (int)(i % c);
    const size_t px = i / c;
    const int x = (int)(px % ww), y = (int)((px / ww) % hh), img = (int)(px / ((size_t)ww * hh));
    auto at = [&](int yy, int xx) { return bf2f(g[(((size_t)img * h + yy) * w + xx) * cg + c_off + ch]); };
    const float s = (at(2 * y, 2 * x) + at(2 * y, 2 * x + 1)) + (at(2 * y + 1, 2 * x) + at(2 * y + 1, 2 * x + 1));
    out[i] = bf_pos(ylow[i]) ? f2bf(s) : (uint16_t)0;
}

// Skip branch: out = g[..][c_off + c] (+ 0.25 gpool[n][y/2][x/2][c] when
// gpool is given), kept where y_skip > 0 (when y_skip is given).
__global__ void skip_mask_kernel(const uint16_t* __restrict__ g, int cg, int c_off, int c, int n, int h, int w,
                                 const uint16_t* __restrict__ gpool, const uint16_t* __restrict__ yskip,
                                 uint16_t* __restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * h * w * c) return;
    const int ch = (int)(i % c);
    const size_t px = i / c;
    const int x = (int)(px % w), y = (int)((px / w) % h), img = (int)(px / ((size_t)w * h));
    float s = bf2f(g[px * cg + c_off + ch]);
    if (gpool) s += 0.25f * bf2f(gpool[(((size_t)img * (h >> 1) + (y >> 1)) * (w >> 1) + (x >> 1)) * c + ch]);
    out[i] = (!yskip || bf_pos(yskip[i])) ? f2bf(s) : (uint16_t)0;
}

// Adam on the fp32 master parameters; refreshes the bf16 copies.
__global__ void adam_kernel(float* __restrict__ w, uint16_t* __restrict__ wb, const float* __restrict__ g,
                            float* __restrict__ m1, float* __restrict__ m2, size_t n, float lr, float b1, float b2,
                            float eps, float c1, float c2)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.0f - b1) * gi;
    const float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    const float v = w[i] - lr * (a * c1) / (sqrtf(b * c2) + eps);
    w[i] = v;
    wb[i] = f2bf(v);
}

__global__ void to_bf16_kernel(const float* __restrict__ w, uint16_t* __restrict__ wb, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) wb[i] = f2bf(w[i]);
}

// ---- tracer outputs -> network input / reference -------------------------

struct Decoded {
    float r, g, b;
};

// SHToIrridiance (ColorPassFrag.glsl:356-368) of SH-Y (4) + CoCg with the
// face normal, as the SVGF chain's final decode.
__device__ inline Decoded sh_decode(const float sh[4], float co, float cg, int face)
{
    const vpt::Vec3 v = vpt::face_normal(face, 1.0f);
    const float x = (sh[0] * v.x + sh[1] * v.y) + sh[2] * v.z;
    float Y = 2.0f * (1.023326f * x + 0.886226f * sh[3]);
    Y = fmaxf(Y, 0.0f);
    const float k = (Y * 0.282095f) / (sh[3] + 1e-6f);
    const float c0 = co * k, c1 = cg * k;
    const float T = Y - c1 * 0.5f;
    const float G = c1 + T;
    const float B = T - c0 * 0.5f;
    const float R = B + c0;
    return Decoded{fmaxf(R, 0.0f), fmaxf(G, 0.0f), fmaxf(B, 0.0f)};
}

__device__ inline void texel_values(const vpt_svgf_texel& t, float v[6])
{
#pragma unroll
    for (int k = 0; k < 6; k++) v[k] = vpt::hw_h2f(t.h[k]);
}

// Network input: ch 0-2 log1p(1-spp irradiance), 3-5 albedo of the hit
// block, 6-8 face normal, 9 log1p(hit distance) / 8, 10-31 zero.  Sky
// pixels: irradiance of their texel, zero albedo / normal / depth.
__global__ void assemble_kernel(int npx, const float* __restrict__ gb_t, const uint8_t* __restrict__ gb_face,
                                const uint8_t* __restrict__ gb_block, const vpt_svgf_texel* __restrict__ tex,
                                const vpt::AlbedoTable albedo, uint16_t* __restrict__ x)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npx) return;
    float v[6];
    texel_values(tex[i], v);
    const int face = gb_face[i];
    const Decoded d = sh_decode(v, v[4], v[5], face);
    uint16_t o[32];
#pragma unroll
    for (int k = 0; k < 32; k++) o[k] = 0;
    o[0] = f2bf(log1pf(d.r));
    o[1] = f2bf(log1pf(d.g));
    o[2] = f2bf(log1pf(d.b));
    const float t = gb_t[i];
    if (t >= 0.0f && face < 6) {
        const int blk = gb_block[i] & 127;
        o[3] = f2bf(albedo.v[blk * 3]);
        o[4] = f2bf(albedo.v[blk * 3 + 1]);
        o[5] = f2bf(albedo.v[blk * 3 + 2]);
        const vpt::Vec3 nrm = vpt::face_normal(face, 0.0f);
        o[6] = f2bf(nrm.x);
        o[7] = f2bf(nrm.y);
        o[8] = f2bf(nrm.z);
        o[9] = f2bf(log1pf(t) * 0.125f);
    }
    uint4* dst = reinterpret_cast<uint4*>(x + (size_t)i * 32);
#pragma unroll
    for (int q = 0; q < 4; q++) dst[q] = *reinterpret_cast<const uint4*>(&o[q * 8]);
}

// acc[i][0..5] (+)= SH-Y, Co, Cg of texel i.
__global__ void accumulate_kernel(int npx, const vpt_svgf_texel* __restrict__ tex, int reset, float* __restrict__ acc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npx) return;
    float v[6];
    texel_values(tex[i], v);
#pragma unroll
    for (int k = 0; k < 6; k++) acc[(size_t)i * 6 + k] = (reset ? 0.0f : acc[(size_t)i * 6 + k]) + v[k];
}

// rgb = decode of the texel (tex) or of acc * scale.
__global__ void decode_irr_kernel(int npx, const uint8_t* __restrict__ gb_face, const vpt_svgf_texel* __restrict__ tex,
                                  const float* __restrict__ acc, float scale, float* __restrict__ rgb)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npx) return;
    float v[6];
    if (tex) {
        texel_values(tex[i], v);
    } else {
#pragma unroll
        for (int k = 0; k < 6; k++) v[k] = acc[(size_t)i * 6 + k] * scale;
    }
    const Decoded d = sh_decode(v, v[4], v[5], gb_face[i]);
    rgb[(size_t)i * 3] = d.r;
    rgb[(size_t)i * 3 + 1] = d.g;
    rgb[(size_t)i * 3 + 2] = d.b;
}

// ---- network description ---------------------------------------------------

enum Buf { X, E0A, E0B, P0, E1A, E1B, P1, B0, B1, D1A, D1B, D0A, D0B, NBUF };

struct Layer {
    int in_a, in_b;   // activation buffers (in_b = -1: none)
    int a_up;         // in_a is at half resolution
    int out, pool;    // output buffer, pooled output buffer (-1: none)
    int cin, cout;
    int level;        // resolution level of the output (0 full, 1 half, 2 quarter)
    size_t w_off, b_off;
};

const int kChannels[NBUF] = {32, 32, 32, 32, 64, 64, 64, 128, 128, 64, 64, 32, 32};
const int kLevel[NBUF] = {0, 0, 0, 1, 1, 1, 2, 2, 2, 1, 1, 0, 0};
constexpr int kLayers = 10;
const Layer kNet[kLayers] = {
    {X, -1, 0, E0A, -1, 32, 32, 0, 0, 0},       {E0A, -1, 0, E0B, P0, 32, 32, 0, 0, 0},
    {P0, -1, 0, E1A, -1, 32, 64, 1, 0, 0},      {E1A, -1, 0, E1B, P1, 64, 64, 1, 0, 0},
    {P1, -1, 0, B0, -1, 64, 128, 2, 0, 0},      {B0, -1, 0, B1, -1, 128, 128, 2, 0, 0},
    {B1, E1B, 1, D1A, -1, 192, 64, 1, 0, 0},    {D1A, -1, 0, D1B, -1, 64, 64, 1, 0, 0},
    {D1B, E0B, 1, D0A, -1, 96, 32, 0, 0, 0},    {D0A, -1, 0, D0B, -1, 32, 32, 0, 0, 0},
};

// Persistent workgroups: one per CU and output-channel block (the wide
// kernel holds one workgroup per CU), never more than the tiles.
int g_cus = 0;

int launch_conv(const ConvArgs& a, hipStream_t s)
{
    const int tiles = a.n * ((a.h + 15) / 16) * ((a.w_ + 31) / 32);
    const int cout_t = a.cout % 64 == 0 ? 64 : 32;
    const int blocks_y = a.cout / cout_t;
    const int per_block = std::max(1, g_cus / blocks_y);
    const unsigned gx = (unsigned)std::min(tiles, per_block);
    if (cout_t == 64) {
        hipLaunchKernelGGL(conv3x3_wide_kernel<64>, dim3(gx, (unsigned)blocks_y), dim3(512), 0, s, a);
    } else {
        hipLaunchKernelGGL(conv3x3_wide_kernel<32>, dim3(gx, (unsigned)blocks_y), dim3(512), 0, s, a);
    }
    HIP_TRY(hipGetLastError());
    return VPT_OK;
}

unsigned blocks_for(size_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

}  // namespace

struct vpt_unet {
    vpt_ctx* ctx = nullptr;
    int device = 0;
    int max_n = 0, max_h = 0, max_w = 0;
    Layer net[kLayers];
    size_t n_params = 0, head_w = 0, head_b = 0;
    float* d_w = nullptr;        // fp32 master parameters
    uint16_t* d_wb = nullptr;    // bf16 copies (conv weights read these)
    float* d_g = nullptr;        // gradients
    float* d_m1 = nullptr;
    float* d_m2 = nullptr;
    uint16_t* d_wt = nullptr;    // flipped / transposed weights of one layer
    uint16_t* act[NBUF] = {};    // activations
    uint16_t* grad[NBUF] = {};   // masked gradients (pre-activation) of each buffer
    uint16_t* d_gcat = nullptr;  // data gradient of a concatenated input
    uint16_t* d_gpool[2] = {};   // gradients of P0, P1
    float* d_pred = nullptr;
    float* d_part = nullptr;     // weight-gradient partials
    size_t part_floats = 0;
    float* d_head_part = nullptr;
    int slices = 0;
    int adam_t = 0;
    float last_loss = 0.0f;
};

namespace {

size_t level_px(const vpt_unet* u, int n, int h, int w, int level) { return (size_t)n * (h >> level) * (w >> level); }

int check_dims(const vpt_unet* u, int n, int h, int w, const char* fn)
{
    if (n < 1 || h < 4 || w < 4 || (h & 3) || (w & 3))
        return set_error(VPT_EINVAL, "%s: %dx%dx%d: sizes must be positive multiples of 4", fn, n, h, w);
    if ((size_t)n * h * w > (size_t)u->max_n * u->max_h * u->max_w)
        return set_error(VPT_EINVAL, "%s: %dx%dx%d exceeds the capacity %dx%dx%d", fn, n, h, w, u->max_n, u->max_h,
                         u->max_w);
    return VPT_OK;
}

ConvArgs conv_args(vpt_unet* u, const Layer& L, int n, int h, int w)
{
    ConvArgs a;
    memset(&a, 0, sizeof(a));
    a.a = u->act[L.in_a];
    a.ca = kChannels[L.in_a];
    a.a_up = L.a_up;
    a.b = L.in_b >= 0 ? u->act[L.in_b] : nullptr;
    a.cb = L.in_b >= 0 ? kChannels[L.in_b] : 0;
    a.w = u->d_wb + L.w_off;
    a.bias = u->d_w + L.b_off;
    a.y = u->act[L.out];
    a.pool = L.pool >= 0 ? u->act[L.pool] : nullptr;
    a.relu = 1;
    a.n = n;
    a.h = h >> L.level;
    a.w_ = w >> L.level;
    a.cout = L.cout;
    return a;
}

int forward(vpt_unet* u, int n, int h, int w, hipStream_t s)
{
    for (int l = 0; l < kLayers; l++)
        if (int rc = launch_conv(conv_args(u, u->net[l], n, h, w), s)) return rc;
    return VPT_OK;
}

int wgrad(vpt_unet* u, const Layer& L, int n, int h, int w, const uint16_t* m, hipStream_t s)
{
    ConvArgs a = conv_args(u, L, n, h, w);
    const int cin = L.cin;
    const size_t wsz = (size_t)9 * L.cout * cin;
    const int slices = u->slices;
    if (L.cout % 64 == 0) {
        hipLaunchKernelGGL(wgrad_kernel<64>, dim3((unsigned)slices, (unsigned)(L.cout / 64), (unsigned)(cin / 32)),
                           dim3(256), 0, s, a, m, cin, u->d_part, u->d_part + (size_t)slices * wsz);
    } else {
        hipLaunchKernelGGL(wgrad_kernel<32>, dim3((unsigned)slices, (unsigned)(L.cout / 32), (unsigned)(cin / 32)),
                           dim3(256), 0, s, a, m, cin, u->d_part, u->d_part + (size_t)slices * wsz);
    }
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(reduce_parts_kernel, dim3(blocks_for(wsz)), dim3(256), 0, s, u->d_part, slices, wsz,
                       u->d_g + L.w_off);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(reduce_parts_kernel, dim3(blocks_for(L.cout)), dim3(256), 0, s,
                       u->d_part + (size_t)slices * wsz, slices, (size_t)L.cout, u->d_g + L.b_off);
    HIP_TRY(hipGetLastError());
    return VPT_OK;
}

// Data gradient of layer L's input: g_in = conv(m_out, flipped W^T), with
// the ReLU mask of the input buffer applied in the epilogue when the input
// is a single activation (mask_buf), else raw (concatenated inputs).
int dgrad(vpt_unet* u, const Layer& L, int n, int h, int w, const uint16_t* m, const uint16_t* mask, uint16_t* out,
          hipStream_t s)
{
    const size_t wsz = (size_t)9 * L.cout * L.cin;
    hipLaunchKernelGGL(flip_transpose_kernel, dim3(blocks_for(wsz)), dim3(256), 0, s, u->d_wb + L.w_off, L.cout, L.cin,
                       u->d_wt);
    HIP_TRY(hipGetLastError());
    ConvArgs a;
    memset(&a, 0, sizeof(a));
    a.a = m;
    a.ca = L.cout;
    a.w = u->d_wt;
    a.mask = mask;
    a.y = out;
    a.n = n;
    a.h = h >> L.level;
    a.w_ = w >> L.level;
    a.cout = L.cin;
    return launch_conv(a, s);
}

int backward(vpt_unet* u, const float* d_target, int n, int h, int w, hipStream_t s)
{
    const size_t npx = level_px(u, n, h, w, 0);
    // head: loss, head gradients, masked gradient of D0B
    const unsigned hb = blocks_for(npx);
    hipLaunchKernelGGL(head_bwd_kernel, dim3(hb), dim3(256), 0, s, u->act[D0B], npx, u->d_w + u->head_w, u->d_pred,
                       d_target, u->grad[D0B], u->d_head_part);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(reduce_parts_kernel, dim3(1), dim3(128), 0, s, u->d_head_part, (int)hb, (size_t)100,
                       u->d_part);
    HIP_TRY(hipGetLastError());
    // d_part[0..95] = dW head, [96..98] = db head, [99] = loss sum (lands in d_g[n_params])
    HIP_TRY(hipMemcpyAsync(u->d_g + u->head_w, u->d_part, sizeof(float) * 100, hipMemcpyDeviceToDevice, s));

    const Layer* N = u->net;
    // layer 9: D0A -> D0B
    if (int rc = wgrad(u, N[9], n, h, w, u->grad[D0B], s)) return rc;
    if (int rc = dgrad(u, N[9], n, h, w, u->grad[D0B], u->act[D0A], u->grad[D0A], s)) return rc;
    // layer 8: cat(up(D1B), E0B) -> D0A
    if (int rc = wgrad(u, N[8], n, h, w, u->grad[D0A], s)) return rc;
    if (int rc = dgrad(u, N[8], n, h, w, u->grad[D0A], nullptr, u->d_gcat, s)) return rc;
    {
        const size_t lo = level_px(u, n, h, w, 1) * 64;
        hipLaunchKernelGGL(sum2x2_mask_kernel, dim3(blocks_for(lo)), dim3(256), 0, s, u->d_gcat, 96, 0, 64, n, h, w,
                           u->act[D1B], u->grad[D1B]);
        HIP_TRY(hipGetLastError());
        // the E0B part waits in grad[E0B] (unmasked) until P0's gradient is known
        const size_t sk = npx * 32;
        hipLaunchKernelGGL(skip_mask_kernel, dim3(blocks_for(sk)), dim3(256), 0, s, u->d_gcat, 96, 64, 32, n, h, w,
                           (const uint16_t*)nullptr, (const uint16_t*)nullptr, u->grad[E0A]);
        HIP_TRY(hipGetLastError());
    }
    // layer 7: D1A -> D1B
    if (int rc = wgrad(u, N[7], n, h, w, u->grad[D1B], s)) return rc;
    if (int rc = dgrad(u, N[7], n, h, w, u->grad[D1B], u->act[D1A], u->grad[D1A], s)) return rc;
    // layer 6: cat(up(B1), E1B) -> D1A
    if (int rc = wgrad(u, N[6], n, h, w, u->grad[D1A], s)) return rc;
    if (int rc = dgrad(u, N[6], n, h, w, u->grad[D1A], nullptr, u->d_gcat, s)) return rc;
    {
        const size_t lo = level_px(u, n, h, w, 2) * 128;
        hipLaunchKernelGGL(sum2x2_mask_kernel, dim3(blocks_for(lo)), dim3(256), 0, s, u->d_gcat, 192, 0, 128, n,
                           h >> 1, w >> 1, u->act[B1], u->grad[B1]);
        HIP_TRY(hipGetLastError());
        const size_t sk = level_px(u, n, h, w, 1) * 64;
        hipLaunchKernelGGL(skip_mask_kernel, dim3(blocks_for(sk)), dim3(256), 0, s, u->d_gcat, 192, 128, 64, n, h >> 1,
                           w >> 1, (const uint16_t*)nullptr, (const uint16_t*)nullptr, u->grad[E1A]);
        HIP_TRY(hipGetLastError());
    }
    // layer 5: B0 -> B1
    if (int rc = wgrad(u, N[5], n, h, w, u->grad[B1], s)) return rc;
    if (int rc = dgrad(u, N[5], n, h, w, u->grad[B1], u->act[B0], u->grad[B0], s)) return rc;
    // layer 4: P1 -> B0; gradient of P1 (no ReLU on a pooled buffer)
    if (int rc = wgrad(u, N[4], n, h, w, u->grad[B0], s)) return rc;
    if (int rc = dgrad(u, N[4], n, h, w, u->grad[B0], nullptr, u->d_gpool[1], s)) return rc;
    {
        // E1B: skip gradient (held in grad[E1A]) + pooled gradient, masked by E1B
        const size_t sk = level_px(u, n, h, w, 1) * 64;
        hipLaunchKernelGGL(skip_mask_kernel, dim3(blocks_for(sk)), dim3(256), 0, s, u->grad[E1A], 64, 0, 64, n, h >> 1,
                           w >> 1, u->d_gpool[1], u->act[E1B], u->grad[E1B]);
        HIP_TRY(hipGetLastError());
    }
    // layer 3: E1A -> E1B
    if (int rc = wgrad(u, N[3], n, h, w, u->grad[E1B], s)) return rc;
    if (int rc = dgrad(u, N[3], n, h, w, u->grad[E1B], u->act[E1A], u->grad[E1A], s)) return rc;
    // layer 2: P0 -> E1A
    if (int rc = wgrad(u, N[2], n, h, w, u->grad[E1A], s)) return rc;
    if (int rc = dgrad(u, N[2], n, h, w, u->grad[E1A], nullptr, u->d_gpool[0], s)) return rc;
    {
        const size_t sk = npx * 32;
        hipLaunchKernelGGL(skip_mask_kernel, dim3(blocks_for(sk)), dim3(256), 0, s, u->grad[E0A], 32, 0, 32, n, h, w,
                           u->d_gpool[0], u->act[E0B], u->grad[E0B]);
        HIP_TRY(hipGetLastError());
    }
    // layer 1: E0A -> E0B
    if (int rc = wgrad(u, N[1], n, h, w, u->grad[E0B], s)) return rc;
    if (int rc = dgrad(u, N[1], n, h, w, u->grad[E0B], u->act[E0A], u->grad[E0A], s)) return rc;
    // layer 0: X -> E0A
    return wgrad(u, N[0], n, h, w, u->grad[E0A], s);
}

}  // namespace

// ---- C ABI -------------------------------------------------------------------

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

int unet_upload(vpt_unet* u, const float* host)
{
    HIP_TRY(hipMemcpy(u->d_w, host, sizeof(float) * u->n_params, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(u->n_params)), dim3(256), 0, nullptr, u->d_w, u->d_wb,
                       u->n_params);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemset(u->d_m1, 0, sizeof(float) * u->n_params));
    HIP_TRY(hipMemset(u->d_m2, 0, sizeof(float) * u->n_params));
    HIP_TRY(hipDeviceSynchronize());
    u->adam_t = 0;
    return VPT_OK;
}

}  // namespace

extern "C" VPT_API int vpt_unet_destroy(vpt_unet* u)
{
    if (!u) return VPT_OK;
    DeviceGuard g(u->device);
    (void)hipFree(u->d_w);
    (void)hipFree(u->d_wb);
    (void)hipFree(u->d_g);
    (void)hipFree(u->d_m1);
    (void)hipFree(u->d_m2);
    (void)hipFree(u->d_wt);
    for (int b = 0; b < NBUF; b++) {
        (void)hipFree(u->act[b]);
        (void)hipFree(u->grad[b]);
    }
    (void)hipFree(u->d_gcat);
    (void)hipFree(u->d_gpool[0]);
    (void)hipFree(u->d_gpool[1]);
    (void)hipFree(u->d_pred);
    (void)hipFree(u->d_part);
    (void)hipFree(u->d_head_part);
    delete u;
    return VPT_OK;
}

extern "C" VPT_API int vpt_unet_create(vpt_ctx* ctx, int max_n, int max_h, int max_w, vpt_unet** out)
{
    if (!ctx || !out) return set_error(VPT_EINVAL, "vpt_unet_create: null argument");
    if (max_n < 1 || max_h < 4 || max_w < 4 || (max_h & 3) || (max_w & 3) ||
        (size_t)max_n * max_h * max_w > ((size_t)1 << 26))
        return set_error(VPT_EINVAL, "vpt_unet_create: capacity %dx%dx%d (multiples of 4, at most 2^26 pixels)",
                         max_n, max_h, max_w);
    DeviceGuard g(ctx->device);
    vpt_unet* u = new vpt_unet();
    if (!g_cus) g_cus = vpt::device_cus(ctx->device);
    u->ctx = ctx;
    u->device = ctx->device;
    u->max_n = max_n;
    u->max_h = max_h;
    u->max_w = max_w;
    size_t off = 0, max_w_layer = 0;
    for (int l = 0; l < kLayers; l++) {
        u->net[l] = kNet[l];
        u->net[l].w_off = off;
        off += (size_t)9 * kNet[l].cout * kNet[l].cin;
        u->net[l].b_off = off;
        off += (size_t)kNet[l].cout;
        max_w_layer = std::max(max_w_layer, (size_t)9 * kNet[l].cout * kNet[l].cin);
    }
    u->head_w = off;
    u->head_b = off + 96;
    u->n_params = off + 99;
    const size_t full = (size_t)max_n * max_h * max_w;
    u->slices = 256;
    u->part_floats = (size_t)u->slices * (max_w_layer + 128);
    hipError_t e = hipSuccess;
    auto alloc = [&](void** p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, std::max<size_t>(bytes, 256));
    };
    alloc((void**)&u->d_w, sizeof(float) * u->n_params);
    alloc((void**)&u->d_wb, sizeof(uint16_t) * u->n_params);
    alloc((void**)&u->d_g, sizeof(float) * (u->n_params + 1));
    alloc((void**)&u->d_m1, sizeof(float) * u->n_params);
    alloc((void**)&u->d_m2, sizeof(float) * u->n_params);
    alloc((void**)&u->d_wt, sizeof(uint16_t) * max_w_layer);
    for (int b = 0; b < NBUF; b++) {
        const size_t px = full >> (2 * kLevel[b]);
        alloc((void**)&u->act[b], sizeof(uint16_t) * px * kChannels[b]);
        if (b != X && b != P0 && b != P1) alloc((void**)&u->grad[b], sizeof(uint16_t) * px * kChannels[b]);
    }
    alloc((void**)&u->d_gcat, sizeof(uint16_t) * full * 96);
    alloc((void**)&u->d_gpool[0], sizeof(uint16_t) * (full >> 2) * 32);
    alloc((void**)&u->d_gpool[1], sizeof(uint16_t) * (full >> 4) * 64);
    alloc((void**)&u->d_pred, sizeof(float) * full * 3);
    alloc((void**)&u->d_part, sizeof(float) * u->part_floats);
    alloc((void**)&u->d_head_part, sizeof(float) * 100 * ((full + 255) / 256));
    if (e != hipSuccess) {
        vpt_unet_destroy(u);
        return set_error(VPT_ENOMEM, "vpt_unet_create: allocation failed (%s)", hipGetErrorString(e));
    }
    *out = u;
    return VPT_OK;
}

extern "C" VPT_API int vpt_unet_param_count(const vpt_unet* u, size_t* n)
{
    if (!u || !n) return set_error(VPT_EINVAL, "vpt_unet_param_count: null argument");
    *n = u->n_params;
    return VPT_OK;
}

extern "C" VPT_API int vpt_unet_init(vpt_unet* u, uint32_t seed)
{
    if (!u) return set_error(VPT_EINVAL, "vpt_unet_init: null network");
    std::vector<float> w(u->n_params, 0.0f);
    std::mt19937 rng(seed);
    for (int l = 0; l < kLayers; l++) {
        const Layer& L = u->net[l];
        std::normal_distribution<float> nd(0.0f, sqrtf(2.0f / (9.0f * (float)L.cin)));   // He initialisation
        for (size_t i = 0; i < (size_t)9 * L.cout * L.cin; i++) w[L.w_off + i] = nd(rng);
    }
    std::normal_distribution<float> nh(0.0f, sqrtf(1.0f / 32.0f));
    for (int i = 0; i < 96; i++) w[u->head_w + i] = nh(rng);
    DeviceGuard g(u->device);
    return unet_upload(u, w.data());
}

extern "C" VPT_API int vpt_unet_set_params(vpt_unet* u, const float* host, size_t n)
{
    if (!u || !host || n != u->n_params)
        return set_error(VPT_EINVAL, "vpt_unet_set_params: expected %zu floats", u ? u->n_params : (size_t)0);
    DeviceGuard g(u->device);
    return unet_upload(u, host);
}

extern "C" VPT_API int vpt_unet_get_params(vpt_unet* u, float* host, size_t n)
{
    if (!u || !host || n != u->n_params)
        return set_error(VPT_EINVAL, "vpt_unet_get_params: expected %zu floats", u ? u->n_params : (size_t)0);
    DeviceGuard g(u->device);
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(host, u->d_w, sizeof(float) * n, hipMemcpyDeviceToHost));
    return VPT_OK;
}

extern "C" VPT_API int vpt_unet_get_grads(vpt_unet* u, float* host, size_t n)
{
    if (!u || !host || n != u->n_params)
        return set_error(VPT_EINVAL, "vpt_unet_get_grads: expected %zu floats", u ? u->n_params : (size_t)0);
    DeviceGuard g(u->device);
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(host, u->d_g, sizeof(float) * n, hipMemcpyDeviceToHost));
    return VPT_OK;
}

extern "C" VPT_API int vpt_unet_forward(vpt_unet* u, const uint16_t* d_x, int n, int h, int w, int mode, float* d_out,
                                        void* stream)
{
    if (!u || !d_x || !d_out) return set_error(VPT_EINVAL, "vpt_unet_forward: null argument");
    if (int rc = check_dims(u, n, h, w, "vpt_unet_forward")) return rc;
    if (mode != 0 && mode != 1) return set_error(VPT_EINVAL, "vpt_unet_forward: mode 0 (log space) or 1 (irradiance)");
    DeviceGuard g(u->device);
    hipStream_t s = (hipStream_t)stream;
    const size_t npx = (size_t)n * h * w;
    HIP_TRY(hipMemcpyAsync(u->act[X], d_x, sizeof(uint16_t) * npx * 32, hipMemcpyDeviceToDevice, s));
    if (int rc = forward(u, n, h, w, s)) return rc;
    hipLaunchKernelGGL(head_kernel, dim3(blocks_for(npx)), dim3(256), 0, s, u->act[D0B], npx, u->d_w + u->head_w,
                       u->d_w + u->head_b, mode, d_out);
    HIP_TRY(hipGetLastError());
    return VPT_OK;
}

extern "C" VPT_API int vpt_unet_backward(vpt_unet* u, const uint16_t* d_x, const float* d_target, int n, int h, int w,
                                         float* loss, void* stream)
{
    if (!u || !d_x || !d_target) return set_error(VPT_EINVAL, "vpt_unet_backward: null argument");
    if (int rc = check_dims(u, n, h, w, "vpt_unet_backward")) return rc;
    DeviceGuard g(u->device);
    hipStream_t s = (hipStream_t)stream;
    const size_t npx = (size_t)n * h * w;
    HIP_TRY(hipMemcpyAsync(u->act[X], d_x, sizeof(uint16_t) * npx * 32, hipMemcpyDeviceToDevice, s));
    if (int rc = forward(u, n, h, w, s)) return rc;
    hipLaunchKernelGGL(head_kernel, dim3(blocks_for(npx)), dim3(256), 0, s, u->act[D0B], npx, u->d_w + u->head_w,
                       u->d_w + u->head_b, 0, u->d_pred);
    HIP_TRY(hipGetLastError());
    if (int rc = backward(u, d_target, n, h, w, s)) return rc;
    if (loss) {
        float sum = 0.0f;
        HIP_TRY(hipMemcpyAsync(&sum, u->d_g + u->n_params, sizeof(float), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        *loss = sum / (3.0f * (float)npx);
    }
    return VPT_OK;
}

extern "C" VPT_API int vpt_unet_adam_step(vpt_unet* u, float lr, void* stream)
{
    if (!u || !(lr > 0.0f)) return set_error(VPT_EINVAL, "vpt_unet_adam_step: bad arguments");
    DeviceGuard g(u->device);
    const float b1 = 0.9f, b2 = 0.999f;
    u->adam_t++;
    const float c1 = 1.0f / (1.0f - powf(b1, (float)u->adam_t));
    const float c2 = 1.0f / (1.0f - powf(b2, (float)u->adam_t));
    hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(u->n_params)), dim3(256), 0, (hipStream_t)stream, u->d_w, u->d_wb,
                       u->d_g, u->d_m1, u->d_m2, u->n_params, lr, b1, b2, 1e-8f, c1, c2);
    HIP_TRY(hipGetLastError());
    return VPT_OK;
}

extern "C" VPT_API int vpt_unet_assemble(vpt_unet* u, const vpt_gbuffer* gb, const vpt_svgf_texel* d_tex,
                                         uint16_t* d_x, void* stream)
{
    if (!u || !gb || !d_tex || !d_x || !gb->d_t || !gb->d_face || !gb->d_block)
        return set_error(VPT_EINVAL, "vpt_unet_assemble: null argument");
    vpt::AlbedoTable albedo;
    {
        std::lock_guard<std::mutex> lk(u->ctx->mu);
        if (!u->ctx->albedo_valid)
            return set_error(VPT_ESTATE, "vpt_unet_assemble: no albedo table yet (trace GI first)");
        memcpy(albedo.v, u->ctx->albedo_host, sizeof(albedo.v));
    }
    DeviceGuard g(u->device);
    const int npx = gb->width * gb->height;
    hipLaunchKernelGGL(assemble_kernel, dim3(blocks_for((size_t)npx)), dim3(256), 0, (hipStream_t)stream, npx,
                       gb->d_t, gb->d_face, gb->d_block, d_tex, albedo, d_x);
    HIP_TRY(hipGetLastError());
    return VPT_OK;
}

extern "C" VPT_API int vpt_unet_denoise(vpt_unet* u, const vpt_gbuffer* gb, const vpt_svgf_texel* d_tex, float* d_rgb,
                                        void* stream)
{
    if (!u || !gb || !d_rgb) return set_error(VPT_EINVAL, "vpt_unet_denoise: null argument");
    if (int rc = check_dims(u, 1, gb->height, gb->width, "vpt_unet_denoise")) return rc;
    if (int rc = vpt_unet_assemble(u, gb, d_tex, u->act[X], stream)) return rc;
    DeviceGuard g(u->device);
    hipStream_t s = (hipStream_t)stream;
    const size_t npx = (size_t)gb->width * gb->height;
    if (int rc = forward(u, 1, gb->height, gb->width, s)) return rc;
    hipLaunchKernelGGL(head_kernel, dim3(blocks_for(npx)), dim3(256), 0, s, u->act[D0B], npx, u->d_w + u->head_w,
                       u->d_w + u->head_b, 1, d_rgb);
    HIP_TRY(hipGetLastError());
    return VPT_OK;
}

extern "C" VPT_API int vpt_texel_accumulate(const vpt_svgf_texel* d_tex, int npx, float* d_acc, int reset, void* stream)
{
    if (!d_tex || !d_acc || npx < 0) return set_error(VPT_EINVAL, "vpt_texel_accumulate: bad arguments");
    if (npx == 0) return VPT_OK;
    hipLaunchKernelGGL(accumulate_kernel, dim3(blocks_for((size_t)npx)), dim3(256), 0, (hipStream_t)stream, npx,
                       d_tex, reset, d_acc);
    HIP_TRY(hipGetLastError());
    return VPT_OK;
}

extern "C" VPT_API int vpt_irradiance_decode(const vpt_gbuffer* gb, const vpt_svgf_texel* d_tex, const float* d_acc,
                                             float acc_scale, float* d_rgb, void* stream)
{
    if (!gb || !gb->d_face || !d_rgb || (!d_tex && !d_acc))
        return set_error(VPT_EINVAL, "vpt_irradiance_decode: null argument");
    const int npx = gb->width * gb->height;
    hipLaunchKernelGGL(decode_irr_kernel, dim3(blocks_for((size_t)npx)), dim3(256), 0, (hipStream_t)stream, npx,
                       gb->d_face, d_tex, d_acc, acc_scale, d_rgb);
    HIP_TRY(hipGetLastError());
    return VPT_OK;
}

extern "C" VPT_API int vpt_unet_read_activation(vpt_unet* u, int index, int n, int h, int w, uint16_t* host)
{
    if (!u || !host || index < 0 || index >= NBUF) return set_error(VPT_EINVAL, "vpt_unet_read_activation: bad arguments");
    if (int rc = check_dims(u, n, h, w, "vpt_unet_read_activation")) return rc;
    DeviceGuard g(u->device);
    const size_t count = ((size_t)n * (h >> kLevel[index]) * (w >> kLevel[index])) * kChannels[index];
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(host, u->act[index], sizeof(uint16_t) * count, hipMemcpyDeviceToHost));
    return VPT_OK;
}
