// Internal object layouts of the C ABI (host side).  Device pointers are
// plain pointers here so that this header is usable from host-only files.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "vpt/vpt.h"

namespace vpt {

int set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

// A world in HBM.  Grid and DF live in 8x8x8 bricks (512 B) ordered
// bx-fastest, with a 9-bit Morton index inside each brick, so that the
// 128-byte line fetched for one DF sample covers an 8x4x4 block of
// neighbouring cells.
struct World {
    int device = 0;
    int nx = 0, ny = 0, nz = 0;
    int nbx = 0, nby = 0, nbz = 0;
    uint8_t* d_grid_lin = nullptr;    // linear x-fastest copy (API layout, edits)
    uint8_t* d_grid = nullptr;        // bricked block ids
    uint8_t* d_df = nullptr;          // bricked distance field
    uint8_t* d_tmp = nullptr;         // DF build intermediate (linear)
    unsigned long long* d_counter = nullptr;  // optional DF-sample counter (caller-owned)
    size_t bytes = 0;
    size_t voxels() const { return (size_t)nx * ny * nz; }
};

}  // namespace vpt

struct vpt_ctx {
    int device = 0;
    float* d_albedo = nullptr;       // GI albedo table (128 ids x RGB)
    float albedo_host[128 * 3] = {};
    bool albedo_valid = false;
};

struct vpt_world : vpt::World {
    vpt_ctx* ctx = nullptr;
};
