// Internal object layouts of the C ABI (host side).  Device pointers are
// plain pointers here so that this header is usable from host-only files.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <mutex>
#include <vector>

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

// One contiguous run of a served request's rays inside a persistent launch:
// launch-wide ray ids [begin, next segment's begin) are the request's rays
// first, first + 1, ...
struct ServeSegment {
    const vpt_ray* rays;   // the request's rays
    vpt_hit* hits;         // its results
    uint8_t* vis;          // GI requests: sun visibility at the bounce hit
    uint32_t begin;        // launch-wide id of the segment's first ray
    uint32_t first;        // request-local index of that ray
    uint32_t world;        // world slot
    uint32_t kind;         // VPT_REQ_PRIMARY / SHADOW / GI
    int32_t max_iter;
    int32_t shadow_iter;   // GI: continuation cap
    float sun[3];          // GI: continuation direction
    uint32_t pad;
};

// Enqueues one persistent launch over `total` rays split in `nseg` segments
// (d_segs, and d_begin = their begin ids followed by `total`), reading
// worlds from the table d_worlds of `nworlds` views.  d_head: one zeroed
// queue counter, reset by the launch itself.
int serve_launch(int cus, const vpt_launch_cfg& cfg, const void* d_worlds, uint32_t nworlds,
                 const ServeSegment* d_segs, const uint32_t* d_begin, uint32_t nseg, uint32_t total, uint32_t* d_head,
                 void* stream);

// Number of compute units of `device` (256 when the query fails).
int device_cus(int device);
// Whether `cfg` is a launch shape the persistent kernels accept.
bool launch_cfg_valid(const vpt_launch_cfg& cfg);
// Looks up, with `device` current, the device address of the fixed scratch
// array of vpt_trace_gi_noalloc on that device (once per device).
int gi_noalloc_init(int device);

}  // namespace vpt

namespace vpt {

// Scratch of the GI pass for one stream, one allocation: the queue
// counters (first 256 bytes), then the bounce continuations and the
// shadow-ray jobs (16 bytes each, at most one of each per pixel).  Calls on
// one stream run in order, so they share it; calls on other streams get
// their own.
// `done` is recorded after a call's last kernel and waited for (on the
// device) before the next call's first, so a stream handle reused after its
// stream was destroyed never writes the scratch while earlier work still
// reads it.  Buffers an entry outgrows, or an evicted entry's, are retired
// (kept until the context is destroyed), never freed: a graph captured
// earlier may still replay with them.
struct GiScratch {
    void* stream = nullptr;
    void* done = nullptr;    // hipEvent_t
    uint8_t* d_buf = nullptr;
    size_t cap = 0;          // pixels
    uint64_t last_use = 0;
};

// Most streams a context keeps GI scratch for; the least recently used
// one is freed beyond that.
constexpr size_t kMaxGiScratch = 8;

}  // namespace vpt

struct vpt_ctx {
    int device = 0;
    // guards the members below: the serving worker thread and the caller's
    // threads may trace on one context
    std::mutex mu;
    // the albedo table of the last GI trace (the learned denoiser's input
    // assembly reads it); kernels receive their table by value
    float albedo_host[128 * 3] = {};
    bool albedo_valid = false;
    std::vector<vpt::GiScratch> gi_scratch;
    std::vector<void*> gi_retired;
    uint64_t gi_clock = 0;
    // vpt_trace_frame: the stream the shadow pass runs on beside GI, and the
    // fork / join events (created on first use, default priority)
    void* side_stream = nullptr;
    void* ev_fork = nullptr;
    void* ev_join = nullptr;
};

struct vpt_world : vpt::World {
    vpt_ctx* ctx = nullptr;
    int32_t* d_dfbox = nullptr;   // incremental DF update: box found on the device (8 ints)
};
