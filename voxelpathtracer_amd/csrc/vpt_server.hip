// Serving loop over resident worlds.
//
// Clients submit ray-query and frame requests from any thread; one worker
// thread per server admits them in arrival order while their result memory
// fits in the server's arena (one device block carved first-fit), runs them,
// and stamps arrival / admission / start / completion times on one steady
// clock.
//
// Plain mode: one request at a time through the one-launch kernels, each
// waited for before the next starts.
//
// Batched mode (continuous batching): each launch takes rays from every
// admitted request at once (a fair share each, then whatever room is left in
// admission order; a request larger than a launch spans several) and traces
// them in one kernel: one lane per ray (serve_direct_kernel, the default) or
// the persistent refill kernel (launch shape bit 1 clear).  Up to `max_inflight` launches are queued on the device;
// while they run the worker admits new arrivals and builds the next launch,
// so requests that arrive mid-launch join the following one.  Frame requests
// run through the persistent frame chain after the launch's ray kernel.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "vpt_device.h"
#include "vpt_gi.h"
#include "vpt_internal.h"

using vpt::ServeSegment;
using vpt::set_error;
using vpt::WorldView;

namespace {

constexpr size_t kGranule = 256;
constexpr uint32_t kMaxSegs = 1024;      // segments (request parts) per launch
constexpr int kMaxInflight = 4;

size_t round_up(size_t x) { return (x + kGranule - 1) / kGranule * kGranule; }

// First-fit allocator over one device block.
class Arena {
public:
    uint8_t* base = nullptr;
    size_t size = 0, used = 0, peak = 0;

    bool alloc(size_t bytes, size_t* off)
    {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const size_t rest = it->second - bytes;
            const size_t at = it->first + bytes;
            free_.erase(it);
            if (rest) free_[at] = rest;
            used += bytes;
            peak = std::max(peak, used);
            return true;
        }
        return false;
    }

    void release(size_t off, size_t bytes)
    {
        used -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

    void reset(size_t bytes)
    {
        size = bytes;
        used = peak = 0;
        free_.clear();
        free_[0] = bytes;
    }

private:
    std::map<size_t, size_t> free_;   // offset -> length
};

enum class State { PENDING, ADMITTED, DONE };

struct Req {
    uint64_t ticket = 0;
    vpt_request r{};
    vpt_camera cam{};
    vpt_gi_params gi{};
    uint32_t units = 0;        // rays, or pixels of a frame
    uint32_t scheduled = 0;    // units handed to launches
    uint32_t completed = 0;    // units whose launches retired
    size_t off = 0, bytes = 0;
    State state = State::PENDING;
    int status = VPT_OK;
    vpt_request_result res{};
};

struct Part {
    Req* req;
    uint32_t first, count;
    vpt_frame* frame;      // frame requests: the chain that runs them
    vpt_world* world;
};

// A launch slot: its own stream (launches in different slots overlap on
// the device, so one launch's tail is filled by the next), world table,
// segment table and staging.
struct Slot {
    hipStream_t stream = nullptr;
    WorldView* d_worlds = nullptr;
    bool table_dirty = true;
    ServeSegment* d_segs = nullptr;
    uint32_t* d_begin = nullptr;
    uint32_t* d_head = nullptr;
    ServeSegment* h_segs = nullptr;   // pinned staging
    uint32_t* h_begin = nullptr;
    hipEvent_t ev = nullptr;
    bool busy = false;
};

struct Launch {
    int slot = 0;
    vpt_launch_cfg cfg{};
    uint32_t nworlds = 0;
    std::vector<Part> parts;
};

int64_t clock_ns()
{
    return (int64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
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

bool is_ray_kind(int k) { return k == VPT_REQ_PRIMARY || k == VPT_REQ_SHADOW || k == VPT_REQ_GI; }

// Result memory of a request, laid out in 256-byte aligned pieces.
size_t result_bytes(const vpt_request& r, uint32_t units)
{
    if (r.kind == VPT_REQ_FRAME)
        return round_up((size_t)units * 4) + 2 * round_up(units) + round_up((size_t)units * 4) + round_up(units) +
               round_up((size_t)units * 12);
    size_t b = round_up((size_t)units * sizeof(vpt_hit));
    if (r.kind == VPT_REQ_GI) b += round_up(units);
    return b;
}

}  // namespace

struct vpt_server {
    vpt_ctx* ctx = nullptr;
    int device = 0;
    int cus = 256;
    vpt_server_cfg cfg{};
    vpt_launch_cfg launch{256, 16, 256, 32, 2};   // one lane per ray (bit 1); the persistent kernel is optional

    std::vector<vpt_world*> worlds;
    std::vector<int> world_refs;
    std::vector<WorldView> world_table;
    uint64_t world_bytes = 0;

    // persistent frame chains, one per (slot, W, H): slot << 56 | W << 28 | H
    std::map<uint64_t, vpt_frame*> frames;

    Arena arena;
    Slot slots[kMaxInflight];

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Req*> pending;                 // arrival order
    std::vector<Req*> admitted;               // admission order, not yet completed
    std::unordered_map<uint64_t, Req*> reqs;
    std::deque<Launch> inflight;
    uint64_t next_ticket = 1;
    bool stop = false;
    bool woken = false;
    std::thread worker;

    vpt_server_stats st{};

    // ---- worker ------------------------------------------------------------

    void admit()
    {
        while (!pending.empty()) {
            Req* q = pending.front();
            size_t off = 0;
            if (!arena.alloc(q->bytes, &off)) break;   // first come, first admitted
            pending.pop_front();
            q->off = off;
            q->state = State::ADMITTED;
            q->res.t_admit_ns = clock_ns();
            layout(q);
            admitted.push_back(q);
        }
    }

    void layout(Req* q)
    {
        uint8_t* p = arena.base + q->off;
        vpt_request_result& res = q->res;
        res.kind = q->r.kind;
        res.n = q->units;
        if (q->r.kind == VPT_REQ_FRAME) {
            const uint32_t n = q->units;
            res.gbuffer.width = q->cam.width;
            res.gbuffer.height = q->cam.height;
            res.gbuffer.d_t = (float*)p;
            p += round_up((size_t)n * 4);
            res.gbuffer.d_face = p;
            p += round_up(n);
            res.gbuffer.d_block = p;
            p += round_up(n);
            res.gbuffer.d_voxel = (uint32_t*)p;
            p += round_up((size_t)n * 4);
            res.d_mask = p;
            p += round_up(n);
            res.d_rgb = (float*)p;
        } else {
            res.d_hits = (vpt_hit*)p;
            p += round_up((size_t)q->units * sizeof(vpt_hit));
            if (q->r.kind == VPT_REQ_GI) res.d_vis = p;
        }
    }

    void complete(Req* q, int status)
    {
        q->state = State::DONE;
        if (status && !q->status) q->status = status;
        q->res.status = q->status;
        q->res.t_done_ns = clock_ns();
        auto it = std::find(admitted.begin(), admitted.end(), q);
        if (it != admitted.end()) admitted.erase(it);
        st.completed++;
        cv_done.notify_all();
    }

    void retire_front(int status)
    {
        Launch L = std::move(inflight.front());
        inflight.pop_front();
        slots[L.slot].busy = false;
        for (const Part& p : L.parts) {
            Req* q = p.req;
            if (status && !q->status) q->status = status;
            q->completed += p.count;
            if (q->completed == q->units) complete(q, VPT_OK);
        }
    }

    bool has_unscheduled() const
    {
        for (const Req* q : admitted)
            if (q->scheduled < q->units) return true;
        return false;
    }

    // Chooses the next launch's parts (called with the lock held).
    bool build(Launch& L)
    {
        int slot = -1;
        for (int i = 0; i < cfg.max_inflight; i++)
            if (!slots[i].busy) {
                slot = i;
                break;
            }
        if (slot < 0) return false;
        L.slot = slot;
        L.cfg = launch;
        L.nworlds = (uint32_t)worlds.size();
        L.parts.clear();
        std::vector<Req*> rays;
        bool frame_taken = false;
        for (Req* q : admitted) {
            if (q->scheduled >= q->units) continue;
            if (q->r.kind == VPT_REQ_FRAME) {
                if (frame_taken) continue;   // one frame per launch
                const auto fit = frames.find(frame_key(slot, q->cam));
                if (fit == frames.end()) {
                    // cannot happen (submit creates every slot's scratch or
                    // refuses the request); fail the request, not the process
                    q->scheduled = q->units;
                    q->completed = q->units;
                    complete(q, set_error(VPT_ESTATE, "vpt_server: no frame scratch for slot %d", slot));
                    return build(L);
                }
                frame_taken = true;
                L.parts.push_back(Part{q, 0, q->units, fit->second, worlds[q->r.world]});
                q->scheduled = q->units;
                q->res.t_start_ns = clock_ns();
            } else {
                rays.push_back(q);
            }
        }
        uint64_t room = cfg.max_launch_rays;
        std::vector<uint32_t> take(rays.size(), 0);
        if (!rays.empty()) {
            // a fair share for each request, then the rest in admission order
            const uint64_t share = std::max<uint64_t>(room / rays.size(), 4096);
            for (size_t i = 0; i < rays.size() && room; i++) {
                const uint64_t t = std::min<uint64_t>({(uint64_t)(rays[i]->units - rays[i]->scheduled), share, room});
                take[i] = (uint32_t)t;
                room -= t;
            }
            for (size_t i = 0; i < rays.size() && room; i++) {
                const uint64_t left = rays[i]->units - rays[i]->scheduled - take[i];
                const uint64_t t = std::min<uint64_t>(left, room);
                take[i] += (uint32_t)t;
                room -= t;
            }
        }
        uint32_t nseg = 0;
        for (size_t i = 0; i < rays.size(); i++) {
            if (!take[i] || nseg == kMaxSegs) continue;
            Req* q = rays[i];
            if (q->scheduled == 0) q->res.t_start_ns = clock_ns();
            L.parts.push_back(Part{q, q->scheduled, take[i], nullptr, worlds[q->r.world]});
            q->scheduled += take[i];
            nseg++;
        }
        // rays of one world next to each other in the launch: the wavefronts
        // in flight at any moment then walk few worlds' distance fields
        std::stable_sort(L.parts.begin(), L.parts.end(), [](const Part& a, const Part& b) {
            const bool fa = a.req->r.kind == VPT_REQ_FRAME, fb = b.req->r.kind == VPT_REQ_FRAME;
            if (fa != fb) return fb;   // frames last (they run after the ray kernel anyway)
            return a.req->r.world < b.req->r.world;
        });
        return !L.parts.empty();
    }

    // Enqueues a built launch (called without the lock; the parts' requests
    // are not touched by other threads until the launch retires).
    int issue(const Launch& L, const std::vector<WorldView>* table)
    {
        Slot& sl = slots[L.slot];
        hipStream_t stream = sl.stream;
        if (table) {
            if (hipMemcpyAsync(sl.d_worlds, table->data(), sizeof(WorldView) * table->size(), hipMemcpyHostToDevice,
                               stream) != hipSuccess)
                return set_error(VPT_EDEVICE, "vpt_server: world table upload failed");
        }
        uint32_t nseg = 0, total = 0;
        for (const Part& p : L.parts) {
            const Req* q = p.req;
            if (q->r.kind == VPT_REQ_FRAME) continue;
            ServeSegment& s = sl.h_segs[nseg];
            s.rays = q->r.d_rays;
            s.hits = q->res.d_hits;
            s.vis = q->res.d_vis;
            s.begin = total;
            s.first = p.first;
            s.world = (uint32_t)q->r.world;
            s.kind = (uint32_t)q->r.kind;
            s.max_iter = q->r.max_iter;
            s.shadow_iter = q->r.shadow_iter;
            s.sun[0] = q->r.sun_dir[0];
            s.sun[1] = q->r.sun_dir[1];
            s.sun[2] = q->r.sun_dir[2];
            s.pad = 0;
            sl.h_begin[nseg] = total;
            total += p.count;
            nseg++;
        }
        if (nseg) {
            sl.h_begin[nseg] = total;
            if (hipMemcpyAsync(sl.d_segs, sl.h_segs, sizeof(ServeSegment) * nseg, hipMemcpyHostToDevice, stream) !=
                    hipSuccess ||
                hipMemcpyAsync(sl.d_begin, sl.h_begin, sizeof(uint32_t) * (nseg + 1), hipMemcpyHostToDevice,
                               stream) != hipSuccess)
                return set_error(VPT_EDEVICE, "vpt_server: segment upload failed");
            if (int rc = vpt::serve_launch(cus, L.cfg, sl.d_worlds, L.nworlds, sl.d_segs, sl.d_begin, nseg, total,
                                           sl.d_head, stream))
                return rc;
        }
        for (const Part& p : L.parts) {
            const Req* q = p.req;
            if (q->r.kind != VPT_REQ_FRAME) continue;
            if (int rc = vpt_frame_trace(p.frame, p.world, &q->cam, &q->gi, q->r.primary_iter,
                                         q->r.frame_shadow_iter, &q->res.gbuffer, q->res.d_mask, q->res.d_rgb,
                                         nullptr, 0, stream))
                return rc;
        }
        if (hipEventRecord(sl.ev, stream) != hipSuccess) return set_error(VPT_EDEVICE, "vpt_server: event record failed");
        return VPT_OK;
    }

    int run_plain(Req* q, vpt_world* w)
    {
        hipStream_t stream = slots[0].stream;
        const vpt_request& r = q->r;
        int rc = VPT_OK;
        switch (r.kind) {
        case VPT_REQ_PRIMARY:
        case VPT_REQ_SHADOW:
            rc = vpt_trace_rays(w, r.d_rays, (int)r.n, r.max_iter, q->res.d_hits, stream);
            break;
        case VPT_REQ_GI:
            rc = vpt_trace_gi_rays(w, r.d_rays, (int)r.n, r.max_iter, r.sun_dir, r.shadow_iter, q->res.d_hits,
                                   q->res.d_vis, stream);
            break;
        default:
            rc = vpt_trace_primary(w, &q->cam, r.primary_iter, &q->res.gbuffer, stream);
            if (!rc)
                rc = vpt_trace_shadow(w, &q->cam, &q->res.gbuffer, q->gi.sun_dir, r.frame_shadow_iter, q->res.d_mask,
                                      stream);
            if (!rc) rc = vpt_trace_gi(w, &q->cam, &q->res.gbuffer, &q->gi, q->res.d_rgb, nullptr, stream);
            break;
        }
        if (!rc && hipStreamSynchronize(stream) != hipSuccess)
            rc = set_error(VPT_EDEVICE, "vpt_server: request failed on the device");
        return rc;
    }

    // The world table for a launch in `slot` when that slot's copy is stale.
    const std::vector<WorldView>* take_table(int slot, std::vector<WorldView>& copy)
    {
        if (!slots[slot].table_dirty) return nullptr;
        copy = world_table;
        slots[slot].table_dirty = false;
        return &copy;
    }

    void tables_dirty()
    {
        for (Slot& sl : slots) sl.table_dirty = true;
    }

    static uint64_t frame_key(int slot, const vpt_camera& c)
    {
        return ((uint64_t)slot << 56) | ((uint64_t)c.width << 28) | (uint64_t)c.height;
    }

    void run()
    {
        DeviceGuard g(device);
        std::unique_lock<std::mutex> lk(mu);
        std::vector<WorldView> table_copy;
        while (true) {
            // retire the launches that have finished
            while (!inflight.empty()) {
                hipError_t e = hipEventQuery(slots[inflight.front().slot].ev);
                if (e == hipErrorNotReady) break;
                retire_front(e == hipSuccess ? VPT_OK : VPT_EDEVICE);
            }
            admit();
            if (cfg.mode == VPT_SERVE_PLAIN) {
                Req* q = nullptr;
                for (Req* a : admitted)
                    if (a->scheduled == 0) {
                        q = a;
                        break;
                    }
                if (q) {
                    q->scheduled = q->units;
                    q->res.t_start_ns = clock_ns();
                    vpt_world* w = worlds[q->r.world];
                    lk.unlock();
                    const int rc = run_plain(q, w);
                    lk.lock();
                    st.launches++;
                    q->completed = q->units;
                    complete(q, rc);
                    continue;
                }
            } else if ((int)inflight.size() < cfg.max_inflight && has_unscheduled()) {
                Launch L;
                if (build(L)) {
                    slots[L.slot].busy = true;
                    const std::vector<WorldView>* tbl = take_table(L.slot, table_copy);
                    lk.unlock();
                    const int rc = issue(L, tbl);
                    lk.lock();
                    st.launches++;
                    inflight.push_back(std::move(L));
                    if (rc) {
                        // the device stream is in an unknown state: wait it out, fail the parts
                        lk.unlock();
                        for (int i = 0; i < cfg.max_inflight; i++) (void)hipStreamSynchronize(slots[i].stream);
                        lk.lock();
                        while (!inflight.empty()) retire_front(rc);
                    }
                    continue;
                }
            }
            if (stop && admitted.empty() && inflight.empty()) {
                // shutting down: what cannot be admitted any more fails
                while (!pending.empty()) {
                    Req* q = pending.front();
                    pending.pop_front();
                    complete(q, VPT_ENOMEM);
                }
                break;
            }
            if (!inflight.empty()) {
                const bool can_launch = (int)inflight.size() < cfg.max_inflight;
                if (!can_launch) {
                    hipEvent_t e = slots[inflight.front().slot].ev;
                    lk.unlock();
                    (void)hipEventSynchronize(e);
                    lk.lock();
                } else {
                    // poll the oldest launch; new work wakes the loop at once
                    woken = false;
                    cv_work.wait_for(lk, std::chrono::microseconds(20), [this] { return woken; });
                }
            } else {
                woken = false;
                cv_work.wait(lk, [this] { return woken || stop; });
            }
        }
    }

    void wake()
    {
        woken = true;
        cv_work.notify_one();
    }
};

extern "C" VPT_API int64_t vpt_now_ns(void) { return clock_ns(); }

extern "C" VPT_API int vpt_server_destroy(vpt_server* s)
{
    if (!s) return VPT_OK;
    if (s->worker.joinable()) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
            s->wake();
        }
        s->worker.join();
    }
    DeviceGuard g(s->device);
    for (auto& kv : s->reqs) delete kv.second;
    for (auto& kv : s->frames) vpt_frame_destroy(kv.second);
    for (Slot& sl : s->slots) {
        (void)hipFree(sl.d_segs);
        (void)hipFree(sl.d_begin);
        (void)hipFree(sl.d_head);
        (void)hipHostFree(sl.h_segs);
        (void)hipHostFree(sl.h_begin);
        (void)hipFree(sl.d_worlds);
        if (sl.ev) (void)hipEventDestroy(sl.ev);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    (void)hipFree(s->arena.base);
    delete s;
    return VPT_OK;
}

extern "C" VPT_API int vpt_server_create(vpt_ctx* ctx, const vpt_server_cfg* cfg, vpt_server** out)
{
    if (!ctx || !cfg || !out) return set_error(VPT_EINVAL, "vpt_server_create: null argument");
    if (cfg->mode != VPT_SERVE_PLAIN && cfg->mode != VPT_SERVE_BATCHED)
        return set_error(VPT_EINVAL, "vpt_server_create: unknown mode %d", cfg->mode);
    if (cfg->max_worlds < 1 || cfg->max_worlds > 256)
        return set_error(VPT_EINVAL, "vpt_server_create: max_worlds in [1, 256]");
    if (cfg->arena_bytes < kGranule) return set_error(VPT_EINVAL, "vpt_server_create: arena too small");
    if (cfg->max_launch_rays < 4096 || cfg->max_launch_rays > (1u << 30))
        return set_error(VPT_EINVAL, "vpt_server_create: max_launch_rays in [4096, 2^30]");
    if (cfg->max_inflight < 1 || cfg->max_inflight > kMaxInflight)
        return set_error(VPT_EINVAL, "vpt_server_create: max_inflight in [1, %d]", kMaxInflight);
    DeviceGuard g(ctx->device);
    vpt_server* s = new vpt_server();
    s->ctx = ctx;
    s->device = ctx->device;
    s->cus = vpt::device_cus(ctx->device);
    s->cfg = *cfg;
    s->worlds.assign(cfg->max_worlds, nullptr);
    s->world_refs.assign(cfg->max_worlds, 0);
    s->world_table.resize(cfg->max_worlds);
    for (WorldView& v : s->world_table) {
        memset(&v, 0, sizeof(v));
        v.xmax = v.ymax = v.zmax = -1.0f;   // empty slot: every ray leaves at once
    }
    const size_t arena = round_up(cfg->arena_bytes);
    hipError_t e = hipMalloc(&s->arena.base, arena);
    for (int i = 0; i < cfg->max_inflight && e == hipSuccess; i++) {
        Slot& sl = s->slots[i];
        e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipMalloc(&sl.d_worlds, sizeof(WorldView) * cfg->max_worlds);
        if (e == hipSuccess) e = hipMalloc(&sl.d_segs, sizeof(ServeSegment) * kMaxSegs);
        if (e == hipSuccess) e = hipMalloc(&sl.d_begin, sizeof(uint32_t) * (kMaxSegs + 1));
        if (e == hipSuccess) e = hipMalloc(&sl.d_head, sizeof(uint32_t) * 4);
        if (e == hipSuccess) e = hipHostMalloc(&sl.h_segs, sizeof(ServeSegment) * kMaxSegs, hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostMalloc(&sl.h_begin, sizeof(uint32_t) * (kMaxSegs + 1), hipHostMallocDefault);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming);
    }
    if (e != hipSuccess) {
        vpt_server_destroy(s);
        return set_error(VPT_ENOMEM, "vpt_server_create: allocation failed (%s)", hipGetErrorString(e));
    }
    s->arena.reset(arena);
    s->st.arena_bytes = arena;
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return VPT_OK;
}

extern "C" VPT_API int vpt_server_add_world(vpt_server* s, vpt_world* w, int* slot)
{
    if (!s || !w || !slot) return set_error(VPT_EINVAL, "vpt_server_add_world: null argument");
    if (w->device != s->device) return set_error(VPT_EINVAL, "vpt_server_add_world: world on another device");
    std::lock_guard<std::mutex> lk(s->mu);
    int free_slot = -1;
    for (int i = 0; i < (int)s->worlds.size(); i++) {
        if (s->worlds[i] == w) return set_error(VPT_EINVAL, "vpt_server_add_world: world already resident");
        if (!s->worlds[i] && free_slot < 0) free_slot = i;
    }
    if (free_slot < 0) return set_error(VPT_ENOMEM, "vpt_server_add_world: all %d slots in use", (int)s->worlds.size());
    if (s->world_bytes + w->bytes > s->cfg.world_budget)
        return set_error(VPT_ENOMEM, "vpt_server_add_world: %zu more bytes exceed the world budget (%llu of %llu used)",
                         w->bytes, (unsigned long long)s->world_bytes, (unsigned long long)s->cfg.world_budget);
    s->worlds[free_slot] = w;
    s->world_table[free_slot] = vpt::make_view(w);
    s->tables_dirty();
    s->world_bytes += w->bytes;
    s->st.world_bytes = s->world_bytes;
    *slot = free_slot;
    return VPT_OK;
}

extern "C" VPT_API int vpt_server_remove_world(vpt_server* s, int slot)
{
    if (!s) return set_error(VPT_EINVAL, "vpt_server_remove_world: null server");
    std::lock_guard<std::mutex> lk(s->mu);
    if (slot < 0 || slot >= (int)s->worlds.size() || !s->worlds[slot])
        return set_error(VPT_EINVAL, "vpt_server_remove_world: slot %d holds no world", slot);
    if (s->world_refs[slot] > 0)
        return set_error(VPT_EBUSY, "vpt_server_remove_world: %d requests still name slot %d", s->world_refs[slot],
                         slot);
    s->world_bytes -= s->worlds[slot]->bytes;
    s->st.world_bytes = s->world_bytes;
    s->worlds[slot] = nullptr;
    memset(&s->world_table[slot], 0, sizeof(WorldView));
    s->world_table[slot].xmax = s->world_table[slot].ymax = s->world_table[slot].zmax = -1.0f;
    s->tables_dirty();
    return VPT_OK;
}

extern "C" VPT_API int vpt_server_set_launch_cfg(vpt_server* s, const vpt_launch_cfg* cfg)
{
    if (!s || !cfg) return set_error(VPT_EINVAL, "vpt_server_set_launch_cfg: null argument");
    if (!vpt::launch_cfg_valid(*cfg)) return set_error(VPT_EINVAL, "vpt_server_set_launch_cfg: invalid launch shape");
    std::lock_guard<std::mutex> lk(s->mu);
    s->launch = *cfg;
    return VPT_OK;
}

extern "C" VPT_API int vpt_server_submit(vpt_server* s, const vpt_request* req, uint64_t* ticket)
{
    if (!s || !req || !ticket) return set_error(VPT_EINVAL, "vpt_server_submit: null argument");
    const vpt_request& r = *req;
    if (!is_ray_kind(r.kind) && r.kind != VPT_REQ_FRAME)
        return set_error(VPT_EINVAL, "vpt_server_submit: unknown kind %d", r.kind);
    uint32_t units = 0;
    if (r.kind == VPT_REQ_FRAME) {
        if (!r.cam || !r.gi) return set_error(VPT_EINVAL, "vpt_server_submit: frame without camera or GI constants");
        if (int rc = vpt::check_cam(r.cam, "vpt_server_submit")) return rc;
        if (r.primary_iter < 0 || r.primary_iter > 65535 || r.frame_shadow_iter < 0 || r.frame_shadow_iter > 65535 ||
            r.gi->gi_iter < 0 || r.gi->gi_iter > 65535 || r.gi->shadow_iter < 0 || r.gi->shadow_iter > 65535)
            return set_error(VPT_EINVAL, "vpt_server_submit: traversal cap out of range");
        units = (uint32_t)r.cam->width * (uint32_t)r.cam->height;
    } else {
        if (r.n > (1u << 30)) return set_error(VPT_EINVAL, "vpt_server_submit: %u rays is too many", r.n);
        if (r.n > 0 && !r.d_rays) return set_error(VPT_EINVAL, "vpt_server_submit: null ray list");
        if (r.max_iter < 0 || r.max_iter > 65535 ||
            (r.kind == VPT_REQ_GI && (r.shadow_iter < 0 || r.shadow_iter > 65535)))
            return set_error(VPT_EINVAL, "vpt_server_submit: traversal cap out of range");
        units = r.n;
    }
    Req* q = new Req();
    q->r = r;
    if (r.kind == VPT_REQ_FRAME) {
        q->cam = *r.cam;
        q->gi = *r.gi;
        q->r.cam = nullptr;
        q->r.gi = nullptr;
    }
    q->units = units;
    q->bytes = std::max(result_bytes(r, units), kGranule);
    DeviceGuard g(s->device);
    std::lock_guard<std::mutex> lk(s->mu);
    if (r.world < 0 || r.world >= (int)s->worlds.size() || !s->worlds[r.world]) {
        delete q;
        return set_error(VPT_EINVAL, "vpt_server_submit: world slot %d is empty", r.world);
    }
    if (q->bytes > s->arena.size) {
        const size_t b = q->bytes;
        delete q;
        return set_error(VPT_ENOMEM, "vpt_server_submit: %zu result bytes exceed the arena (%zu)", b, s->arena.size);
    }
    if (r.kind == VPT_REQ_FRAME && !s->frames.count(vpt_server::frame_key(0, q->cam))) {
        // scratch of the persistent frame chain for this resolution, one per
        // launch slot (slots run concurrently), kept resident
        const uint64_t px = units;
        const uint64_t scratch = (px * (2 * sizeof(vpt_qray) + 7) + 4096) * (uint64_t)s->cfg.max_inflight;
        if (s->world_bytes + scratch > s->cfg.world_budget) {
            delete q;
            return set_error(VPT_ENOMEM, "vpt_server_submit: frame scratch exceeds the world budget");
        }
        // all slots or none: a partial set would leave later requests of this
        // size without scratch in some slot
        std::vector<vpt_frame*> made;
        for (int slot = 0; slot < s->cfg.max_inflight; slot++) {
            vpt_frame* f = nullptr;
            if (int rc = vpt_frame_create(s->ctx, q->cam.width, q->cam.height, &f)) {
                for (vpt_frame* m : made) vpt_frame_destroy(m);
                delete q;
                return rc;
            }
            made.push_back(f);
        }
        for (int slot = 0; slot < s->cfg.max_inflight; slot++) s->frames[vpt_server::frame_key(slot, q->cam)] = made[slot];
        s->world_bytes += scratch;
        s->st.world_bytes = s->world_bytes;
    }
    q->ticket = s->next_ticket++;
    q->res.t_arrival_ns = r.arrival_ns ? r.arrival_ns : clock_ns();
    q->r.arrival_ns = q->res.t_arrival_ns;
    s->reqs[q->ticket] = q;
    s->world_refs[r.world]++;
    s->st.submitted++;
    if (r.kind == VPT_REQ_FRAME) s->st.frames++;
    else s->st.rays += r.n;
    if (units == 0) {
        // nothing to trace: complete at once
        q->state = State::DONE;
        q->res.kind = r.kind;
        q->res.t_admit_ns = q->res.t_start_ns = q->res.t_done_ns = clock_ns();
        s->st.completed++;
        *ticket = q->ticket;
        s->cv_done.notify_all();
        return VPT_OK;
    }
    s->pending.push_back(q);
    s->wake();
    *ticket = q->ticket;
    return VPT_OK;
}

extern "C" VPT_API int vpt_server_wait(vpt_server* s, uint64_t ticket, int timeout_ms, vpt_request_result* res)
{
    if (!s || !res) return set_error(VPT_EINVAL, "vpt_server_wait: null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return set_error(VPT_EINVAL, "vpt_server_wait: unknown ticket %llu", (unsigned long long)ticket);
    Req* q = it->second;
    auto done = [q] { return q->state == State::DONE; };
    if (timeout_ms < 0) {
        s->cv_done.wait(lk, done);
    } else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done)) {
        return set_error(VPT_ETIMEOUT, "vpt_server_wait: request %llu not complete after %d ms",
                         (unsigned long long)ticket, timeout_ms);
    }
    *res = q->res;
    return VPT_OK;
}

extern "C" VPT_API int vpt_server_release(vpt_server* s, uint64_t ticket)
{
    if (!s) return set_error(VPT_EINVAL, "vpt_server_release: null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end())
        return set_error(VPT_EINVAL, "vpt_server_release: unknown ticket %llu", (unsigned long long)ticket);
    Req* q = it->second;
    if (q->state != State::DONE)
        return set_error(VPT_ESTATE, "vpt_server_release: request %llu has not completed", (unsigned long long)ticket);
    if (q->units) s->arena.release(q->off, q->bytes);
    s->world_refs[q->r.world]--;
    s->reqs.erase(it);
    delete q;
    s->wake();   // freed memory may admit a waiting request
    return VPT_OK;
}

extern "C" VPT_API int vpt_server_stats_get(vpt_server* s, vpt_server_stats* st)
{
    if (!s || !st) return set_error(VPT_EINVAL, "vpt_server_stats_get: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *st = s->st;
    st->arena_used = s->arena.used;
    st->arena_peak = s->arena.peak;
    st->pending = (uint32_t)s->pending.size();
    st->active = (uint32_t)s->admitted.size();
    return VPT_OK;
}
