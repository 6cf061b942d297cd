// HIP-graph replay of a host's per-frame sequence of calls of this library
// (vpt_graph_*).  The host brackets its calls with vpt_graph_begin /
// vpt_graph_launch; the calls enqueue on an internal capture stream (the
// caller's stream may be the legacy null stream, which cannot be captured),
// ordered against the caller's stream with events.  The first launch
// instantiates the captured graph; later launches update the instantiated
// graph in place with the new capture's kernel arguments (camera, ping-pong
// planes, frame constants), which keeps the per-frame host cost to a
// capture and an update, and re-instantiate only when the sequence changed
// shape.
#include <hip/hip_runtime.h>

#include "vpt_internal.h"

using vpt::set_error;

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return set_error(VPT_EDEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));      \
    } while (0)

struct vpt_graph {
    int device = 0;
    hipStream_t cap = nullptr;        // capture and replay stream (default priority)
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    hipGraphExec_t exec = nullptr;
    bool capturing = false;
    vpt_graph_stats st{};
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

}  // namespace

extern "C" VPT_API int vpt_graph_destroy(vpt_graph* g)
{
    if (!g) return VPT_OK;
    DeviceGuard dg(g->device);
    if (g->capturing) {
        hipGraph_t graph = nullptr;
        (void)hipStreamEndCapture(g->cap, &graph);
        if (graph) (void)hipGraphDestroy(graph);
    }
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->ev_in) (void)hipEventDestroy(g->ev_in);
    if (g->ev_out) (void)hipEventDestroy(g->ev_out);
    if (g->cap) (void)hipStreamDestroy(g->cap);
    delete g;
    return VPT_OK;
}

extern "C" VPT_API int vpt_graph_create(vpt_ctx* ctx, vpt_graph** out)
{
    if (!ctx || !out) return set_error(VPT_EINVAL, "vpt_graph_create: null argument");
    DeviceGuard dg(ctx->device);
    vpt_graph* g = new vpt_graph();
    g->device = ctx->device;
    hipError_t e = hipStreamCreateWithFlags(&g->cap, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&g->ev_in, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&g->ev_out, hipEventDisableTiming);
    if (e != hipSuccess) {
        vpt_graph_destroy(g);
        return set_error(VPT_EDEVICE, "vpt_graph_create: %s", hipGetErrorString(e));
    }
    *out = g;
    return VPT_OK;
}

extern "C" VPT_API int vpt_graph_begin(vpt_graph* g, void* stream, void** capture_stream)
{
    if (!g || !capture_stream) return set_error(VPT_EINVAL, "vpt_graph_begin: null argument");
    if (g->capturing) return set_error(VPT_ESTATE, "vpt_graph_begin: already capturing");
    DeviceGuard dg(g->device);
    // the replay must not start before the caller's earlier work
    HIP_TRY(hipEventRecord(g->ev_in, (hipStream_t)stream));
    HIP_TRY(hipStreamWaitEvent(g->cap, g->ev_in, 0));
    HIP_TRY(hipStreamBeginCapture(g->cap, hipStreamCaptureModeRelaxed));
    g->capturing = true;
    *capture_stream = (void*)g->cap;
    return VPT_OK;
}

extern "C" VPT_API int vpt_graph_launch(vpt_graph* g, void* stream)
{
    if (!g) return set_error(VPT_EINVAL, "vpt_graph_launch: null graph");
    if (!g->capturing) return set_error(VPT_ESTATE, "vpt_graph_launch: vpt_graph_begin was not called");
    DeviceGuard dg(g->device);
    g->capturing = false;
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(g->cap, &graph);
    if (e != hipSuccess || !graph) {
        if (graph) (void)hipGraphDestroy(graph);
        return set_error(VPT_EDEVICE, "vpt_graph_launch: capture failed: %s", hipGetErrorString(e));
    }
    bool fresh = true;
    if (g->exec) {
        hipGraphNode_t bad = nullptr;
        hipGraphExecUpdateResult res = hipGraphExecUpdateError;
        if (hipGraphExecUpdate(g->exec, graph, &bad, &res) == hipSuccess && res == hipGraphExecUpdateSuccess) {
            fresh = false;
            g->st.updates++;
        } else {
            (void)hipGetLastError();   // the failed update's error is not the caller's
            (void)hipGraphExecDestroy(g->exec);
            g->exec = nullptr;
        }
    }
    if (fresh) {
        e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            g->exec = nullptr;
            (void)hipGraphDestroy(graph);
            return set_error(VPT_EDEVICE, "vpt_graph_launch: instantiation failed: %s", hipGetErrorString(e));
        }
        g->st.instantiations++;
    }
    size_t nodes = 0;
    if (hipGraphGetNodes(graph, nullptr, &nodes) == hipSuccess) g->st.nodes = (uint32_t)nodes;
    (void)hipGraphDestroy(graph);
    HIP_TRY(hipGraphLaunch(g->exec, g->cap));
    HIP_TRY(hipEventRecord(g->ev_out, g->cap));
    HIP_TRY(hipStreamWaitEvent((hipStream_t)stream, g->ev_out, 0));
    g->st.launches++;
    return VPT_OK;
}

extern "C" VPT_API int vpt_graph_abort(vpt_graph* g)
{
    if (!g) return set_error(VPT_EINVAL, "vpt_graph_abort: null graph");
    if (!g->capturing) return VPT_OK;
    DeviceGuard dg(g->device);
    g->capturing = false;
    hipGraph_t graph = nullptr;
    (void)hipStreamEndCapture(g->cap, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipGetLastError();
    return VPT_OK;
}

extern "C" VPT_API int vpt_graph_stats_get(const vpt_graph* g, vpt_graph_stats* st)
{
    if (!g || !st) return set_error(VPT_EINVAL, "vpt_graph_stats_get: null argument");
    *st = g->st;
    return VPT_OK;
}
