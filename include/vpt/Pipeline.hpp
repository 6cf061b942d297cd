// C++ façade for the hot-path slice of the reference's frame loop
// (MainPipeline::StartPipeline, Core/Pipeline.cpp:1960-3060): primary trace
// into a ping-pong G-buffer (InitialTraceFBO / InitialTraceFBOPrev,
// :1970-1971), diffuse GI trace (:2171-2281), SVGF (:2284-2615), the sun
// shadow trace with its transversal (:2790-2852), the sun-shadow denoise
// (:2854-2948), the reflection trace at the reference's settings with its hit
// shading from the denoised planes (:3000-3060), and the colour pass's
// composite (:3611-3700).  Device buffers are plain hipMalloc allocations
// owned by the pipeline.  Header-only; links against libvpt.so and the HIP
// runtime.
#pragma once

#include <hip/hip_runtime_api.h>

#include <stdexcept>

#include "World.hpp"
#include "vpt.h"

namespace VoxelRT {

class Pipeline {
public:
    // persistent: trace primary, sun shadow and GI with the persistent engine
    // (vpt_frame_trace) instead of one launch per pass (vpt_trace_primary,
    // vpt_trace_gi; the faster path).  Either way the shadow pass with its
    // transversal, the shadow denoise and the reflection rays
    // (vpt_trace_reflect_deferred) run on a side stream beside GI and SVGF;
    // the reflection shading (vpt_shade_reflect) and the composite follow
    // both.
    // use_graph (one launch per pass only): after the primary pass, the side
    // chain, the main chain and the tail are each captured and replayed as a
    // HIP graph (vpt_graph_*) every frame, the instantiated graphs updated in
    // place.
    Pipeline(Context& ctx, World& world, int width, int height, float sun_tick = 50.0f, bool persistent = false,
             bool use_graph = false)
        : m_Ctx(ctx.Handle()), m_World(world), m_W(width), m_H(height), m_SunTick(sun_tick), m_Persistent(persistent)
    {
        if (use_graph && persistent)
            throw std::invalid_argument("Pipeline: the persistent engine replays its own graph; use_graph is for "
                                        "the one-launch passes");
        const size_t px = (size_t)width * height;
        for (auto& g : m_GBuf) {
            Alloc(&g.d_t, px * sizeof(float));
            Alloc(&g.d_face, px);
            Alloc(&g.d_block, px);
            Alloc(&g.d_voxel, px * sizeof(uint32_t));
            g.width = width;
            g.height = height;
        }
        Alloc(&m_Mask, px);
        Alloc(&m_Transversal, px * sizeof(uint16_t));
        Alloc(&m_ShadowOut, px);
        m_RDiv = width >= 4 && height >= 4 ? 4 : 1;   // ReflectionTraceResolution 0.25
        m_RW = width / m_RDiv;
        m_RH = height / m_RDiv;
        Alloc(&m_ReflDist, (size_t)m_RW * m_RH * sizeof(uint16_t) + 16);
        Alloc(&m_ReflHit, (size_t)m_RW * m_RH * sizeof(uint32_t) + 16);
        Alloc(&m_ReflColor, (size_t)m_RW * m_RH * 4 * sizeof(uint16_t) + 16);
        {
            vpt_reflect_params rp;
            const float up[3] = {0.0f, 1.0f, 0.0f};
            vpt_check(vpt_reflect_params_default(up, 0, &rp));
            Alloc(&m_ReflRecords, (size_t)rp.spp * m_RW * m_RH * VPT_REFLECT_RECORD_BYTES);
        }
        Alloc(&m_Color, px * 3 * sizeof(float));
        Alloc(&m_RGB, px * 3 * sizeof(float));
        Alloc(&m_Denoised, px * 3 * sizeof(float));
        Alloc(&m_Texels, px * sizeof(vpt_svgf_texel));
        vpt_svgf_params sp;
        vpt_check(vpt_svgf_params_default(&sp));
        vpt_check(vpt_svgf_create(ctx.Handle(), width, height, &sp, &m_Svgf));
        vpt_check(vpt_shadow_denoise_create(ctx.Handle(), width, height, &m_ShadowDn));
        if (hipStreamCreateWithFlags(&m_Side, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&m_Fork, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&m_Join, hipEventDisableTiming) != hipSuccess)
            throw std::runtime_error("Pipeline: side stream");
        if (persistent) vpt_check(vpt_frame_create(ctx.Handle(), width, height, &m_Frame));
        if (use_graph)
            for (vpt_graph*& g : m_Graphs) vpt_check(vpt_graph_create(ctx.Handle(), &g));
    }

    ~Pipeline()
    {
        if (m_Side) (void)hipStreamSynchronize(m_Side);
        for (vpt_graph* g : m_Graphs)
            if (g) vpt_graph_destroy(g);
        if (m_Frame) vpt_frame_destroy(m_Frame);
        vpt_shadow_denoise_destroy(m_ShadowDn);
        if (m_Side) (void)hipStreamDestroy(m_Side);
        if (m_Fork) (void)hipEventDestroy(m_Fork);
        if (m_Join) (void)hipEventDestroy(m_Join);
        vpt_svgf_destroy(m_Svgf);
        for (auto& g : m_GBuf) {
            (void)hipFree(g.d_t);
            (void)hipFree(g.d_face);
            (void)hipFree(g.d_block);
            (void)hipFree(g.d_voxel);
        }
        (void)hipFree(m_Mask);
        (void)hipFree(m_Transversal);
        (void)hipFree(m_ShadowOut);
        (void)hipFree(m_ReflDist);
        (void)hipFree(m_ReflHit);
        (void)hipFree(m_ReflColor);
        (void)hipFree(m_ReflRecords);
        (void)hipFree(m_Color);
        (void)hipFree(m_RGB);
        (void)hipFree(m_Denoised);
        (void)hipFree(m_Texels);
    }

    Pipeline(const Pipeline&) = delete;
    Pipeline& operator=(const Pipeline&) = delete;

    // One frame for a camera at `pos` looking along (yaw, pitch) degrees.
    void RenderFrame(const float pos[3], float yaw, float pitch, void* stream = nullptr)
    {
        vpt_camera cam;
        float view[16], proj[16];
        vpt_check(vpt_camera_make(pos, yaw, pitch, 60.0f, m_W, m_H, 0.1f, 1000.0f, &cam, view, proj));
        vpt_gbuffer& cur = m_GBuf[m_Frame_ % 2];
        vpt_gbuffer& prev = m_Frame_ == 0 ? cur : m_GBuf[(m_Frame_ + 1) % 2];
        if (m_Frame_ == 0) {
            for (int i = 0; i < 16; i++) {
                m_PrevView[i] = view[i];
                m_PrevProj[i] = proj[i];
            }
        }
        vpt_gi_params prm;
        vpt_check(vpt_gi_params_default(m_SunTick, (uint32_t)m_Frame_, 0x5EED0003u, &prm));
        hipStream_t s = (hipStream_t)stream;
        if (m_Persistent)
            vpt_check(vpt_frame_trace(m_Frame, m_World.Handle(), &cam, &prm, 350, 350, &cur, m_Mask, m_RGB, m_Texels,
                                      1, stream));
        else
            vpt_check(vpt_trace_primary(m_World.Handle(), &cam, 350, &cur, stream));
        float prev_vp[16], cur_vp[16];
        ViewProj(m_PrevView, m_PrevProj, prev_vp);
        ViewProj(view, proj, cur_vp);
        // the side chain (shadow with its transversal, shadow denoise,
        // reflection rays) reads only the G-buffer and the shadow pass and
        // runs beside the main chain (GI, SVGF); the tail (reflection hits
        // shaded from both chains' outputs, composite) follows both
        vpt_reflect_params rp;
        vpt_check(vpt_reflect_params_default(prm.sun_dir, (uint32_t)m_Frame_, &rp));
        rp.div = m_RDiv;
        auto side_chain = [&](void* ss) {
            vpt_check(vpt_trace_shadow2(m_World.Handle(), &cam, &cur, prm.sun_dir, 350, m_Mask, m_Transversal, ss));
            vpt_check(vpt_shadow_denoise_frame(m_ShadowDn, &cam, prev_vp, &cur, &prev, m_Mask, m_Transversal, 1.0f,
                                               m_ShadowOut, ss));
            vpt_check(vpt_trace_reflect_deferred(m_World.Handle(), &cam, &cur, &rp, cur_vp, m_ReflDist, m_ReflHit,
                                                 m_ReflRecords, ss));
        };
        auto main_chain = [&](void* ms) {
            if (!m_Persistent) vpt_check(vpt_trace_gi(m_World.Handle(), &cam, &cur, &prm, m_RGB, m_Texels, ms));
            vpt_check(vpt_svgf_frame(m_Svgf, &cam, prev_vp, &cur, &prev, m_Texels, (float)m_Frame_ / 60.0f,
                                     m_Denoised, nullptr, ms));
        };
        auto tail = [&](void* ts) {
            const vpt_svgf_texel* diffuse = nullptr;
            vpt_check(vpt_svgf_final(m_Svgf, &diffuse));
            vpt_reflect_shading rs;
            vpt_check(vpt_reflect_shading_from_gi(&prm, cur_vp, diffuse, m_ShadowOut, &rs));
            vpt_check(vpt_shade_reflect(m_Ctx, &cam, &cur, &rp, &rs, m_ReflRecords, m_ReflHit, m_ReflColor, ts));
            vpt_composite_params cp;
            vpt_check(vpt_composite_params_from_gi(&prm, &cp));
            vpt_check(vpt_composite(m_Ctx, &cam, &cur, &cp, diffuse, m_ShadowOut, m_ReflColor, m_RDiv, m_Color, ts));
        };
        // the main chain (the longer one) is enqueued first; the side chain
        // starts from the fork event after the primary pass
        if (hipEventRecord(m_Fork, s) != hipSuccess || hipStreamWaitEvent(m_Side, m_Fork, 0) != hipSuccess)
            throw std::runtime_error("Pipeline: fork");
        if (!m_Graphs[0]) {
            main_chain(stream);
            side_chain(m_Side);
            if (hipEventRecord(m_Join, m_Side) != hipSuccess) throw std::runtime_error("Pipeline: join");
            if (hipStreamWaitEvent(s, m_Join, 0) != hipSuccess) throw std::runtime_error("Pipeline: join");
            tail(stream);
        } else {
            // each chain is captured and replayed as a graph of its own,
            // its completion joined on the stream it started from (the first
            // two overlap); the tail's replay waits for both
            Replay(m_Graphs[1], stream, stream, main_chain);
            Replay(m_Graphs[0], m_Side, m_Side, side_chain);
            if (hipEventRecord(m_Join, m_Side) != hipSuccess) throw std::runtime_error("Pipeline: join");
            if (hipStreamWaitEvent(s, m_Join, 0) != hipSuccess) throw std::runtime_error("Pipeline: join");
            Replay(m_Graphs[2], stream, stream, tail);
        }
        for (int i = 0; i < 16; i++) {
            m_PrevView[i] = view[i];
            m_PrevProj[i] = proj[i];
        }
        ++m_Frame_;
    }

    const vpt_graph* Graph(int chain) const { return m_Graphs[chain]; }

private:
    template <typename F>
    static void Replay(vpt_graph* g, void* after, void* join, F&& chain)
    {
        void* cap = nullptr;
        vpt_check(vpt_graph_begin(g, after, &cap));
        try {
            chain(cap);
        } catch (...) {
            vpt_graph_abort(g);
            throw;
        }
        vpt_check(vpt_graph_launch(g, join));
    }

public:
    const vpt_gbuffer& GBuffer() const { return m_GBuf[(m_Frame_ + 1) % 2]; }
    const float* DenoisedRGB() const { return m_Denoised; }
    // the composite: 3 floats per pixel (vpt_composite)
    const float* Color() const { return m_Color; }
    const float* RawRGB() const { return m_RGB; }
    const uint8_t* ShadowMask() const { return m_Mask; }
    // the denoised sun shadow (unorm8 per pixel)
    const uint8_t* DenoisedShadow() const { return m_ShadowOut; }
    // reflections at (width / 4) x (height / 4) (full size below 4 pixels):
    // o_HitDistance (binary16), the hit record and the RGBA binary16 colour
    // of vpt_trace_reflect_shaded
    const uint16_t* ReflectionHitDistance() const { return m_ReflDist; }
    const uint32_t* ReflectionHits() const { return m_ReflHit; }
    const uint16_t* ReflectionColor() const { return m_ReflColor; }
    int ReflectionWidth() const { return m_RW; }
    int ReflectionHeight() const { return m_RH; }
    int FrameIndex() const { return m_Frame_; }

private:
    template <typename T>
    static void Alloc(T** p, size_t bytes)
    {
        if (hipMalloc(reinterpret_cast<void**>(p), bytes) != hipSuccess) throw std::runtime_error("hipMalloc failed");
    }

    // proj * view, column-major, summed in k order (u_PrevProjection * u_PrevView).
    static void ViewProj(const float* v, const float* p, float* out)
    {
        for (int c = 0; c < 4; c++)
            for (int r = 0; r < 4; r++) {
                float acc = 0.0f;
                for (int k = 0; k < 4; k++) acc = acc + p[k * 4 + r] * v[c * 4 + k];
                out[c * 4 + r] = acc;
            }
    }

    vpt_ctx* m_Ctx;
    World& m_World;
    int m_W, m_H;
    float m_SunTick;
    bool m_Persistent;
    int m_Frame_ = 0;
    vpt_gbuffer m_GBuf[2] = {};
    uint8_t* m_Mask = nullptr;
    uint16_t* m_Transversal = nullptr;
    uint8_t* m_ShadowOut = nullptr;
    uint16_t* m_ReflDist = nullptr;
    uint32_t* m_ReflHit = nullptr;
    uint16_t* m_ReflColor = nullptr;
    void* m_ReflRecords = nullptr;   // the deferred reflection pass's per-sample records
    float* m_Color = nullptr;
    int m_RDiv = 4, m_RW = 0, m_RH = 0;
    vpt_shadow_denoise* m_ShadowDn = nullptr;
    hipStream_t m_Side = nullptr;   // shadow, shadow denoise, reflection rays (default priority)
    hipEvent_t m_Fork = nullptr, m_Join = nullptr;
    float* m_RGB = nullptr;
    float* m_Denoised = nullptr;
    vpt_svgf_texel* m_Texels = nullptr;
    vpt_svgf* m_Svgf = nullptr;
    vpt_frame* m_Frame = nullptr;
    vpt_graph* m_Graphs[3] = {nullptr, nullptr, nullptr};   // side chain, main chain, tail
    float m_PrevView[16] = {}, m_PrevProj[16] = {};
};

}  // namespace VoxelRT
