// A C++ host driving the hot path through the VoxelRT::World /
// VoxelRT::Pipeline façades (include/vpt/World.hpp, Pipeline.hpp): terrain
// world, GPU distance field, N frames of primary + shadow + GI + SVGF +
// shadow denoise + shaded reflections + composite along the SVGF workload's
// camera path.  Prints one JSON line; with --dump FILE writes the last
// frame's denoised RGB (float32, H x W x 3), then its denoised shadow (uint8,
// H x W), its reflection hit distances (binary16 bits, H/4 x W/4), its
// reflection colour (RGBA binary16 bits, H/4 x W/4 x 4) and its composite
// (float32, H x W x 3).
//
//   vpt_frame_demo [W H FRAMES [--plain | --graph] [--warmup N] [--dump FILE]]
//
// --plain: one launch per pass; --graph: one launch per pass, every frame
// replayed as a HIP graph; default: the persistent engine.  --warmup N:
// N frames of the same camera path before the FRAMES timed ones.
#include <hip/hip_runtime_api.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "vpt/Pipeline.hpp"
#include "vpt/World.hpp"

int main(int argc, char** argv)
{
    int W = argc > 1 ? atoi(argv[1]) : 640;
    int H = argc > 2 ? atoi(argv[2]) : 360;
    int frames = argc > 3 ? atoi(argv[3]) : 8;
    bool persistent = true, graph = false;
    const char* dump = nullptr;
    int warmup = 0;
    for (int i = 4; i < argc; i++) {
        if (!strcmp(argv[i], "--warmup") && i + 1 < argc) warmup = atoi(argv[++i]);
        if (!strcmp(argv[i], "--plain")) persistent = false;
        if (!strcmp(argv[i], "--graph")) {
            persistent = false;
            graph = true;
        }
        if (!strcmp(argv[i], "--dump") && i + 1 < argc) dump = argv[++i];
    }
    try {
        VoxelRT::Context ctx(0);
        VoxelRT::World world(ctx, 384, 128, 384);
        VoxelRT::vpt_check(vpt_generate_terrain(384, 128, 384, 0x5EED0001u, 0x5EED0002u, 20, 90, 10,
                                                reinterpret_cast<uint8_t*>(world.m_WorldData.data())));
        world.Buffer();
        world.InitializeDistanceGenerator();
        world.GenerateDistanceField();
        VoxelRT::Pipeline pipe(ctx, world, W, H, 50.0f, persistent, graph);
        hipEvent_t e0, e1;
        (void)hipEventCreate(&e0);
        (void)hipEventCreate(&e1);
        for (int f = 0; f < warmup; f++) {
            const float pos[3] = {192.5f + 0.1f * (float)f, 100.5f, 192.5f};
            pipe.RenderFrame(pos, 45.0f + 0.2f * (float)f, -20.0f);
        }
        (void)hipDeviceSynchronize();
        (void)hipEventRecord(e0, nullptr);
        for (int f = warmup; f < warmup + frames; f++) {
            const float pos[3] = {192.5f + 0.1f * (float)f, 100.5f, 192.5f};
            pipe.RenderFrame(pos, 45.0f + 0.2f * (float)f, -20.0f);
        }
        (void)hipEventRecord(e1, nullptr);
        (void)hipEventSynchronize(e1);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        std::vector<float> rgb((size_t)W * H * 3);
        (void)hipMemcpy(rgb.data(), pipe.DenoisedRGB(), rgb.size() * sizeof(float), hipMemcpyDeviceToHost);
        double mean = 0.0;
        for (float v : rgb) mean += v;
        mean /= (double)rgb.size();
        // block picking from the camera toward the view centre
        const float p0[3] = {192.5f, 100.5f, 192.5f};
        const float d0[3] = {0.664463f, -0.342020f, 0.664463f};
        auto pick = world.RaycastDetect(p0, d0);
        if (dump) {
            std::vector<uint8_t> shadow((size_t)W * H);
            std::vector<uint16_t> refl((size_t)pipe.ReflectionWidth() * pipe.ReflectionHeight());
            std::vector<uint16_t> refl_color(refl.size() * 4);
            std::vector<float> color((size_t)W * H * 3);
            (void)hipMemcpy(shadow.data(), pipe.DenoisedShadow(), shadow.size(), hipMemcpyDeviceToHost);
            (void)hipMemcpy(refl.data(), pipe.ReflectionHitDistance(), refl.size() * sizeof(uint16_t),
                            hipMemcpyDeviceToHost);
            (void)hipMemcpy(refl_color.data(), pipe.ReflectionColor(), refl_color.size() * sizeof(uint16_t),
                            hipMemcpyDeviceToHost);
            (void)hipMemcpy(color.data(), pipe.Color(), color.size() * sizeof(float), hipMemcpyDeviceToHost);
            FILE* fp = fopen(dump, "wb");
            if (!fp) throw std::runtime_error("cannot open dump file");
            fwrite(rgb.data(), sizeof(float), rgb.size(), fp);
            fwrite(shadow.data(), 1, shadow.size(), fp);
            fwrite(refl.data(), sizeof(uint16_t), refl.size(), fp);
            fwrite(refl_color.data(), sizeof(uint16_t), refl_color.size(), fp);
            fwrite(color.data(), sizeof(float), color.size(), fp);
            fclose(fp);
        }
        printf("{\"frames\": %d, \"width\": %d, \"height\": %d, \"path\": \"%s\", \"ms_per_frame\": %.4f, "
               "\"denoised_mean\": %.6f, \"pick\": [%d, %d, %d, %d]}\n",
               frames, W, H, persistent ? "persistent" : (graph ? "graph" : "plain"), ms / frames, mean, pick[0],
               pick[1], pick[2],
               pick[3]);
    } catch (const std::exception& e) {
        fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
