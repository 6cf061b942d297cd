// Serving benchmark: 16 resident worlds (128^3 up to 512x256x512 terrain),
// a seeded stream of ray-query requests with log-uniform sizes in [min, max]
// rays and kinds primary : shadow : GI = 1 : 1 : 2, served by vpt_server in
// plain mode (one launch per request, in order) and in batched mode
// (continuous batching through the persistent kernel).
//
// Per mode it measures
//   - saturated throughput: every request arrives at once; rays / makespan;
//   - latency under load: Poisson arrivals at a fixed offered rate (rays/s),
//     latency = completion - scheduled arrival (so a late generator cannot
//     hide queueing), reported as p50 / p99 with the delivered rays/s.
// Rays are in device memory before the timed region (ray pools per world
// and kind, built from the worlds' own primary hits); results stay in HBM.
// Prints one JSON object.
//
//   vpt_serve_bench [--requests N] [--min-rays A] [--max-rays B] [--seed S]
//                   [--loads f1,f2,..] [--worlds K]
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "vpt/vpt.h"

namespace {

void check(int rc, const char* what)
{
    if (rc != VPT_OK) {
        fprintf(stderr, "%s failed: %d %s\n", what, rc, vpt_last_error());
        exit(1);
    }
}

void hcheck(hipError_t e, const char* what)
{
    if (e != hipSuccess) {
        fprintf(stderr, "%s failed: %s\n", what, hipGetErrorString(e));
        exit(1);
    }
}

struct Dims {
    int nx, ny, nz;
};

// 16 worlds spanning 128^3 .. 512x256x512.
const Dims kWorlds[16] = {{128, 128, 128}, {160, 128, 160}, {192, 128, 192}, {256, 64, 256},
                          {256, 128, 256}, {320, 128, 320}, {384, 128, 384}, {384, 64, 384},
                          {256, 256, 256}, {448, 128, 448}, {512, 128, 512}, {512, 64, 512},
                          {384, 256, 384}, {448, 256, 448}, {512, 192, 512}, {512, 256, 512}};

struct Pool {
    vpt_ray* d[3] = {nullptr, nullptr, nullptr};   // primary, shadow, GI rays
};

struct Req {
    int world, kind;
    uint32_t n, offset;
    int64_t arrival;   // ns after the start of the run
};

float face_n(int face, int axis)
{
    static const float N[6][3] = {{0, 0, 1}, {0, 0, -1}, {0, 1, 0}, {0, -1, 0}, {-1, 0, 0}, {1, 0, 0}};
    return N[face][axis];
}

// Ray pools of one world: camera rays over the world, then shadow and
// cosine-weighted GI rays from their hit points (or random points above the
// terrain for the misses).
void build_pool(vpt_world* w, const Dims& d, uint32_t n, uint32_t seed, const float sun[3], Pool& pool)
{
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    std::vector<vpt_ray> prim(n), shad(n), gi(n);
    const double cx = d.nx * 0.5 + 0.5, cz = d.nz * 0.5 + 0.5, cy = d.ny * 0.8 + 0.5;
    const int W = 1920, H = (int)((n + W - 1) / W);
    const double fov = 60.0 * M_PI / 180.0, aspect = (double)W / H;
    for (uint32_t i = 0; i < n; i++) {
        const int px = (int)(i % W), py = (int)(i / W);
        // camera looking along +x+z, pitched 20 degrees down, from above the centre
        const double yaw = 45.0 * M_PI / 180.0;
        const double pitch = -20.0 * M_PI / 180.0;
        const double fx = cos(pitch) * cos(yaw), fy = sin(pitch), fz = cos(pitch) * sin(yaw);
        // right = normalize(cross(front, up)), up' = cross(right, front)
        double rx = -fz, ry = 0.0, rz = fx;
        const double rl = sqrt(rx * rx + rz * rz);
        rx /= rl;
        rz /= rl;
        const double ux = ry * fz - rz * fy, uy = rz * fx - rx * fz, uz = rx * fy - ry * fx;
        const double sx = (2.0 * (px + 0.5) / W - 1.0) * tan(fov / 2) * aspect;
        const double sy = (2.0 * (py + 0.5) / H - 1.0) * tan(fov / 2);
        double dx = fx + sx * rx + sy * ux, dy = fy + sx * ry + sy * uy, dz = fz + sx * rz + sy * uz;
        const double l = sqrt(dx * dx + dy * dy + dz * dz);
        vpt_ray r;
        r.origin[0] = (float)(cx - 0.3 * d.nx);
        r.origin[1] = (float)cy;
        r.origin[2] = (float)(cz - 0.3 * d.nz);
        r.dir[0] = (float)(dx / l);
        r.dir[1] = (float)(dy / l);
        r.dir[2] = (float)(dz / l);
        prim[i] = r;
    }
    hcheck(hipMalloc(&pool.d[0], sizeof(vpt_ray) * n), "hipMalloc");
    hcheck(hipMalloc(&pool.d[1], sizeof(vpt_ray) * n), "hipMalloc");
    hcheck(hipMalloc(&pool.d[2], sizeof(vpt_ray) * n), "hipMalloc");
    hcheck(hipMemcpy(pool.d[0], prim.data(), sizeof(vpt_ray) * n, hipMemcpyHostToDevice), "hipMemcpy");
    vpt_hit* d_hits = nullptr;
    hcheck(hipMalloc(&d_hits, sizeof(vpt_hit) * n), "hipMalloc");
    check(vpt_trace_rays(w, pool.d[0], (int)n, 350, d_hits, nullptr), "vpt_trace_rays");
    std::vector<vpt_hit> hits(n);
    hcheck(hipMemcpy(hits.data(), d_hits, sizeof(vpt_hit) * n, hipMemcpyDeviceToHost), "hipMemcpy");
    hcheck(hipFree(d_hits), "hipFree");
    for (uint32_t i = 0; i < n; i++) {
        const vpt_hit& h = hits[i];
        float o[3];
        float nrm[3] = {0.0f, 1.0f, 0.0f};
        if (h.t > 0.0f && h.block > 0) {
            for (int a = 0; a < 3; a++) {
                nrm[a] = face_n(h.face, a);
                o[a] = prim[i].origin[a] + prim[i].dir[a] * h.t + nrm[a] * 0.06f;
            }
        } else {
            o[0] = (float)(U(rng) * d.nx);
            o[1] = (float)(d.ny * (0.75 + 0.2 * U(rng)));
            o[2] = (float)(U(rng) * d.nz);
        }
        vpt_ray s;
        memcpy(s.origin, o, sizeof(o));
        memcpy(s.dir, sun, sizeof(float) * 3);
        shad[i] = s;
        // cosine-weighted direction around nrm
        const double r1 = U(rng), r2 = U(rng);
        const double a = 2.0 * M_PI * r1, rr = sqrt(r2);
        double t0[3];
        if (fabs(nrm[1]) < 0.9) {
            t0[0] = nrm[2];
            t0[1] = 0.0;
            t0[2] = -nrm[0];
        } else {
            t0[0] = 0.0;
            t0[1] = -nrm[2];
            t0[2] = nrm[1];
        }
        double tl = sqrt(t0[0] * t0[0] + t0[1] * t0[1] + t0[2] * t0[2]);
        if (tl == 0.0) {
            t0[0] = 1.0;
            t0[1] = t0[2] = 0.0;
            tl = 1.0;
        }
        for (double& c : t0) c /= tl;
        const double b[3] = {nrm[1] * t0[2] - nrm[2] * t0[1], nrm[2] * t0[0] - nrm[0] * t0[2],
                             nrm[0] * t0[1] - nrm[1] * t0[0]};
        double g[3];
        for (int k = 0; k < 3; k++) g[k] = rr * cos(a) * t0[k] + rr * sin(a) * b[k] + sqrt(1.0 - r2) * nrm[k];
        const double gl = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
        vpt_ray q;
        memcpy(q.origin, o, sizeof(o));
        for (int k = 0; k < 3; k++) q.dir[k] = (float)(g[k] / gl);
        gi[i] = q;
    }
    hcheck(hipMemcpy(pool.d[1], shad.data(), sizeof(vpt_ray) * n, hipMemcpyHostToDevice), "hipMemcpy");
    hcheck(hipMemcpy(pool.d[2], gi.data(), sizeof(vpt_ray) * n, hipMemcpyHostToDevice), "hipMemcpy");
}

double pct(std::vector<double> v, double p)
{
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double idx = p * (double)(v.size() - 1);
    const size_t lo = (size_t)idx, hi = std::min(lo + 1, v.size() - 1);
    return v[lo] + (v[hi] - v[lo]) * (idx - (double)lo);
}

struct RunResult {
    double rays_per_s = 0, makespan_ms = 0, p50_ms = 0, p99_ms = 0, mean_ms = 0, offered = 0;
    uint64_t launches = 0, arena_peak = 0;
};

// Serves `reqs` (arrival offsets relative to the start); a collector thread
// waits for the tickets in order and releases them.
RunResult serve(vpt_server* srv, const std::vector<int>& slots, const std::vector<Pool>& pools,
                const std::vector<Req>& reqs, const float sun[3])
{
    const size_t n = reqs.size();
    std::vector<uint64_t> tickets(n, 0);
    std::vector<int> submitted(n, 0);
    std::vector<double> lat(n, 0.0);
    std::vector<int64_t> done(n, 0);
    vpt_server_stats s0;
    check(vpt_server_stats_get(srv, &s0), "stats");
    const int64_t t0 = vpt_now_ns() + 2000000;   // start 2 ms from now
    std::atomic<size_t> n_sub{0};
    std::thread collector([&] {
        for (size_t i = 0; i < n; i++) {
            while (n_sub.load(std::memory_order_acquire) <= i) std::this_thread::yield();
            vpt_request_result res;
            check(vpt_server_wait(srv, tickets[i], -1, &res), "vpt_server_wait");
            check(res.status, "request");
            lat[i] = (double)(res.t_done_ns - res.t_arrival_ns) * 1e-6;
            done[i] = res.t_done_ns;
            check(vpt_server_release(srv, tickets[i]), "vpt_server_release");
        }
    });
    for (size_t i = 0; i < n; i++) {
        const Req& r = reqs[i];
        const int64_t at = t0 + r.arrival;
        // sleep until shortly before the arrival, then spin
        while (true) {
            const int64_t now = vpt_now_ns();
            if (now >= at) break;
            if (at - now > 200000) std::this_thread::sleep_for(std::chrono::nanoseconds(at - now - 100000));
        }
        vpt_request q;
        memset(&q, 0, sizeof(q));
        q.world = slots[r.world];
        q.kind = r.kind;
        q.n = r.n;
        q.max_iter = r.kind == VPT_REQ_GI ? 48 : 350;
        q.d_rays = pools[r.world].d[r.kind] + r.offset;
        memcpy(q.sun_dir, sun, sizeof(float) * 3);
        q.shadow_iter = 128;
        q.arrival_ns = at;
        check(vpt_server_submit(srv, &q, &tickets[i]), "vpt_server_submit");
        n_sub.store(i + 1, std::memory_order_release);
    }
    collector.join();
    vpt_server_stats s1;
    check(vpt_server_stats_get(srv, &s1), "stats");
    RunResult out;
    uint64_t rays = 0;
    for (const Req& r : reqs) rays += r.n;
    const int64_t last = *std::max_element(done.begin(), done.end());
    out.makespan_ms = (double)(last - t0) * 1e-6;
    out.rays_per_s = (double)rays / ((double)(last - t0) * 1e-9);
    out.p50_ms = pct(lat, 0.50);
    out.p99_ms = pct(lat, 0.99);
    double m = 0;
    for (double v : lat) m += v;
    out.mean_ms = m / (double)n;
    out.launches = s1.launches - s0.launches;
    out.arena_peak = s1.arena_peak;
    return out;
}

std::vector<Req> make_trace(int count, uint32_t min_rays, uint32_t max_rays, uint32_t pool_n, uint64_t seed,
                            int nworlds)
{
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    std::vector<Req> v(count);
    const double lo = log((double)min_rays), hi = log((double)max_rays);
    for (Req& r : v) {
        r.world = (int)(U(rng) * nworlds) % nworlds;
        const int k = (int)(U(rng) * 4) % 4;   // primary : shadow : GI = 1 : 1 : 2
        r.kind = k == 0 ? VPT_REQ_PRIMARY : k == 1 ? VPT_REQ_SHADOW : VPT_REQ_GI;
        r.n = (uint32_t)std::min<double>(std::max<double>(exp(lo + (hi - lo) * U(rng)), min_rays), max_rays);
        r.offset = (uint32_t)(U(rng) * (double)(pool_n - r.n));
        r.arrival = 0;
    }
    return v;
}

// Poisson arrivals for an offered load of `rays_per_s`.
void set_arrivals(std::vector<Req>& v, double rays_per_s, uint64_t seed)
{
    double mean = 0;
    for (const Req& r : v) mean += r.n;
    mean /= (double)v.size();
    const double rate = rays_per_s / mean;   // requests per second
    std::mt19937_64 rng(seed);
    std::exponential_distribution<double> E(rate);
    double t = 0;
    for (Req& r : v) {
        r.arrival = (int64_t)(t * 1e9);
        t += E(rng);
    }
}

}  // namespace

int main(int argc, char** argv)
{
    int count = 6000, nworlds = 16;
    uint32_t min_rays = 1024, max_rays = 2u << 20;
    uint64_t seed = 0x5EED0005;
    std::vector<double> loads = {0.5};
    for (int i = 1; i < argc; i++) {
        if (!strcmp(argv[i], "--requests") && i + 1 < argc) count = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--min-rays") && i + 1 < argc) min_rays = (uint32_t)atol(argv[++i]);
        else if (!strcmp(argv[i], "--max-rays") && i + 1 < argc) max_rays = (uint32_t)atol(argv[++i]);
        else if (!strcmp(argv[i], "--seed") && i + 1 < argc) seed = strtoull(argv[++i], nullptr, 0);
        else if (!strcmp(argv[i], "--worlds") && i + 1 < argc) nworlds = std::max(1, std::min(16, atoi(argv[++i])));
        else if (!strcmp(argv[i], "--loads") && i + 1 < argc) {
            loads.clear();
            for (char* p = strtok(argv[++i], ","); p; p = strtok(nullptr, ",")) loads.push_back(atof(p));
        }
    }
    const float sun[3] = {0.3f / 1.0478f, 0.85f / 1.0478f, 0.42f / 1.0478f};
    vpt_ctx* ctx = nullptr;
    check(vpt_ctx_create(0, &ctx), "vpt_ctx_create");

    // resident worlds + their ray pools
    const uint32_t pool_n = max_rays;
    std::vector<vpt_world*> worlds(nworlds);
    std::vector<Pool> pools(nworlds);
    uint64_t world_bytes = 0;
    for (int i = 0; i < nworlds; i++) {
        const Dims& d = kWorlds[i];
        std::vector<uint8_t> grid((size_t)d.nx * d.ny * d.nz);
        check(vpt_generate_terrain(d.nx, d.ny, d.nz, 0x5EED1000u + i, 0x5EED2000u + i, d.ny * 20 / 128,
                                   d.ny * 90 / 128, 10, grid.data()),
              "vpt_generate_terrain");
        check(vpt_world_create(ctx, d.nx, d.ny, d.nz, &worlds[i]), "vpt_world_create");
        check(vpt_world_upload(worlds[i], grid.data(), nullptr), "vpt_world_upload");
        check(vpt_df_build(worlds[i], nullptr), "vpt_df_build");
        size_t b = 0;
        check(vpt_world_bytes(worlds[i], &b), "vpt_world_bytes");
        world_bytes += b;
        build_pool(worlds[i], d, pool_n, 0x5EED3000u + i, sun, pools[i]);
        fprintf(stderr, "world %d %dx%dx%d ready\n", i, d.nx, d.ny, d.nz);
    }
    hcheck(hipDeviceSynchronize(), "sync");

    std::vector<Req> trace = make_trace(count, min_rays, max_rays, pool_n, seed, nworlds);
    uint64_t total_rays = 0;
    int kinds[3] = {0, 0, 0};
    for (const Req& r : trace) {
        total_rays += r.n;
        kinds[r.kind]++;
    }

    // plain: one launch per request; batched: continuous batching through
    // the persistent kernel (lane refill) or one lane per ray (direct)
    const int M = 3;
    const char* names[M] = {"plain", "batched", "batched_direct"};
    RunResult sat[M];
    std::vector<std::vector<RunResult>> lat(M);
    std::vector<vpt_server*> servers(M, nullptr);
    std::vector<std::vector<int>> slots(M);
    for (int m = 0; m < M; m++) {
        vpt_server_cfg cfg;
        memset(&cfg, 0, sizeof(cfg));
        cfg.mode = m == 0 ? VPT_SERVE_PLAIN : VPT_SERVE_BATCHED;
        cfg.max_worlds = 16;
        cfg.arena_bytes = 4ull << 30;
        cfg.world_budget = 64ull << 30;
        cfg.max_launch_rays = 4u << 20;
        cfg.max_inflight = 2;
        check(vpt_server_create(ctx, &cfg, &servers[m]), "vpt_server_create");
        vpt_launch_cfg lc = {256, 16, 256, 32, m == 2 ? 2 : 0};
        check(vpt_server_set_launch_cfg(servers[m], &lc), "vpt_server_set_launch_cfg");
        for (int i = 0; i < nworlds; i++) {
            int s = -1;
            check(vpt_server_add_world(servers[m], worlds[i], &s), "vpt_server_add_world");
            slots[m].push_back(s);
        }
        // warm-up: a few requests
        std::vector<Req> warm(trace.begin(), trace.begin() + std::min<size_t>(64, trace.size()));
        for (Req& r : warm) r.arrival = 0;
        (void)serve(servers[m], slots[m], pools, warm, sun);
        // saturated: everything arrives at once
        std::vector<Req> all = trace;
        for (Req& r : all) r.arrival = 0;
        sat[m] = serve(servers[m], slots[m], pools, all, sun);
        fprintf(stderr, "%s saturated: %.3f Grays/s, makespan %.1f ms, %llu launches\n", names[m],
                sat[m].rays_per_s * 1e-9, sat[m].makespan_ms, (unsigned long long)sat[m].launches);
    }
    const int best = sat[2].rays_per_s > sat[1].rays_per_s ? 2 : 1;
    // the same Poisson traces for every mode, at offered loads relative to
    // the plain path's saturated throughput
    for (double f : loads) {
        std::vector<Req> v = trace;
        set_arrivals(v, f * sat[0].rays_per_s, seed ^ 0x9E3779B97F4A7C15ull);
        for (int m = 0; m < M; m++) {
            RunResult r = serve(servers[m], slots[m], pools, v, sun);
            r.offered = f * sat[0].rays_per_s;
            lat[m].push_back(r);
            fprintf(stderr, "%s load %.2fx plain: %.3f Grays/s p50 %.3f ms p99 %.3f ms\n", names[m], f,
                    r.rays_per_s * 1e-9, r.p50_ms, r.p99_ms);
        }
    }
    // the faster batched path at the same fraction of its own capacity
    std::vector<RunResult> own;
    for (double f : loads) {
        std::vector<Req> v = trace;
        set_arrivals(v, f * sat[best].rays_per_s, seed ^ 0x7F4A7C159E3779B9ull);
        RunResult r = serve(servers[best], slots[best], pools, v, sun);
        r.offered = f * sat[best].rays_per_s;
        own.push_back(r);
        fprintf(stderr, "%s load %.2fx own: %.3f Grays/s p50 %.3f ms p99 %.3f ms\n", names[best], f,
                r.rays_per_s * 1e-9, r.p50_ms, r.p99_ms);
    }

    auto res_json = [](const RunResult& r) {
        char b[512];
        snprintf(b, sizeof(b),
                 "{\"offered_grays_s\": %.4f, \"grays_s\": %.4f, \"makespan_ms\": %.3f, \"p50_ms\": %.4f, "
                 "\"p99_ms\": %.4f, \"mean_ms\": %.4f, \"launches\": %llu}",
                 r.offered * 1e-9, r.rays_per_s * 1e-9, r.makespan_ms, r.p50_ms, r.p99_ms, r.mean_ms,
                 (unsigned long long)r.launches);
        return std::string(b);
    };
    printf("{\"requests\": %d, \"rays\": %llu, \"kinds\": [%d, %d, %d], \"min_rays\": %u, \"max_rays\": %u, "
           "\"worlds\": %d, \"world_bytes\": %llu, \"best_batched\": \"%s\", \"saturated\": {",
           count, (unsigned long long)total_rays, kinds[0], kinds[1], kinds[2], min_rays, max_rays, nworlds,
           (unsigned long long)world_bytes, names[best]);
    for (int m = 0; m < M; m++) printf("%s\"%s\": %s", m ? ", " : "", names[m], res_json(sat[m]).c_str());
    printf("}, \"ratio_saturated\": %.3f, \"loads\": [", sat[best].rays_per_s / sat[0].rays_per_s);
    for (size_t i = 0; i < loads.size(); i++) {
        printf("%s{\"load_vs_plain\": %.2f", i ? ", " : "", loads[i]);
        for (int m = 0; m < M; m++) printf(", \"%s\": %s", names[m], res_json(lat[m][i]).c_str());
        printf(", \"best_batched_own_load\": %s}", res_json(own[i]).c_str());
    }
    printf("]}\n");
    fflush(stdout);
    for (vpt_server* s : servers) vpt_server_destroy(s);
    for (int i = 0; i < nworlds; i++) {
        for (vpt_ray* p : pools[i].d) (void)hipFree(p);
        vpt_world_destroy(worlds[i]);
    }
    vpt_ctx_destroy(ctx);
    return 0;
}
