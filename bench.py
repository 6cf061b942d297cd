"""Benchmark of the hot path on one MI355X.

Workload (one step = one frame of rays, inputs resident in HBM):
  world 384x128x384 uint8 terrain (seeded value noise, heights 20-90,
  grass/dirt/stone layering, 1% of solid voxels given random ids 1-100),
  distance field built on the GPU; at 1920x1080, 1 spp: one primary ray per
  pixel, one sun-shadow ray per lit pixel, one cosine-weighted diffuse GI
  ray per hit pixel plus a sun-shadow ray at the GI hit, run as the library
  runs a frame (vpt_trace_frame: the shadow pass on a second stream beside
  the GI pass, from the end of its bounce kernel; the passes one after
  another are reported too, with each pass's time).  Grays/s counts every ray whose traversal loop is entered.  SVGF ms/frame (pre-spatial,
  temporal, variance, 5 a-trous, decode at 1920x1080) and DF-build ms are
  reported beside it, with the wall time of an incremental DF update after
  a one-block edit, the sun-shadow denoiser's ms/frame over the same
  sequence, and the reflection pass (reflected rays + sun shadow at their
  hits, 64 / 150 steps) off the frame's G-buffer with the reference's
  settings (1/4 resolution, 2 rough samples) and as full-resolution
  mirrors; and the whole pipeline frame (all of the above that a frame
  runs, as Pipeline.render runs them).

Prints one JSON line (see the "roofline" and "cpu_baseline" objects).
"""
from __future__ import annotations

import argparse
import gc
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
# the kernels each frame pass launches (the roofline covers the whole pass)
PASS_KERNELS = {"primary": "primary_kernel", "shadow": "shadow_kernel",
                "gi": "gi pass: gi_bounce_kernel + gi_cont_kernel + gi_shadow_kernel"}
NX, NY, NZ = 384, 128, 384
CAM_POS = (192.5, 100.5, 192.5)
CAM_YAW, CAM_PITCH = 45.0, -20.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--svgf-frames", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving benchmark")
    ap.add_argument("--serve-requests", type=int, default=3000)
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-denoiser section")
    ap.add_argument("--no-final-image", action="store_true", help="skip the final-image PSNR section")
    ap.add_argument("--unet-pairs", type=int, default=2048)
    ap.add_argument("--unet-steps", type=int, default=2000)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="also run a plain device copy of MIB MiB (calibrates rocprof FETCH_SIZE/WRITE_SIZE)")
    return ap.parse_args()


def copy_probe(mib: int) -> None:
    import torch
    n = mib * (1 << 20)
    a = torch.ones(n, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "bench.py measures one GPU only; run with --gpus 1 "
                          "(the path would shard by pixel tiles / requests with no collective)"}))
        sys.exit(2)
    import torch
    from voxelpathtracer_amd import (SVGF, Context, ShadowDenoiser, World, generate_terrain, gi_params, make_camera,
                                     viewproj)
    from voxelpathtracer_amd.world import DeviceGBuffer

    W, H = args.width, args.height
    ctx = Context(0)
    grid = generate_terrain(NX, NY, NZ)
    world = World(ctx, NX, NY, NZ)
    world.m_WorldData[...] = grid
    world.Buffer()
    stream = torch.cuda.current_stream()

    def ev():
        return torch.cuda.Event(enable_timing=True)

    # ---- DF build --------------------------------------------------------
    for _ in range(3):
        world.GenerateDistanceField()
    torch.cuda.synchronize()
    df_reps = 20
    e0, e1 = ev(), ev()
    e0.record(stream)
    for _ in range(df_reps):
        world.GenerateDistanceField()
    e1.record(stream)
    torch.cuda.synchronize()
    df_ms = e0.elapsed_time(e1) / df_reps
    df_update = df_update_timing(world, grid, torch)

    # ---- the whole frame as the pipelines run it (Pipeline.render: primary,
    # shadow + transversal beside GI, SVGF, shadow denoise, reflections at the
    # reference's settings) along the SVGF camera path, 16 frames after 4.
    # Timed before the sections below create their streams: with more HIP
    # streams than the GPU's hardware queues (4 per process), streams share
    # queues, and a side-chain graph whose start is ordered through a stream
    # that shares the main chain's queue waits for the main chain's replay.
    from voxelpathtracer_amd import Pipeline
    pipe = Pipeline(ctx, world, W, H)
    for f in range(4):
        pipe.render((CAM_POS[0] + 0.1 * f, CAM_POS[1], CAM_POS[2]), CAM_YAW + 0.2 * f, CAM_PITCH)
    torch.cuda.synchronize()
    # the host loop allocates Python objects every frame; a cyclic garbage
    # collection over this process's large heap inside the timed frames
    # would stall the enqueueing (as timeit, the timed loops run without it)
    gc.collect()
    gc.disable()
    p0, p1 = ev(), ev()
    p0.record(stream)
    for f in range(4, 20):
        pipe.render((CAM_POS[0] + 0.1 * f, CAM_POS[1], CAM_POS[2]), CAM_YAW + 0.2 * f, CAM_PITCH)
    p1.record(stream)
    torch.cuda.synchronize()
    gc.enable()
    pipeline_ms = p0.elapsed_time(p1) / 16
    del pipe
    # the same frames replayed as a HIP graph each (Pipeline(use_graph=True))
    pipe = Pipeline(ctx, world, W, H, use_graph=True)
    for f in range(4):
        pipe.render((CAM_POS[0] + 0.1 * f, CAM_POS[1], CAM_POS[2]), CAM_YAW + 0.2 * f, CAM_PITCH)
    torch.cuda.synchronize()
    gc.collect()
    gc.disable()
    p0, p1 = ev(), ev()
    p0.record(stream)
    t_host = time.perf_counter()
    for f in range(4, 20):
        pipe.render((CAM_POS[0] + 0.1 * f, CAM_POS[1], CAM_POS[2]), CAM_YAW + 0.2 * f, CAM_PITCH)
    pipeline_graph_host_ms = (time.perf_counter() - t_host) * 1e3 / 16
    p1.record(stream)
    torch.cuda.synchronize()
    gc.enable()
    pipeline_graph_ms = p0.elapsed_time(p1) / 16
    pipeline_graph_stats = {"side_chain": pipe.graphs[0].stats(), "main_chain": pipe.graphs[1].stats(),
                            "tail": pipe.graphs[2].stats()}
    del pipe
    pipeline_cpp = cpp_pipeline(W, H)

    # ---- rays --------------------------------------------------------------
    cam, view, proj = make_camera(CAM_POS, CAM_YAW, CAM_PITCH, W, H)
    prm = gi_params(50.0, 0)
    sun = list(prm.sun_dir)
    gb = DeviceGBuffer(W, H)
    mask = torch.empty((H, W), dtype=torch.uint8, device="cuda")
    rgb = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
    tex = torch.empty((H, W, 8), dtype=torch.int16, device="cuda")

    def step():
        # the frame as the library runs it: the shadow pass beside GI
        world.trace_frame(cam, gb, prm, mask, rgb, tex, 350, 350)

    def step_seq(events=None):
        if events is not None:
            events[0].record(stream)
        world.trace_primary(cam, gb, 350)
        if events is not None:
            events[1].record(stream)
        world.trace_shadow(cam, gb, sun, mask, 350)
        if events is not None:
            events[2].record(stream)
        world.trace_gi(cam, gb, prm, rgb, tex)
        if events is not None:
            events[3].record(stream)

    # work counts (untimed): DF samples and traced rays per pass
    counts = {}
    counter = torch.zeros(2, dtype=torch.int64, device="cuda")
    world.set_sample_counter(counter)
    for name, fn in (("primary", lambda: world.trace_primary(cam, gb, 350)),
                     ("shadow", lambda: world.trace_shadow(cam, gb, sun, mask, 350)),
                     ("gi", lambda: world.trace_gi(cam, gb, prm, rgb, tex))):
        counter.zero_()
        fn()
        torch.cuda.synchronize()
        c = counter.cpu().numpy()
        counts[name] = dict(samples=int(c[0]), rays=int(c[1]))
    world.set_sample_counter(None)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t_start, t_end = ev(), ev()
    t_start.record(stream)
    for k in range(args.steps):
        step()
    t_end.record(stream)
    torch.cuda.synchronize()
    total_ms = t_start.elapsed_time(t_end)
    ms_per_step = total_ms / args.steps
    # each pass on its own (one stream, in order) for the per-pass times
    for _ in range(args.warmup):
        step_seq()
    evs = [[ev() for _ in range(4)] for _ in range(args.steps)]
    q0, q1 = ev(), ev()
    torch.cuda.synchronize()
    q0.record(stream)
    for k in range(args.steps):
        step_seq(evs[k])
    q1.record(stream)
    torch.cuda.synchronize()
    ms_sequential = q0.elapsed_time(q1) / args.steps
    pass_ms = {
        "primary": float(np.mean([e[0].elapsed_time(e[1]) for e in evs])),
        "shadow": float(np.mean([e[1].elapsed_time(e[2]) for e in evs])),
        "gi": float(np.mean([e[2].elapsed_time(e[3]) for e in evs])),
    }
    rays_per_step = sum(c["rays"] for c in counts.values())
    grays_plain = rays_per_step / (ms_per_step * 1e-3) / 1e9
    ms_plain = ms_per_step
    # the same frame captured each step and replayed as a HIP graph
    # (vpt_graph: capture, in-place update of the instantiated graph, launch)
    from voxelpathtracer_amd import Graph
    graph = Graph(ctx)

    def step_graph():
        with graph.capture() as s:
            world.trace_frame(cam, gb, prm, mask, rgb, tex, 350, 350, stream=s)

    for _ in range(args.warmup):
        step_graph()
    torch.cuda.synchronize()
    g0, g1 = ev(), ev()
    g0.record(stream)
    t_host = time.perf_counter()
    for _ in range(args.steps):
        step_graph()
    host_graph_ms = (time.perf_counter() - t_host) * 1e3 / args.steps
    g1.record(stream)
    torch.cuda.synchronize()
    ms_graph = g0.elapsed_time(g1) / args.steps
    graph_stats = graph.stats()
    graph.close()

    # ---- persistent frame path (same outputs bit for bit): launch shape
    # autotuned first (untimed), then the pass chain replayed as a HIP graph
    from voxelpathtracer_amd import Frame
    fr = Frame(ctx, W, H)
    tuned_ms = fr.autotune(world, cam, prm, gb, mask, rgb, tex)
    counter.zero_()
    world.set_sample_counter(counter)
    fr.trace(world, cam, prm, gb, mask, rgb, tex, use_graph=True)
    torch.cuda.synchronize()
    pc = counter.cpu().numpy()
    world.set_sample_counter(None)
    persist_counts = dict(samples=int(pc[0]), rays=int(pc[1]))
    for _ in range(args.warmup):
        fr.trace(world, cam, prm, gb, mask, rgb, tex, use_graph=True)
    torch.cuda.synchronize()
    p0, p1 = ev(), ev()
    p0.record(stream)
    for _ in range(args.steps):
        fr.trace(world, cam, prm, gb, mask, rgb, tex, use_graph=True)
    p1.record(stream)
    torch.cuda.synchronize()
    ms_persist = p0.elapsed_time(p1) / args.steps
    grays_persist = persist_counts["rays"] / (ms_persist * 1e-3) / 1e9
    cfg = fr.cfg
    persistent = {"ms_per_frame": round(ms_persist, 4), "grays": round(grays_persist, 4), "work": persist_counts,
                  "launch": {"block": cfg.block, "waves_per_cu": cfg.waves_per_cu, "chunk": cfg.chunk,
                             "refill": cfg.refill, "tile_primary": bool(cfg.flags & 1)},
                  "autotune_best_ms": round(tuned_ms, 4)}
    if ms_persist < ms_plain:
        ms_per_step, grays, path = ms_persist, grays_persist, "persistent"
    else:
        grays, path = grays_plain, "plain"

    # algorithmic bytes per launch: 1 B per DF sample + 1 B final voxel per
    # ray + ray I/O (primary out 10 B/px; shadow in 5 + out 1 B/px; GI in
    # 5 B/px + out 12 B RGB + 16 B SVGF texel per px)
    px = W * H
    io = {"primary": 10 * px, "shadow": 6 * px, "gi": (5 + 12 + 16) * px}
    alg = {k: counts[k]["samples"] + counts[k]["rays"] + io[k] for k in counts}
    dom = max(pass_ms, key=lambda k: pass_ms[k])
    achieved = alg[dom] / (pass_ms[dom] * 1e-3) / 1e9

    # ---- SVGF: the 64-frame sequence (camera +0.1 voxel/frame in x and
    # +0.2 deg yaw/frame); G-buffers and 1-spp GI inputs of every frame are
    # rendered first and stay resident, then the denoiser alone is timed.
    nf = args.svgf_frames
    seq = []
    shadow_in = []   # per frame: shadow mask and transversal (vpt_trace_shadow2)
    cur_vps = []     # per frame: its own projection * view
    prev_vm = None
    for f in range(nf):
        pos = (CAM_POS[0] + 0.1 * f, CAM_POS[1], CAM_POS[2])
        c_f, v_f, p_f = make_camera(pos, CAM_YAW + 0.2 * f, CAM_PITCH, W, H)
        g_f = DeviceGBuffer(W, H)
        t_f = torch.empty((H, W, 8), dtype=torch.int16, device="cuda")
        world.trace_primary(c_f, g_f, 350)
        world.trace_gi(c_f, g_f, gi_params(50.0, f), rgb, t_f)
        m_f = torch.empty((H, W), dtype=torch.uint8, device="cuda")
        tv_f = torch.empty((H, W), dtype=torch.int16, device="cuda")
        world.trace_shadow2(c_f, g_f, sun, m_f, tv_f, 350)
        vp_f = viewproj(*(prev_vm if prev_vm else (v_f, p_f)))
        seq.append((c_f, vp_f, g_f, t_f))
        shadow_in.append((m_f, tv_f))
        cur_vps.append(viewproj(v_f, p_f))
        prev_vm = (v_f, p_f)
    torch.cuda.synchronize()
    sv = SVGF(ctx, W, H)
    den = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")

    eval_frames = [f for f in (15, 31, 47, 63) if f < nf]
    svgf_out = {}
    svgf_texels = {}   # eval frames: the final SVGF texels (the composite's diffuse)

    def svgf_sequence(keep=False):
        sv.reset()
        for f, (c_f, vp_f, g_f, t_f) in enumerate(seq):
            g_prev = seq[f - 1][2] if f else g_f
            keep_f = keep and f in eval_frames
            if keep_f:
                svgf_texels[f] = torch.empty((H, W, 8), dtype=torch.int16, device="cuda")
            sv.frame(c_f, vp_f, g_f, g_prev, t_f, den, time=f / 60.0, texels_out=svgf_texels[f] if keep_f else None)
            if keep_f:
                svgf_out[f] = den.clone()

    svgf_sequence(keep=True)
    torch.cuda.synchronize()
    s0, s1 = ev(), ev()
    s0.record(stream)
    svgf_sequence()
    s1.record(stream)
    torch.cuda.synchronize()
    svgf_ms = s0.elapsed_time(s1) / nf
    den_eager = den.clone()

    # the same sequence with each frame's pass chain replayed as one HIP graph
    # (captured on the first frame, updated in place on the others); the last
    # frame's output must equal the eager sequence's
    svgf_graph = Graph(ctx)

    def svgf_graph_sequence():
        sv.reset()
        for f, (c_f, vp_f, g_f, t_f) in enumerate(seq):
            g_prev = seq[f - 1][2] if f else g_f
            with svgf_graph.capture() as s_cap:
                sv.frame(c_f, vp_f, g_f, g_prev, t_f, den, time=f / 60.0, stream=s_cap)

    svgf_graph_sequence()
    torch.cuda.synchronize()
    s0.record(stream)
    svgf_graph_sequence()
    s1.record(stream)
    torch.cuda.synchronize()
    svgf_graph_ms = s0.elapsed_time(s1) / nf
    svgf_graph_equal = bool(torch.equal(den, den_eager))
    svgf_graph.close()
    del den_eager

    # ---- sun-shadow denoise (ShadowTemporalFilter + ShadowFilter) over the
    # same sequence, inputs resident; and the shadow pass with its
    # transversal output on the frame itself
    sd = ShadowDenoiser(ctx, W, H)
    sh_out = torch.empty((H, W), dtype=torch.uint8, device="cuda")

    shadow_eval = {}   # eval frames: (denoised shadow, raw mask)

    def shadow_sequence(keep=False):
        sd.reset()
        for f, (c_f, vp_f, g_f, _) in enumerate(seq):
            g_prev = seq[f - 1][2] if f else g_f
            m_f, tv_f = shadow_in[f]
            sd.frame(c_f, vp_f, g_f, g_prev, m_f, tv_f, sh_out)
            if keep and f in eval_frames:
                shadow_eval[f] = (sh_out.clone(), m_f.clone())

    shadow_sequence(keep=True)
    torch.cuda.synchronize()
    s0, s1 = ev(), ev()
    s0.record(stream)
    shadow_sequence()
    s1.record(stream)
    torch.cuda.synchronize()
    shadow_denoise_ms = s0.elapsed_time(s1) / nf
    m_tmp, tv_tmp = shadow_in[0]
    s0, s1 = ev(), ev()
    s0.record(stream)
    for _ in range(10):
        world.trace_shadow2(cam, gb, sun, m_tmp, tv_tmp, 350)
    s1.record(stream)
    torch.cuda.synchronize()
    shadow2_ms = s0.elapsed_time(s1) / 10
    sd.close()
    del shadow_in

    # ---- reflections (the reflection pass's traversal core) off the frame's
    # G-buffer: the reference's settings (1/4 resolution, 2 rough samples)
    # and full-resolution mirror reflections
    from voxelpathtracer_amd.world import reflect_params
    reflect = {}
    for name, kw in (("reference_settings", reflect_params(sun, 0)),
                     ("mirror_full_res", dict(div=1, max_iter=64, shadow_iter=150, spp=1, rough=0))):
        div = kw["div"]
        Wr, Hr = W // div, H // div
        r_dist = torch.empty((Hr, Wr), dtype=torch.int16, device="cuda")
        r_hit = torch.empty((Hr, Wr), dtype=torch.int32, device="cuda")
        r_ctr = torch.zeros(2, dtype=torch.int64, device="cuda")
        world.set_sample_counter(r_ctr)
        world.trace_reflect(cam, gb, sun, r_dist, r_hit, **kw)
        torch.cuda.synchronize()
        world.set_sample_counter(None)
        r_rays = int(r_ctr[1].item())
        s0, s1 = ev(), ev()
        s0.record(stream)
        for _ in range(20):
            world.trace_reflect(cam, gb, sun, r_dist, r_hit, **kw)
        s1.record(stream)
        torch.cuda.synchronize()
        r_ms = s0.elapsed_time(s1) / 20
        reflect[name] = {"target": [Wr, Hr], "spp": kw["spp"], "rough": kw["rough"], "ms": round(r_ms, 4),
                         "rays": r_rays, "samples": int(r_ctr[0].item()), "grays": round(r_rays / r_ms / 1e6, 4)}


    result = {
        "metric": "Grays/s DDA+GI @1080p (bit-exact hits); SVGF ms/frame; served p50/p99 latency",
        "value": round(grays, 4),
        "unit": "Grays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"GPU plain: {NX}x{NY}x{NZ} uint8 terrain + DF, {W}x{H} 1 spp primary + sun shadow "
                               f"+ 1 diffuse GI bounce + GI-hit shadow",
                   "world": [NX, NY, NZ], "image": [W, H], "rays_per_step": rays_per_step,
                   "parallelism": "single GPU"},
        "path": path,
        "plain": {"ms_per_frame": round(ms_plain, 4), "grays": round(grays_plain, 4),
                  "ms_per_frame_graph": round(ms_graph, 4), "graph_host_ms_per_frame": round(host_graph_ms, 4),
                  "graph": graph_stats,
                  "ms_per_frame_passes_in_order": round(ms_sequential, 4),
                  "pass_ms": {k: round(v, 4) for k, v in pass_ms.items()}, "work": counts},
        "persistent": persistent,
        "df_build_ms": round(df_ms, 4),
        "df_update": df_update,
        "svgf_ms_per_frame": round(svgf_ms, 4),
        "svgf_graph": {"ms_per_frame": round(svgf_graph_ms, 4), "matches_eager": svgf_graph_equal},
        "shadow_denoise": {"ms_per_frame": round(shadow_denoise_ms, 4), "shadow_pass_with_transversal_ms":
                           round(shadow2_ms, 4), "frames": nf},
        "reflect": reflect,
        "pipeline_ms_per_frame": round(pipeline_ms, 4),
        "pipeline_graph": {"ms_per_frame": round(pipeline_graph_ms, 4),
                           "host_ms_per_frame": round(pipeline_graph_host_ms, 4), "graphs": pipeline_graph_stats},
        "pipeline_cpp": pipeline_cpp,
        "roofline": {"bound": "hbm", "kernel": PASS_KERNELS[dom], "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": profiled_traffic(f"{dom} pass"), "algorithmic_bytes": alg[dom]},
    }

    unet_out = {}
    if not args.no_learned:
        result["learned"] = learned(ctx, world, seq, eval_frames, svgf_out, svgf_ms, args, keep=unet_out)
    if not args.no_final_image:
        result["final_image"] = final_image(ctx, world, seq, cur_vps, eval_frames, svgf_texels, shadow_eval,
                                            unet_out)
    del seq, svgf_texels, shadow_eval, unet_out
    if not args.no_serve:
        result["serve"] = serve(args)
    if args.copy_probe:
        copy_probe(args.copy_probe)
    if not args.no_cpu_baseline:
        # the frame's GPU outputs again (the buffers served as scratch since)
        world.trace_primary(cam, gb, 350)
        world.trace_shadow(cam, gb, sun, mask, 350)
        world.trace_gi(cam, gb, prm, rgb, tex)
        torch.cuda.synchronize()
        result["cpu_baseline"] = cpu_baseline(grid, gb, mask, rgb, cam, prm, args.cpu_reps)
    print(json.dumps(result), flush=True)


def cpp_pipeline(W: int, H: int, frames: int = 16, warmup: int = 4) -> dict:
    """The same frame loop through the C++ host (VoxelRT::Pipeline, the
    in-tree example binary): one launch per pass, and each frame's three
    chains replayed as HIP graphs; ms per frame over `frames` frames after
    `warmup` (GPU events in the binary)."""
    import subprocess
    from voxelpathtracer_amd.build import DEMO
    out = {}
    for mode in ("plain", "graph"):
        r = subprocess.run([str(DEMO), str(W), str(H), str(frames), f"--{mode}", "--warmup", str(warmup)],
                           capture_output=True, text=True, timeout=120)
        if r.returncode != 0:
            raise RuntimeError(f"vpt_frame_demo --{mode} failed: {r.stderr[-500:]}")
        out[f"{mode}_ms_per_frame"] = json.loads(r.stdout.strip().splitlines()[-1])["ms_per_frame"]
    out.update(frames=frames, warmup=warmup)
    return out


def df_update_timing(world, grid, torch, edits: int = 24):
    """One block placed on the terrain surface, then removed again: GPU time
    (events around the flush + vpt_world_set_blocks + vpt_df_update) and
    wall time to completion, beside the recomputed box's size (read in a
    second pass over the same edits).  The world ends as it started."""
    rng = np.random.default_rng(7)
    sites = []
    while len(sites) < edits:
        x, z = int(rng.integers(0, NX)), int(rng.integers(0, NZ))
        y = int(np.argmin(grid[z, :, x] > 0))
        if not grid[z, y, x]:                 # skip columns solid to the top
            sites.append((x, y, z))
    stream = torch.cuda.current_stream()
    gpu_ms, wall_ms, cells = [], [], []
    for x, y, z in sites:
        for block in (5, 0):
            world.SetBlock(x, y, z, block)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            e0.record(stream)
            world.UpdateDistanceField()
            e1.record(stream)
            torch.cuda.synchronize()
            wall_ms.append((time.perf_counter() - t0) * 1e3)
            gpu_ms.append(e0.elapsed_time(e1))
    for x, y, z in sites:
        for block in (5, 0):
            world.SetBlock(x, y, z, block)
            x0, y0, z0, x1, y1, z1 = world.UpdateDistanceField(region=True)
            cells.append((x1 - x0) * (y1 - y0) * (z1 - z0))
    gpu_ms, wall_ms = gpu_ms[4:], wall_ms[4:]     # first calls warm up
    return {"edit": "one block placed on / removed from the surface", "samples": len(gpu_ms),
            "gpu_ms_median": round(float(np.median(gpu_ms)), 4), "gpu_ms_p90": round(float(np.percentile(gpu_ms, 90)), 4),
            "wall_ms_median": round(float(np.median(wall_ms)), 4),
            "box_cells_median": int(np.median(cells)), "box_cells_max": int(max(cells)),
            "local_fraction": float(np.mean([c < NX * NY * NZ for c in cells]))}


def profiled_traffic(kernel: str):
    """HBM bytes per launch of `kernel` (or per call of a pass, summed over
    its kernels) measured by rocprofv3 (FETCH_SIZE and WRITE_SIZE runs scaled
    by a plain copy's readings), as recorded in
    profiles/frame_1080p_kernels.md; None when not recorded."""
    from pathlib import Path
    p = Path(__file__).resolve().parent / "profiles" / "frame_1080p_kernels.md"
    try:
        for line in p.read_text().splitlines():
            if line.startswith(f"traffic_bytes_per_launch {kernel}:"):
                return int(float(line.split(":", 1)[1]))
    except OSError:
        pass
    return None


def serve(args):
    """The serving benchmark (C++ host over vpt_server): 16 resident worlds,
    ray-query requests of log-uniform size in [1K, 2M] rays, kinds
    primary : shadow : GI = 1 : 1 : 2; plain (one launch per request) vs
    batched (continuous batching, persistent kernel): saturated throughput
    and p50 / p99 latency under Poisson arrivals at half the plain path's
    throughput."""
    import subprocess
    from pathlib import Path
    exe = Path(__file__).resolve().parent / "voxelpathtracer_amd" / "_build" / "vpt_serve_bench"
    cmd = [str(exe), "--requests", str(args.serve_requests), "--loads", "0.5"]
    try:
        res = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    except subprocess.TimeoutExpired:
        return {"error": "serving benchmark timed out"}
    if res.returncode != 0:
        return {"error": f"serving benchmark failed ({res.returncode}): {res.stderr[-400:]}"}
    out = json.loads(res.stdout.strip().splitlines()[-1])
    sat = out["saturated"]
    ld = out["loads"][0]
    best = out["best_batched"]
    out["summary"] = {
        "batched_kernel": best,
        "plain_grays_s": sat["plain"]["grays_s"], "batched_grays_s": sat[best]["grays_s"],
        "throughput_ratio": out["ratio_saturated"],
        "p50_ms": {"plain": ld["plain"]["p50_ms"], "batched": ld[best]["p50_ms"]},
        "p99_ms": {"plain": ld["plain"]["p99_ms"], "batched": ld[best]["p99_ms"]},
        "batched_at_own_half_load": {"grays_s": ld["best_batched_own_load"]["grays_s"],
                                     "p50_ms": ld["best_batched_own_load"]["p50_ms"],
                                     "p99_ms": ld["best_batched_own_load"]["p99_ms"]},
    }
    return out


def learned(ctx, world, seq, eval_frames, svgf_out, svgf_ms, args, keep=None):
    """Learned denoiser: trained on the GPU by its own kernels on pairs the
    tracer renders (1 spp -> 256 spp, 128x128 windows of a 1080p frustum
    over random seeded worlds and cameras), then run at 1920x1080 on the
    SVGF sequence.  PSNR (tonemapped x/(1+x)) against 256-spp renders of the
    same frames for the raw 1-spp irradiance, SVGF and the network; time per
    frame of the network (assembly + forward) beside SVGF's.  keep: filled
    with the network's irradiance per evaluated frame."""
    import torch
    from voxelpathtracer_amd import gi_params
    from voxelpathtracer_amd.unet import UNet, irradiance_decode, make_dataset, texel_accumulate, tonemap_psnr, train
    H, W = seq[0][2].height, seq[0][2].width
    u = UNet(ctx, 1, H, W)
    u.init(1)
    t0 = time.perf_counter()
    x, t = make_dataset(ctx, u, args.unet_pairs, size=128, spp=256, seed=11, worlds=8)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    losses = train(u, x, t, steps=args.unet_steps, batch=16, lr=2e-3)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    del x, t
    stream = torch.cuda.current_stream()
    rgb = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
    tex = torch.empty((H, W, 8), dtype=torch.int16, device="cuda")
    acc = torch.empty((H, W, 6), dtype=torch.float32, device="cuda")
    ref = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
    noisy = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
    out = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
    psnr = {"input_1spp": [], "svgf": [], "unet": []}
    for f in eval_frames:
        c_f, _, g_f, t_f = seq[f]
        for k in range(256):
            world.trace_gi(c_f, g_f, gi_params(50.0, 100000 + 256 * f + k), rgb, tex)
            texel_accumulate(tex, W * H, acc, k == 0)
        irradiance_decode(g_f, ref, acc=acc, scale=1.0 / 256)
        irradiance_decode(g_f, noisy, tex=t_f)
        u.denoise(g_f, t_f, out)
        torch.cuda.synchronize()
        r = ref.cpu().numpy()
        psnr["input_1spp"].append(tonemap_psnr(noisy.cpu().numpy(), r))
        psnr["svgf"].append(tonemap_psnr(svgf_out[f].cpu().numpy(), r))
        psnr["unet"].append(tonemap_psnr(out.cpu().numpy(), r))
        if keep is not None:
            keep[f] = out.clone()
    # time per frame at 1080p
    c_f, _, g_f, t_f = seq[-1]
    for _ in range(3):
        u.denoise(g_f, t_f, out)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 20
    e0.record(stream)
    for _ in range(reps):
        u.denoise(g_f, t_f, out)
    e1.record(stream)
    torch.cuda.synchronize()
    unet_ms = e0.elapsed_time(e1) / reps
    u.close()
    mean = {k: round(float(np.mean(v)), 3) for k, v in psnr.items()}
    # MFMA work of one 1080p forward: 2 * 9 * cin * cout per output pixel per layer
    flops = 0
    for cin, cout, lv in ((32, 32, 0), (32, 32, 0), (32, 64, 1), (64, 64, 1), (64, 128, 2), (128, 128, 2),
                          (192, 64, 1), (64, 64, 1), (96, 32, 0), (32, 32, 0)):
        flops += 2 * 9 * cin * cout * (W >> lv) * (H >> lv)
    return {"pairs": args.unet_pairs, "pair_size": 128, "ref_spp": 256, "train_steps": args.unet_steps,
            "batch": 16, "dataset_s": round(t1 - t0, 2), "train_s": round(t2 - t1, 2),
            "loss_first": round(float(np.mean(losses[:10])), 5), "loss_last": round(float(np.mean(losses[-50:])), 5),
            "eval_frames": eval_frames, "psnr_db": {k: [round(x, 3) for x in v] for k, v in psnr.items()},
            "psnr_mean_db": mean, "unet_ms_per_frame": round(unet_ms, 4), "svgf_ms_per_frame": round(svgf_ms, 4),
            "unet_tflops": round(flops / (unet_ms * 1e-3) / 1e12, 2), "unet_gflop_per_frame": round(flops / 1e9, 2)}


def _texels_from_irradiance(E, ao_texels):
    """Texels whose SHToIrridiance at any normal is E (H, W, 3): L1 = 0,
    L0 = Y / (2 * 0.886226) and CoCg over the decode's CoCg scale (YCoCg-R
    of E), rounded to binary16; the AO byte from ao_texels."""
    import torch
    r, g, b = E.unbind(-1)
    Y = r / 4 + g / 2 + b / 4
    co, cg = r - b, g - (r + b) / 2
    l0 = (Y / (2.0 * 0.886226)).half().float()
    y_dec = 2.0 * (0.886226 * l0)
    kk = (y_dec * 0.282095) / (l0 + 1e-6)
    kk = torch.where(kk > 0, kk, torch.ones_like(kk))
    out = torch.zeros(E.shape[:2] + (8,), dtype=torch.float16, device=E.device)
    out[..., 3] = l0.half()
    out[..., 4] = (co / kk).half()
    out[..., 5] = (cg / kk).half()
    t = out.view(torch.int16)
    t[..., 6] = 0
    t[..., 7] = ao_texels[..., 7] & 0xFF
    return t


def _accumulate_texels(world, cam, gb, frame, spp, W, H):
    """Mean of `spp` GI texel planes (seeds 100000 + spp * frame + k): SH and
    CoCg in binary16, AO byte rounded."""
    import torch
    from voxelpathtracer_amd import gi_params
    from voxelpathtracer_amd.unet import texel_accumulate
    rgb = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
    tex = torch.empty((H, W, 8), dtype=torch.int16, device="cuda")
    acc = torch.empty((H, W, 6), dtype=torch.float32, device="cuda")
    ao = torch.zeros((H, W), dtype=torch.float32, device="cuda")
    for k in range(spp):
        world.trace_gi(cam, gb, gi_params(50.0, 100000 + spp * frame + k), rgb, tex)
        texel_accumulate(tex, W * H, acc, k == 0)
        ao += (tex[..., 7] & 0xFF).float()
    out = torch.zeros((H, W, 8), dtype=torch.float16, device="cuda")
    out[..., :6] = (acc / spp).half()
    t = out.view(torch.int16)
    t[..., 6] = 0
    t[..., 7] = torch.round(ao / spp).to(torch.int16)
    return t


def final_image(ctx, world, seq, cur_vps, eval_frames, svgf_texels, shadow_eval, unet_out, ref_spp=256):
    """Final-image quality of the frame loop: the composite (vpt_composite)
    of each evaluated 1080p frame of the SVGF sequence with the reflections
    at the reference's settings (1/4 resolution, 2 rough samples) shaded
    from the same planes, for (a) the raw 1-spp GI texels and the raw shadow
    mask, (b) the SVGF chain's texels and the denoised shadow, (c) the
    learned denoiser's irradiance (as texels) and the denoised shadow;
    against the converged frame: 256-spp GI texels, the exact (hard) shadow
    and the reflection colour averaged over 256 seeds.  PSNR of the
    tonemapped x / (1 + x) image (peak 1) over the whole frame."""
    import torch
    from voxelpathtracer_amd import gi_params
    from voxelpathtracer_amd.unet import tonemap_psnr
    from voxelpathtracer_amd.world import composite, composite_params, reflect_params, reflect_shading
    _, _, g0, _ = seq[0]
    H, W = g0.height, g0.width
    sun = list(gi_params(50.0, 0).sun_dir)
    rp = reflect_params(sun, 0)
    div = rp["div"]
    dist = torch.empty((H // div, W // div), dtype=torch.int16, device="cuda")
    rcol = torch.empty((H // div, W // div, 4), dtype=torch.int16, device="cuda")
    racc = torch.empty((H // div, W // div, 4), dtype=torch.float32, device="cuda")
    img = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")

    def render(f, texels, shadow, seeds):
        c_f, _, g_f, _ = seq[f]
        prm = gi_params(50.0, f)
        shading = reflect_shading(prm, cur_vps[f], texels, shadow)
        for k, seed in enumerate(seeds):
            world.trace_reflect_shaded(c_f, g_f, sun, shading, dist, rcol, **{**rp, "frame": seed})
            if len(seeds) > 1:
                v = rcol.view(torch.float16).float()
                if k == 0:
                    racc.copy_(v)
                else:
                    racc.add_(v)
        if len(seeds) > 1:
            rcol.view(torch.float16).copy_((racc / len(seeds)).half())
        composite(ctx, c_f, g_f, composite_params(prm), texels, shadow, rcol, div, img)
        torch.cuda.synchronize()
        return img.cpu().numpy()

    psnr = {"input_1spp": [], "svgf": []}
    if unet_out:
        psnr["unet"] = []
    for f in eval_frames:
        c_f, _, g_f, t_f = seq[f]
        den_shadow, mask = shadow_eval[f]
        ref_tex = _accumulate_texels(world, c_f, g_f, f, ref_spp, W, H)
        ref = render(f, ref_tex, mask * 255, [1000 + ref_spp * f + k for k in range(ref_spp)])
        psnr["input_1spp"].append(tonemap_psnr(render(f, t_f, mask * 255, [f]), ref))
        psnr["svgf"].append(tonemap_psnr(render(f, svgf_texels[f], den_shadow, [f]), ref))
        if unet_out:
            u_tex = _texels_from_irradiance(unet_out[f], svgf_texels[f])
            psnr["unet"].append(tonemap_psnr(render(f, u_tex, den_shadow, [f]), ref))
    return {"eval_frames": eval_frames, "ref_spp": ref_spp, "image": [W, H],
            "psnr_db": {k: [round(x, 3) for x in v] for k, v in psnr.items()},
            "psnr_mean_db": {k: round(float(np.mean(v)), 3) for k, v in psnr.items()}}


def cpu_baseline(grid, gb, mask, rgb, cam, prm, reps):
    """The oracle (C++ transcription, OpenMP over rows) on the same 1080p
    DDA+GI frame on this host's cores: DF build, primary, shadow and GI,
    median of `reps` after a warm-up.  Its outputs are also compared with
    the GPU's (bit-exact hits)."""
    import oracle
    import torch
    iv = np.array(cam.inv_view, np.float32)
    ip = np.array(cam.inv_proj, np.float32)
    W, H = cam.width, cam.height
    sun = np.array(prm.sun_dir, np.float32)
    oprm = oracle.gi_params_from(prm)
    times = []
    for r in range(reps + 1):
        t0 = time.perf_counter()
        df = oracle.df_build(grid)
        t1 = time.perf_counter()
        p = oracle.trace_primary(grid, df, iv, ip, W, H, 350)
        s = oracle.trace_shadow(grid, df, iv, ip, p["t"], p["face"], sun, 350)
        g = oracle.trace_gi(grid, df, iv, ip, p["t"], p["face"], oprm)
        t2 = time.perf_counter()
        if r > 0:
            times.append((t1 - t0, t2 - t1))
    df_s = float(np.median([a for a, _ in times]))
    ray_s = float(np.median([b for _, b in times]))
    rays = int(W * H + s["rays"].sum() + g["rays"].sum())
    host = gb.numpy()
    parity = bool(np.array_equal(host["face"], p["face"]) and np.array_equal(host["voxel"], p["voxel"])
                  and np.array_equal(host["t"], p["t"]) and np.array_equal(mask.cpu().numpy(), s["mask"])
                  and np.array_equal(rgb.cpu().numpy(), g["rgb"]))
    threads = oracle.threads()
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    # the CPU plain-method config: DF build + 640x360 pinhole primary rays on
    # the same world, 1 warm-up + median of 5 (steady_clock inside the oracle)
    cam640, _, _ = _make_camera_640()
    plain = oracle.time_plain(grid, np.array(cam640.inv_view, np.float32), np.array(cam640.inv_proj, np.float32),
                              640, 360, 350, reps=5)
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = None
    return {"value": round(rays / ray_s / 1e9, 6), "unit": "Grays/s", "cores": threads, "kind": "port",
            "sample": f"one {W}x{H} frame (primary+shadow+GI, {rays} rays traced by the oracle: it also traces the "
                      f"shadow rays whose origin cell is solid, which the GPU skips as their mask is 1 either way), "
                      f"median of {reps} after 1 warm-up; DF build {df_s * 1e3:.1f} ms",
            "df_build_ms": round(df_s * 1e3, 2), "cpu_model": model, "gpu_matches_oracle": parity,
            "nproc": os.cpu_count(), "affinity_cpus": affinity, "omp_threads": threads,
            "cpu_plain_method": {"config": "384x128x384 terrain, DF build + 640x360 primary rays (350 steps), "
                                           "1 warm-up + median of 5",
                                 "df_build_ms": round(plain["df_ms"], 3), "primary_ms": round(plain["primary_ms"], 3),
                                 "primary_grays_s": round(640 * 360 / (plain["primary_ms"] * 1e-3) / 1e9, 6),
                                 "df_samples": int(plain["samples"]), "omp_threads": plain["threads"]}}


def _make_camera_640():
    from voxelpathtracer_amd import make_camera
    return make_camera(CAM_POS, CAM_YAW, CAM_PITCH, 640, 360)


if __name__ == "__main__":
    main()
