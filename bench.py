"""Benchmark: VoxelRT's base frame pipeline at 1920x1080 on a 256^3 grid.

One step = one frame of a 64-frame camera orbit: pathtrace (1 spp, 2 bounces,
sun-cone NEE) -> temporal reprojection (rgba16f history, cap 64) -> a-trous
steps 1/2/4, on one MI355X, scene and frame targets resident in HBM.

Prints one JSON line:
  value        Mrays/s over the timed frames (primary + shadow + bounce rays,
               counted in the kernel; the counts equal the oracle's)
  ms_per_step  ms per frame (wall clock over the timed frames)
  roofline     the path-trace kernel (dominant): algorithmic bytes per launch
               (4 B per brick step + 4 B per voxel step + 4 B per hit voxel +
               28 B of stores per pixel) / its HIP-event duration, vs 8 TB/s
  batched      the batched form (64 views of 8 scenes in one call per pass)
               vs the same views rendered one call each (--workload batch)
  served       the served form (continuous batching of a request stream over
               16 scenes up to 1024^3): saturated Mrays/s vs serial plain
               rendering, p50/p99 frame latency at 0.8 x plain capacity
               (--workload serve)
  cpu_baseline the CPU oracle on --cpu-frames (8) frames of the same orbit
               (all host cores, OpenMP over rows); beside it the oracle on one
               core and the reference's VoxelBench trace() on the traversal
               workload

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-forms]
                [--workload frame_1080|traverse|batch|serve|profile]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Mrays/s & ms/frame @1080p 256³; batched/served p50/p99 latency vs plain"


def _profile_value(name: str, field: str, key: str):
    """A figure of one kernel variant from a committed rocprofv3 --pmc summary
    (profiles/<name>: a line '<field>[<key>]: <number>'), or None when the
    summary does not hold that variant.  Fields: traffic_bytes_per_launch
    (FETCH_SIZE + WRITE_SIZE, calibrated), lanes_per_valu_inst,
    valu_lane_op_fraction."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        for line in open(path):
            if line.startswith(f"{field}[{key}]:"):
                v = float(line.split(":", 1)[1])
                return int(v) if field == "traffic_bytes_per_launch" else v
    except OSError:
        pass
    return None


def _pmc_traffic(name: str, key: str):
    return _profile_value(name, "traffic_bytes_per_launch", key)


def _pt_kernel_name(t):
    """The path-tracer kernel a tuning launches at frame_1080 (vrt_tuning's
    pt_kernel / pt_tile; the single-loop DDA fits LDS at 256^3)."""
    k, f = t["pt_kernel"], t["pt_tile"]
    if k == 8:
        return "k_pathtrace_wave"
    if (f in (2, 3) and k in (4, 7)) or (f in (4, 5) and k == 7):
        return "k_pathtrace_refill"
    if f == 1 and k == 7:
        return "k_pathtrace8"
    return "k_pathtrace"


def _oracle_frames(g, frames, w, h, fclamp, desc, threads):
    """Mrays/s and ms/frame of the CPU oracle chain on the first `frames`
    frames of the orbit."""
    import oracle
    from tests.chain import OracleChain
    from voxelrt_amd import configs as cf

    chain = OracleChain(g, w, h, 0, threads=threads)
    prev_vp = prev_pos = None
    rays = 0
    t0 = time.perf_counter()
    for i in range(frames):
        pos, yaw, pitch = cf.world_orbit(desc, i)
        u, vp = oracle.make_uniforms(pos, yaw, pitch, w, h, frame=i, bounces=2, flags=3,
                                     prev_view_proj=prev_vp, prev_cam_pos=prev_pos)
        out = chain.frame(u, fclamp)
        rays += out["counters"].rays
        prev_vp, prev_pos = vp, pos
    dt = time.perf_counter() - t0
    return rays / dt / 1e6, dt / frames * 1e3


def cpu_baseline(vox, frames, w, h, fclamp, desc):
    """CPU baselines of the workload: the oracle chain on `frames` frames on
    all host cores (the headline figure) and on one frame on one core, and
    the reference's own VoxelBench brick DDA, trace() (core.cpp:52-130,
    compiled from its sources into oracle/_ref), on 2^20 of the 2^24-ray
    traversal workload at p = 0.25 on all host cores."""
    import oracle
    from voxelrt_amd import configs as cf

    threads = min(os.cpu_count() or 1, 64)
    g = oracle.OracleGrid(vox)
    mrays, ms = _oracle_frames(g, frames, w, h, fclamp, desc, threads)
    out = {"value": round(mrays, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
           "ms_per_frame": round(ms, 1),
           "sample": f"{frames} frames of the 1920x1080 orbit (pathtrace + temporal + 3 a-trous), "
                     f"oracle C++ FP32, OpenMP over rows"}
    mrays1, ms1 = _oracle_frames(g, 1, w, h, fclamp, desc, 1)
    out["one_core"] = {"value": round(mrays1, 3), "unit": "Mrays/s", "cores": 1, "kind": "port",
                       "ms_per_frame": round(ms1, 1), "sample": "frame 0 of the same orbit, one thread"}
    rays = cf.traverse_rays(1 << 24, seed=77)
    tr = _traverse_cpu(cf.traverse_scene(0.25), rays, 0.25, sample=1 << 20, one_core=False)
    for b in tr["cpu_baseline"]:
        if b["kind"] == "reference":
            out["reference_trace"] = dict(b, workload="traversal 256^3, p = 0.25")
    return out


def traverse(args):
    """Traversal microbench: 256^3 random bricks, 2^24 rays resident in HBM,
    hits (voxel, face, steps, t) written back to HBM.  One JSON line per
    occupancy p.  Algorithmic bytes per ray: 24 in + 16 out + 4 per brick step
    + 4 per voxel step + 4 per hit."""
    import numpy as np

    import voxelrt_amd as v
    from voxelrt_amd import configs as cf

    n_rays = 1 << 24
    rays = cf.traverse_rays(n_rays, seed=77)
    ctx = v.Context(0)
    d_rays = ctx.device_array(rays.nbytes).upload(rays)
    d_hits = ctx.device_array(n_rays * 16)
    variants = {"lane": dict(trace_kernel=0),
                "lane_fast": dict(trace_kernel=4, dda_batch=1),
                "lane_fast_b16": dict(trace_kernel=4, dda_batch=16),
                "lane_fast_b16_w6": dict(trace_kernel=4, dda_batch=16, trace_waves=6),
                "persistent_fast_b16": dict(trace_kernel=5, trace_blocks=1024, trace_steps=8, trace_min_refill=16,
                                            dda_batch=16),
                "wave_m16_f32": dict(PROFILED_TRAVERSE),
                "wave_m16_f32_768": dict(PROFILED_TRAVERSE, trace_blocks=768),
                "wave_m16_f32_w6": dict(PROFILED_TRAVERSE, trace_waves=6)}
    # every variant from the same base (an autotuned tuning of the previous
    # scene must not leak into the next scene's fixed variants)
    base = dict(trace_waves=0, empty_skip=0, dda_batch=1)
    for p in (0.05, 0.25, 0.6):
        vox = cf.traverse_scene(p)
        sc = ctx.upload(vox)
        if not args.no_cpu:
            print(json.dumps(_traverse_cpu(vox, rays, p)), flush=True)
        for name in list(variants) + ["autotuned"]:
            if name == "autotuned":
                tuned = ctx.autotune(sc, n_rays).as_dict()  # tuned at the measured size
            else:
                ctx.set_tuning(**dict(base, **variants[name]))
            _traverse_one(args, ctx, sc, p, name, d_rays, d_hits, n_rays,
                          tuned if name == "autotuned" else variants[name])
        sc.release()
    ctx.close()


def profile(args):
    """Target of the committed rocprofv3 summaries (kernel traces and --pmc
    passes): (1) --steps frames of frame_1080 with the path-tracer variant
    given by --pt-kernel / --pt-waves / --pt-tile / --pt-order /
    --pt-min-refill / --pt-blocks; (2) for each --traverse-p occupancy, two
    launches over 2^24 rays of the traversal variant --trace-kernel
    (--dda-batch, 1024 blocks, --trace-steps / --trace-min-refill for the
    persistent kernels), rays and hits resident in HBM; (3) two device-to-device copies
    of 256 MiB (vrt_copy), the calibration for FETCH_SIZE / WRITE_SIZE."""
    import numpy as np

    import voxelrt_amd as v
    from voxelrt_amd import configs as cf

    print("profile target start", flush=True)
    ctx = v.Context(0)
    desc = dict(cf.FRAME_1080_SCENE)
    sc = ctx.upload(cf.frame_1080_scene())
    p = cf.FRAME_1080
    r = v.Renderer(ctx, sc, p["width"], p["height"], bounces=p["bounces"], temporal=True, denoise=True,
                   fclamp=p["fclamp"], max_history=p["max_history"], view_id=1)
    kernel = 8 if args.pt_kernel == "auto" else int(args.pt_kernel)
    ctx.set_tuning(pt_kernel=kernel, pt_waves=args.pt_waves, pt_order=args.pt_order, pt_tile=args.pt_tile,
                   pt_min_refill=args.pt_min_refill, pt_blocks=args.pt_blocks, pt_shade_min=args.pt_shade_min)
    for i in range(args.steps):
        r.set_pose(*cf.world_orbit(desc, i))
        r.frame()
    ctx.stats()
    print(f"{args.steps} frames done", flush=True)
    if args.traverse_p:
        n = 1 << 24
        rays = cf.traverse_rays(n, seed=77)
        d_rays = ctx.device_array(rays.nbytes).upload(rays)
        d_hits = ctx.device_array(n * 16)
        ctx.set_tuning(trace_kernel=args.trace_kernel, trace_blocks=1024, trace_steps=args.trace_steps,
                       trace_min_refill=args.trace_min_refill, dda_batch=args.dda_batch)
        for occ in args.traverse_p:
            sc2 = ctx.upload(cf.traverse_scene(occ))
            for _ in range(2):
                ctx.trace_rays_device(sc2, d_rays.ptr, n, d_hits.ptr)
            sc2.release()
            print(f"traverse p={occ} done", flush=True)
    a = ctx.device_array(256 << 20).upload(np.ones(256 << 20, np.uint8))
    b = ctx.device_array(256 << 20)
    for _ in range(2):
        ctx.copy_device(b, a)
    print("copy done (256 MiB device to device, x2)", flush=True)
    ctx.close()


def _traverse_cpu(vox, rays, p, sample=1 << 20, one_core=True):
    """CPU baselines on the first `sample` rays of the same workload: the
    oracle's FP32 DDA (all host cores and one core) and the reference's own
    VoxelBench brick DDA (core.cpp:52-130, double t, 4^3 bricks) compiled from
    its sources in oracle/_ref, on all host cores."""
    import numpy as np

    import oracle

    r = rays[:sample]
    g = oracle.OracleGrid(vox)
    cores = os.cpu_count() or 1
    out = {"metric": "traversal CPU baseline Mrays/s (256^3)", "p": p, "sample": f"first {len(r)} of the 2^24 rays",
           "cpu_baseline": []}
    for threads in ((cores, 1) if one_core else (cores,)):
        rr = r if threads > 1 else r[: len(r) // 8]
        t0 = time.perf_counter()
        oracle.trace_rays(g, rr, threads=threads)
        dt = time.perf_counter() - t0
        out["cpu_baseline"].append({"value": round(len(rr) / dt / 1e6, 2), "unit": "Mrays/s", "cores": threads,
                                    "kind": "port", "sample": f"{len(rr)} rays, oracle FP32 DDA (voxel.wgsl)"})
    if oracle.ref_lib() is not None:
        rv = np.ascontiguousarray(r.copy())
        rv[:, :3] /= 0.0625  # VoxelBench works in voxel units
        oracle.vb_trace(vox, rv[:1], naive=False, threads=cores)  # builds its grid + brickmap once
        t0 = time.perf_counter()
        oracle.vb_trace(vox, rv, naive=False, threads=cores)
        dt = time.perf_counter() - t0
        out["cpu_baseline"].append({"value": round(len(rv) / dt / 1e6, 2), "unit": "Mrays/s", "cores": cores,
                                    "kind": "reference",
                                    "sample": f"{len(rv)} rays, VoxelBench trace() (core.cpp:52-130), occupancy only"})
    return out


# the traversal variant whose HBM traffic profiles/traverse_pmc.txt holds
PROFILED_TRAVERSE = dict(trace_kernel=7, trace_blocks=1024, trace_steps=32, trace_min_refill=16)


def _traverse_one(args, ctx, sc, p, name, d_rays, d_hits, n_rays, tuning):
    import numpy as np
    for _ in range(args.warmup):
        ctx.trace_rays_device(sc, d_rays.ptr, n_rays, d_hits.ptr)
    s0 = ctx.stats()
    ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ms.append(ctx.trace_rays_device(sc, d_rays.ptr, n_rays, d_hits.ptr))
    s1 = ctx.stats()
    wall = time.perf_counter() - t0
    K = args.steps
    bs = (s1["brick_steps"] - s0["brick_steps"]) / K
    vs = (s1["voxel_steps"] - s0["voxel_steps"]) / K
    hits = (s1["hits"] - s0["hits"]) / K
    alg = 40 * n_rays + 4 * bs + 4 * vs + 4 * hits
    kms = float(np.mean(ms))
    gbs = alg / (kms * 1e-3) / 1e9
    print(json.dumps({"metric": "traversal Mrays/s (256^3, 2^24 rays)", "p": p, "variant": name,
                      "tuning": tuning,
                      "value": round(n_rays / (kms * 1e-3) / 1e6, 1), "unit": "Mrays/s",
                      "kernel_ms": round(kms, 4), "wall_ms_per_launch": round(wall / K * 1e3, 4),
                      "brick_steps_per_ray": round(bs / n_rays, 3),
                      "voxel_steps_per_ray": round(vs / n_rays, 3), "hit_frac": round(hits / n_rays, 4),
                      "roofline": {"bound": "hbm", "kernel": {5: "k_trace_persistent", 7: "k_trace_wave"}.get(
                                       tuning.get("trace_kernel"), "k_trace_rays"), "achieved": round(gbs, 1),
                                   "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                   "frac": round(gbs / HBM_PEAK_GBS, 4),
                                   "traffic": _pmc_traffic("traverse_pmc.txt", f"trace_kernel=7,p={p}")
                                   if tuning == PROFILED_TRAVERSE else None,
                                   "valu_lane_op_fraction": _profile_value("traverse_pmc.txt", "valu_lane_op_fraction",
                                                                           f"trace_kernel=7,p={p}")
                                   if tuning == PROFILED_TRAVERSE else None,
                                   "alg_bytes_per_launch": int(alg)}}), flush=True)


def batch(args, emit=True):
    """Batched form: 8 scenes of {64,128,256,512}^3 and 64 views of uneven
    size (256x256 .. 1920x1080, aspect 1:1 .. 21:9) and bounce count (1..6),
    temporal + a-trous on.  'plain' renders the views one call each (one
    launch per pass per view); 'batched' renders all 64 in one call (one
    launch per pass).  Aggregate Mrays/s over the timed frames."""
    import voxelrt_amd as v
    from voxelrt_amd import configs as cf

    descs, specs = cf.batch_views(64)
    ctx = v.Context(0)
    t0 = time.perf_counter()
    scenes = [ctx.upload(v.generate_scene(**d)) for d in descs]
    upload_s = time.perf_counter() - t0
    prev = {}

    def views_for(frame, base_id):
        views = []
        for k, sp in enumerate(specs):
            d = descs[sp["scene"]]
            pos, yaw, pitch = cf.world_orbit(d, (sp["orbit"] + frame) % cf.ORBIT_FRAMES)
            key = base_id + k
            pvp, ppos = prev.get(key, (None, None))
            u, vp = v.make_uniforms(pos, yaw, pitch, sp["width"], sp["height"], frame=sp["seed"] + frame,
                                    bounces=sp["bounces"], flags=3, prev_view_proj=pvp, prev_cam_pos=ppos)
            prev[key] = (vp, pos)
            views.append(v.View(scenes[sp["scene"]].handle, u, 48.0, key, 1, 0))
        return views

    results = {}
    pixels = sum(sp["width"] * sp["height"] for sp in specs)
    # kernel variant: the project's autotuner on the first view's pose and size
    d0 = descs[specs[0]["scene"]]
    u0, _ = v.make_uniforms(*cf.world_orbit(d0, specs[0]["orbit"]), 1920, 1080, bounces=2, flags=3)
    tuned = ctx.autotune(scenes[specs[0]["scene"]], 0, u0).as_dict()
    # batched form: launch mode, tile form and persistent grid timed on the
    # batch itself (its own view states; outside the timed region)
    tuned_b = ctx.autotune_batch(views_for(0, 90000)).as_dict()
    # plain per-view baseline: whichever of the single-view autotuner's choice
    # (it saw one view only) and the refill-tile kernels renders the 64 views
    # one by one fastest
    cands = [tuned] + [dict(tuned, pt_kernel=7, pt_waves=4, pt_tile=f) for f in (3, 5)]
    plain_ms = []
    for c, cand in enumerate(cands):
        ctx.set_tuning(**cand)
        vs = views_for(0, 80000 + 100 * c)
        for view in vs:
            ctx.render([view])
        ctx.stats()
        t0 = time.perf_counter()
        for view in vs:
            ctx.render([view])
        ctx.stats()
        plain_ms.append(time.perf_counter() - t0)
        for view in vs:
            ctx.release_view(view.view_id)
    tuned_p = cands[min(range(len(cands)), key=plain_ms.__getitem__)]
    modes = {"plain": 0, "batched": 0, "batched_queue": 1, "batched_graph": 2}
    for i, (mode, bm) in enumerate(modes.items()):
        ctx.set_tuning(**(tuned_p if mode == "plain" else dict(tuned_b, batch_mode=bm)))
        base = 10000 * (i + 1)
        for f in range(args.warmup):
            vs = views_for(f, base)
            if mode == "plain":
                for view in vs:
                    ctx.render([view])
            else:
                ctx.render(vs)
        s0 = ctx.stats()
        t0 = time.perf_counter()
        for f in range(args.steps):
            vs = views_for(args.warmup + f, base)
            if mode == "plain":
                for view in vs:
                    ctx.render([view])
            else:
                ctx.render(vs)
        s1 = ctx.stats()
        wall = time.perf_counter() - t0
        rays = s1["rays"] - s0["rays"]
        results[mode] = {"mrays_s": round(rays / wall / 1e6, 1), "ms_per_batch": round(wall / args.steps * 1e3, 3),
                         "gpu_ms_per_batch": round((s1["sum_ms_frame"] - s0["sum_ms_frame"]) / args.steps, 3),
                         "rays_per_batch": rays // args.steps}
    best = max(("batched", "batched_queue", "batched_graph"), key=lambda m: results[m]["mrays_s"])
    out = {"metric": "batched aggregate Mrays/s vs plain per-view (64 views, 8 scenes)", "unit": "Mrays/s",
           "value": results[best]["mrays_s"], "best_mode": best, **results,
           "speedup": round(results[best]["mrays_s"] / max(results["plain"]["mrays_s"], 1e-9), 2),
           "tuning": {k: tuned[k] for k in ("pt_kernel", "pt_waves", "pt_tile")},
           "tuning_plain": {k: tuned_p[k] for k in ("pt_kernel", "pt_waves", "pt_tile")},
           "tuning_batched": {k: tuned_b[k] for k in ("batch_mode", "pt_kernel", "pt_tile", "batch_tile",
                                                      "batch_blocks")},
           "views": len(specs), "pixels_per_batch": pixels, "scene_upload_s": round(upload_s, 2),
           "steps": args.steps, "warmup": args.warmup}
    for sc in scenes:
        sc.release()
    ctx.close()
    if emit:
        print(json.dumps(out), flush=True)
    return out


def _pct(xs, q):
    import numpy as np
    return round(float(np.percentile(np.asarray(xs, dtype=np.float64), q)), 3) if len(xs) else None


def _sample_stats(xs):
    """n, mean, median, sample stddev, cv, min, p95 (linear between order
    statistics) and max of one figure over repeats, as the reference's
    harness summarises them (test/bench.mjs:533-553)."""
    import numpy as np
    a = np.asarray([x for x in xs if np.isfinite(x)], dtype=np.float64)
    if a.size == 0:
        return None
    mean = float(a.mean())
    sd = float(a.std(ddof=1)) if a.size > 1 else 0.0
    r = lambda x: round(float(x), 4)
    return {"n": int(a.size), "mean": r(mean), "median": r(np.percentile(a, 50)), "stddev": r(sd),
            "cv": r(sd / mean if mean else 0.0), "min": r(a.min()), "p95": r(np.percentile(a, 95)),
            "max": r(a.max())}


def _perf_stats(frames, w, h, wall_ms_per_frame, pass_ms):
    """Per-frame figures of one timed run in the shape of the reference's
    performanceStats (test/bench.mjs:480-512): wall and GPU ms and fps,
    GPU ms per megapixel, the run's GPU cost, each pass's share of the GPU
    frame and the dominant pass.  pass_ms: mean GPU ms per frame by pass."""
    mp = w * h / 1e6
    total = sum(pass_ms.values())
    pct = {k: round(100.0 * ms / total, 2) for k, ms in pass_ms.items()}
    dom = max(pass_ms, key=pass_ms.get)
    return {"frames": frames, "pixels": w * h, "megapixels": round(mp, 4),
            "wall_ms_per_frame": round(wall_ms_per_frame, 4), "wall_fps": round(1e3 / wall_ms_per_frame, 1),
            "wall_run_ms": round(wall_ms_per_frame * frames, 2), "gpu_ms_per_frame": round(total, 4),
            "gpu_fps": round(1e3 / total, 1), "gpu_ms_per_megapixel": round(total / mp, 4),
            "gpu_run_ms": round(total * frames, 2), "timing_frames": frames,
            "gpu_pass_percent": pct, "dominant_pass": dom, "dominant_pass_percent": pct[dom]}


def serve(args, emit=True):
    """Served form vs plain serial rendering on one arrival trace.

    16 scenes (64^3 .. 1024^3), 32 camera streams (360p .. 1440p, spp 1..4,
    bounces 1..4, temporal + denoise), `--requests` frames.  Four runs, each
    with fresh view histories:
      plain_saturated   every request rendered alone with vrt_render, back to
                        back (the plain per-view kernel's capacity)
      served_saturated  every request submitted at t=0 (continuous batching)
      plain_poisson     Poisson arrivals at load x plain capacity, one request
                        at a time in arrival order (latency = done - arrival)
      served_poisson    the same arrivals through vrt_submit / vrt_wait
    Every run reads each frame's HDR back to the host."""
    import threading

    import numpy as np

    import voxelrt_amd as v
    from voxelrt_amd import configs as cf

    descs, cams, order = cf.serve_workload(args.requests)
    ctx = v.Context(0)
    ctx.serve_configure(streams=args.serve_streams, max_batch_views=args.serve_views,
                        max_batch_pixels=args.serve_pixels)
    t0 = time.perf_counter()
    # evictable: under budget pressure the least recently used idle grids
    # leave the device and come back on their next request
    scenes = [ctx.upload(v.generate_scene(**d), evictable=True) for d in descs]
    upload_s = time.perf_counter() - t0
    # kernel variant: the project's autotuner on a 256^3 scene at 1080p (both
    # plain and served use it)
    s256 = next(i for i, d in enumerate(descs) if d["n"] == 256)
    u0, _ = v.make_uniforms(*cf.world_orbit(descs[s256], 0), 1920, 1080, bounces=2, flags=3)
    tuned = ctx.autotune(scenes[s256], 0, u0).as_dict()
    # read-back destinations in page-locked memory (DMA at link speed; a
    # pageable destination's staged copies made the timing depend on the
    # driver's staging state)
    hdr = [ctx.host_array((c["height"], c["width"], 3), np.float32) for c in cams]

    def views_for(seq, base):
        prev, frame_of, out = {}, {}, []
        for s in seq:
            cam = cams[s]
            f = frame_of.get(s, 0)
            frame_of[s] = f + 1
            pos, yaw, pitch = cf.world_orbit(descs[cam["scene"]], (cam["orbit"] + f) % cf.ORBIT_FRAMES)
            pvp, ppos = prev.get(s, (None, None))
            u, vp = v.make_uniforms(pos, yaw, pitch, cam["width"], cam["height"], frame=cam["seed"] + f,
                                    bounces=cam["bounces"], flags=3, prev_view_proj=pvp, prev_cam_pos=ppos)
            prev[s] = (vp, pos)
            out.append(v.View(scenes[cam["scene"]].handle, u, 48.0, base + s, cam["spp"], 0))
        return out

    warm = list(range(len(cams)))  # one frame per stream: allocates every view's state
    # the served form's batched launch shape: the batch autotuner on 8 of the
    # streams' views (about a saturated batch), in --batch-mode (-1: any mode)
    modes = (0, 1, 2) if args.batch_mode < 0 else (args.batch_mode,)
    served_tuning = ctx.autotune_batch(views_for(warm[:8], 90000), modes).as_dict()
    # the server's batches are mostly of one or two views (its limits and the
    # arrival rate cut them), not the eight the batch autotuner timed: keep
    # whichever of its choice and the step-level refill batch form (batch_tile
    # 0) serves a saturated burst of 48 requests faster
    cands = [served_tuning]
    if served_tuning["pt_kernel"] == 8 and served_tuning["batch_tile"] != 0:
        cands.append(dict(served_tuning, batch_tile=0, batch_blocks=0))
    if len(cands) > 1:
        burst_s = []
        for c, cand in enumerate(cands):
            ctx.set_tuning(**cand)
            vs = views_for(warm + order[:48], 95000 + 100 * c)
            for k in range(len(warm)):
                ctx.wait(ctx.submit(vs[k]), hdr[warm[k]])
            t0 = time.perf_counter()
            tk = [ctx.submit(view) for view in vs[len(warm):]]
            for t, s in zip(tk, order[:48]):
                ctx.wait(t, hdr[s])
            burst_s.append(time.perf_counter() - t0)
            for k in range(len(warm)):
                ctx.release_view(vs[k].view_id)
        served_tuning = cands[min(range(len(cands)), key=burst_s.__getitem__)]

    def run_plain(base, arrivals=None):
        ctx.set_tuning(**tuned)
        views = views_for(warm + order, base)
        for k in range(len(warm)):
            ctx.render([views[k]], [{"hdr": hdr[warm[k]]}])
        views = views[len(warm):]
        lat, service, rays = [], [], 0
        start = time.perf_counter()
        for k, view in enumerate(views):
            if arrivals is not None:
                wait = start + arrivals[k] - time.perf_counter()
                if wait > 0:
                    time.sleep(wait)
            ts = time.perf_counter()
            fr = ctx.render([view], [{"hdr": hdr[order[k]]}])
            te = time.perf_counter()
            rays += fr[0].rays
            service.append((te - ts) * 1e3)
            lat.append((te - start - (arrivals[k] if arrivals is not None else 0.0)) * 1e3)
        span = time.perf_counter() - start
        return dict(mrays_s=round(rays / span / 1e6, 1), span_s=round(span, 3), rays=rays,
                    p50_ms=_pct(lat, 50), p99_ms=_pct(lat, 99), mean_service_ms=round(float(np.mean(service)), 3))

    def run_served(base, arrivals=None):
        ctx.set_tuning(**served_tuning)
        views = views_for(warm + order, base)
        for k in range(len(warm)):
            ctx.wait(ctx.submit(views[k]), hdr[warm[k]])
        views = views[len(warm):]
        tickets = [None] * len(views)
        ready = [threading.Event() for _ in views]
        start = time.perf_counter()

        def submitter():
            for k, view in enumerate(views):
                if arrivals is not None:
                    wait = start + arrivals[k] - time.perf_counter()
                    if wait > 0:
                        time.sleep(wait)
                tickets[k] = ctx.submit(view)
                ready[k].set()

        th = threading.Thread(target=submitter)
        th.start()
        lat, rays, sizes = [], 0, []
        for k in range(len(views)):
            ready[k].wait()
            res = ctx.wait(tickets[k], hdr[order[k]])
            rays += res["rays"]
            lat.append(res["timing"]["latency_ms"])
            sizes.append(res["timing"]["batch_views"])
        span = time.perf_counter() - start
        th.join()
        return dict(mrays_s=round(rays / span / 1e6, 1), span_s=round(span, 3), rays=rays,
                    p50_ms=_pct(lat, 50), p99_ms=_pct(lat, 99), mean_batch=round(float(np.mean(sizes)), 2))

    res = {}
    res["plain_saturated"] = run_plain(100000)
    res["served_saturated"] = run_served(200000)
    rate = args.load * 1e3 / res["plain_saturated"]["mean_service_ms"]  # requests / s
    gaps = np.random.default_rng(5).exponential(1.0 / rate, len(order))
    arrivals = np.cumsum(gaps) - gaps[0]
    res["plain_poisson"] = run_plain(300000, arrivals)
    res["served_poisson"] = run_served(400000, arrivals)
    st = ctx.serve_stats()
    out = {"metric": "served aggregate Mrays/s and p50/p99 frame latency vs plain serial rendering",
           "unit": "Mrays/s", "value": res["served_saturated"]["mrays_s"],
           "speedup_saturated": round(res["served_saturated"]["mrays_s"] / res["plain_saturated"]["mrays_s"], 2),
           "poisson_rate_per_s": round(rate, 2), "load_vs_plain": args.load,
           "runs": res, "serve_stats": st, "requests": len(order), "scenes": len(descs),
           "scene_bytes": sum(sc.nbytes() for sc in scenes),
           "scene_evictions": ctx.stats()["scene_evictions"],
           "camera_streams": len(cams), "scene_upload_s": round(upload_s, 2),
           "serve_config": {"streams": args.serve_streams, "max_batch_views": args.serve_views,
                            "max_batch_pixels": args.serve_pixels},
           "tuning_served": {k: served_tuning[k] for k in ("batch_mode", "pt_kernel", "pt_tile", "batch_tile",
                                                           "batch_blocks")},
           "tuning": {k: tuned[k] for k in ("pt_kernel", "pt_waves", "pt_tile")}}
    ctx.close()
    if emit:
        print(json.dumps(out), flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=32)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-forms", action="store_true", help="frame_1080: leave out the batched and served forms")
    ap.add_argument("--cpu-frames", type=int, default=8)
    ap.add_argument("--repeats", type=int, default=3,
                    help="frame_1080: runs of --steps frames summarised in perf (the timed run is the first)")
    ap.add_argument("--workload", default="frame_1080", choices=["frame_1080", "traverse", "batch", "serve", "profile"])
    ap.add_argument("--pt-order", type=int, default=2, help="profile workload: path-tracer tile order")
    ap.add_argument("--trace-kernel", type=int, default=7, help="profile workload: traversal variant")
    ap.add_argument("--dda-batch", type=int, default=1, help="profile workload: brick-entry batch")
    ap.add_argument("--trace-steps", type=int, default=32,
                    help="profile workload: trace 5 steps per refill check, trace 7 finished rays per pass")
    ap.add_argument("--trace-min-refill", type=int, default=16,
                    help="profile workload: trace 5 idle lanes per refill, trace 7 lanes waiting per level switch")
    ap.add_argument("--traverse-p", type=float, nargs="*", default=[], help="profile workload: occupancies")
    ap.add_argument("--requests", type=int, default=256)
    ap.add_argument("--load", type=float, default=0.8, help="Poisson rate as a fraction of plain capacity")
    # served form: four streams with batches of at most 8 views / 4M pixel-samples
    # (measured against 2-3 streams and 16-64 views: the same saturated rate,
    # p99 at the heaviest requests' own service time instead of above it)
    ap.add_argument("--serve-streams", type=int, default=4)
    ap.add_argument("--serve-views", type=int, default=8)
    ap.add_argument("--serve-pixels", type=int, default=4 << 20)
    ap.add_argument("--batch-mode", type=int, default=2, choices=[-1, 0, 1, 2],
                    help="served form's batch mode (-1: the batch autotuner's choice among all)")
    ap.add_argument("--pt-kernel", default="auto", choices=["auto", "0", "4", "7", "8"])
    ap.add_argument("--pt-waves", type=int, default=4, choices=[0, 4, 5, 6])
    ap.add_argument("--pt-min-refill", type=int, default=16, help="pt-kernel 8: lanes waiting before a level-switch pass")
    ap.add_argument("--pt-blocks", type=int, default=0, help="pt-kernel 8: persistent waves (0 = 16 per CU)")
    ap.add_argument("--pt-shade-min", type=int, default=40, help="pt-kernel 8: finished rays shaded together")
    ap.add_argument("--pt-tile", type=int, default=0, choices=[0, 1, 2, 3, 4, 5],
                    help="path-tracer tile form with a fixed --pt-kernel")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (--gpus {args.gpus} "
              f"requested). The frame shards per view; see the design notes.", file=sys.stderr)
        sys.exit(2)

    if args.workload == "traverse":
        return traverse(args)
    if args.workload == "batch":
        return batch(args)
    if args.workload == "serve":
        return serve(args)
    if args.workload == "profile":
        return profile(args)

    import voxelrt_amd as v
    from voxelrt_amd import configs as cf

    desc = dict(cf.FRAME_1080_SCENE)
    p = cf.FRAME_1080
    w, h = p["width"], p["height"]
    vox = cf.frame_1080_scene()
    ctx = v.Context(0)
    sc = ctx.upload(vox)
    r = v.Renderer(ctx, sc, w, h, bounces=p["bounces"], temporal=True, denoise=True,
                   fclamp=p["fclamp"], max_history=p["max_history"], view_id=1)

    def step(i):
        r.set_pose(*cf.world_orbit(desc, i % cf.ORBIT_FRAMES))
        r.frame()

    # launch shapes: the project's autotuner on the first orbit view (outside
    # the timed region), or a fixed kernel
    if args.pt_kernel == "auto":
        r.set_pose(*cf.world_orbit(desc, 0))
        tuned = ctx.autotune(sc, 0, r.view().u).as_dict()
    else:
        tuned = ctx.set_tuning(pt_kernel=int(args.pt_kernel), pt_waves=args.pt_waves,
                               pt_tile=args.pt_tile, pt_min_refill=args.pt_min_refill,
                               pt_blocks=args.pt_blocks, pt_shade_min=args.pt_shade_min).as_dict()

    for i in range(args.warmup):
        step(i)
    s0 = ctx.stats()  # stream sync: barrier before the timed region
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(args.warmup + k)
    s1 = ctx.stats()  # stream sync: barrier after (also reads the frames' HIP events)
    wall = time.perf_counter() - t0
    assert s1["timed_calls"] - s0["timed_calls"] == args.steps
    pt_ms = s1["sum_ms_pathtrace"] - s0["sum_ms_pathtrace"]
    tp_ms = s1["sum_ms_temporal"] - s0["sum_ms_temporal"]
    at_ms = s1["sum_ms_atrous"] - s0["sum_ms_atrous"]
    fr_ms = s1["sum_ms_frame"] - s0["sum_ms_frame"]

    K = args.steps
    rays = s1["rays"] - s0["rays"]
    bsteps = s1["brick_steps"] - s0["brick_steps"]
    vsteps = s1["voxel_steps"] - s0["voxel_steps"]
    hits = s1["hits"] - s0["hits"]
    ms_frame = wall / K * 1e3
    mrays = rays / wall / 1e6
    pkey = f"pt_kernel={tuned['pt_kernel']},pt_waves={tuned['pt_waves']},pt_tile={tuned['pt_tile']}"
    # path-trace kernel: algorithmic bytes per launch
    alg_bytes = (4 * bsteps + 4 * vsteps + 4 * hits) / K + 28 * w * h
    pt_avg_ms = pt_ms / K
    achieved = alg_bytes / (pt_avg_ms * 1e-3) / 1e9
    out = {
        "metric": METRIC, "value": round(mrays, 2), "unit": "Mrays/s", "n_gpus": 1,
        "steps": K, "warmup": args.warmup, "ms_per_step": round(ms_frame, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded 256^3 terrain + Cornell room + 3 lanterns + 40 trees)",
        "config": {"workload": "frame_1080: 1920x1080, 256^3 grid, 1 spp, 2 bounces, temporal "
                               "(rgba16f, cap 64) + a-trous 1/2/4, 64-frame orbit",
                   "width": w, "height": h, "grid": 256, "spp": 1, "bounces": 2},
        "roofline": {"bound": "hbm", "kernel": _pt_kernel_name(tuned), "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": _pmc_traffic("pathtrace_frame1080_pmc.txt", pkey),
                     "valu_lanes_per_inst": _profile_value("pathtrace_frame1080_pmc.txt", "lanes_per_valu_inst", pkey),
                     "valu_lane_op_fraction": _profile_value("pathtrace_frame1080_pmc.txt", "valu_lane_op_fraction",
                                                             pkey),
                     "alg_bytes_per_launch": int(alg_bytes), "avg_launch_ms": round(pt_avg_ms, 4)},
        "tuning": {k: tuned[k] for k in ("pt_kernel", "pt_waves", "pt_order", "pt_tile", "pt_blocks",
                                          "pt_min_refill", "pt_shade_min")},
        "breakdown": {"gpu_ms_pathtrace": round(pt_avg_ms, 4), "gpu_ms_temporal": round(tp_ms / K, 4),
                      "gpu_ms_atrous3": round(at_ms / K, 4), "gpu_ms_frame": round(fr_ms / K, 4),
                      "rays_per_frame": rays // K, "brick_steps_per_ray": round(bsteps / max(rays, 1), 2),
                      "voxel_steps_per_ray": round(vsteps / max(rays, 1), 2)},
    }
    # the reference harness's per-run figures for the timed run, and their
    # spread over repeats: the timed run above and args.repeats - 1 more runs
    # of K frames continuing the orbit (after the timed region; they do not
    # enter `value`)
    pass_ms = {"pathtrace": pt_ms / K, "temporal": tp_ms / K, "atrous": at_ms / K}
    perf = _perf_stats(K, w, h, ms_frame, pass_ms)
    walls, gpus = [ms_frame], [perf["gpu_ms_per_frame"]]
    nxt = args.warmup + K
    for _ in range(max(args.repeats, 1) - 1):
        a = ctx.stats()
        t1 = time.perf_counter()
        for k in range(K):
            step(nxt + k)
        b = ctx.stats()
        nxt += K
        walls.append((time.perf_counter() - t1) / K * 1e3)
        gpus.append(sum(b[f] - a[f] for f in ("sum_ms_pathtrace", "sum_ms_temporal", "sum_ms_atrous")) / K)
    perf["repeats"] = len(walls)
    perf["wall_ms_per_frame_stats"] = _sample_stats(walls)
    perf["gpu_ms_per_frame_stats"] = _sample_stats(gpus)
    out["perf"] = perf
    sc.release()
    ctx.close()
    if not args.no_forms:
        # the batched and served forms on their own workloads (a few seconds
        # each), summarised; --workload batch / serve print them in full
        ns = argparse.Namespace(**vars(args))
        ns.steps, ns.warmup = 8, 2
        b = batch(ns, emit=False)
        out["batched"] = {"value": b["value"], "unit": "Mrays/s", "plain_mrays_s": b["plain"]["mrays_s"],
                          "speedup_vs_plain": b["speedup"], "best_mode": b["best_mode"],
                          "ms_per_batch": b[b["best_mode"]]["ms_per_batch"],
                          "plain_ms_per_batch": b["plain"]["ms_per_batch"],
                          "workload": f"{b['views']} views of 8 scenes (64^3..512^3), 256x256..1920x1080, "
                                      f"bounces 1..6, temporal + a-trous, {b['steps']} timed batches"}
        sv = serve(ns, emit=False)
        runs = sv["runs"]
        out["served"] = {"value": sv["value"], "unit": "Mrays/s",
                         "saturated_speedup_vs_plain": sv["speedup_saturated"],
                         "p50_ms": runs["served_poisson"]["p50_ms"], "p99_ms": runs["served_poisson"]["p99_ms"],
                         "plain_p50_ms": runs["plain_poisson"]["p50_ms"],
                         "plain_p99_ms": runs["plain_poisson"]["p99_ms"],
                         "load_vs_plain": sv["load_vs_plain"], "poisson_rate_per_s": sv["poisson_rate_per_s"],
                         "workload": f"{sv['requests']} requests over {sv['camera_streams']} camera streams of "
                                     f"{sv['scenes']} scenes (64^3..1024^3), 360p..1440p, spp 1..4, bounces 1..4"}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(vox, args.cpu_frames, w, h, p["fclamp"], desc)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
