"""Profiling target for rocprofv3 runs (kernel traces and --pmc passes).

python voxelrt_amd/profile_target.py [--pt-kernel K] [--pt-waves W] [--pt-order O]
        [--pt-tile F] [--trace-kernel T] [--dda-batch B] [--traverse-p P ...]
  1. renders 3 frames of frame_1080 (1920x1080, 256^3, 2 bounces, temporal +
     a-trous) with the given path-tracer variant (vrt_tuning.pt_kernel /
     pt_waves / pt_order / pt_tile; default the one-pixel-per-lane kernel on
     16x16 tiles, tiles from the last one),
  2. for each traversal grid occupancy P (default 0.25), traces 2^24 rays
     twice with the given traversal variant (default the persistent kernel,
     1024 blocks, 8 steps per refill check, refill at 16 idle lanes; brick
     entries batched per --dda-batch for the single-loop kernels 4 and 5),
  3. copies 256 MiB device to device twice (the calibration copy for the
     FETCH_SIZE / WRITE_SIZE counters).
"""
import argparse
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import torch  # noqa: E402

import voxelrt_amd as v  # noqa: E402
from voxelrt_amd import configs as cf  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--pt-kernel", type=int, default=0)
    ap.add_argument("--pt-waves", type=int, default=0)
    ap.add_argument("--pt-order", type=int, default=2)
    ap.add_argument("--pt-tile", type=int, default=0)
    ap.add_argument("--trace-kernel", type=int, default=1)
    ap.add_argument("--dda-batch", type=int, default=1)
    ap.add_argument("--traverse-p", type=float, nargs="*", default=[0.25])
    args = ap.parse_args()
    print("profile target start", flush=True)
    ctx = v.Context(0)
    desc = dict(cf.FRAME_1080_SCENE)
    sc = ctx.upload(cf.frame_1080_scene())
    r = v.Renderer(ctx, sc, 1920, 1080, bounces=2, temporal=True, denoise=True, fclamp=48.0, view_id=1)
    ctx.set_tuning(pt_kernel=args.pt_kernel, pt_waves=args.pt_waves, pt_order=args.pt_order, pt_tile=args.pt_tile)
    for i in range(3):
        r.set_pose(*cf.world_orbit(desc, i))
        r.frame()
    ctx.stats()
    print("frames done", flush=True)
    n = 1 << 24
    d_rays = torch.from_numpy(cf.traverse_rays(n, seed=77)).to("cuda:0")
    d_hits = torch.empty(n * 4, dtype=torch.int32, device="cuda:0")
    torch.cuda.synchronize()
    ctx.set_tuning(trace_kernel=args.trace_kernel, trace_blocks=1024, trace_steps=8, trace_min_refill=16,
                   dda_batch=args.dda_batch)
    for p in args.traverse_p:
        sc2 = ctx.upload(cf.traverse_scene(p))
        for _ in range(2):
            ctx.trace_rays_device(sc2, d_rays.data_ptr(), n, d_hits.data_ptr())
        sc2.release()
        print(f"traverse p={p} done", flush=True)
    a = torch.empty(256 << 20, dtype=torch.uint8, device="cuda:0")
    b = torch.empty_like(a)
    a.fill_(1)
    torch.cuda.synchronize()
    for _ in range(2):
        b.copy_(a)
    torch.cuda.synchronize()
    print("copy done (256 MiB device to device, x2)", flush=True)


if __name__ == "__main__":
    main()
