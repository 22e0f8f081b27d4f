"""Profiling target for rocprofv3 runs (kernel traces and --pmc passes).

python voxelrt_amd/profile_target.py [--pt-kernel K] [--pt-waves W] [--trace-kernel T]
  1. renders 3 frames of frame_1080 (1920x1080, 256^3, 2 bounces, temporal +
     a-trous) with the given path-tracer variant (vrt_tuning.pt_kernel /
     pt_waves; default the one-pixel-per-lane kernel),
  2. traces 2^24 rays twice on the p=0.25 traversal grid with the given
     traversal variant (default the persistent kernel, 1024 blocks, 4 steps
     per refill check, refill at 1 idle lane),
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
    ap.add_argument("--trace-kernel", type=int, default=1)
    args = ap.parse_args()
    print("profile target start", flush=True)
    ctx = v.Context(0)
    desc = dict(cf.FRAME_1080_SCENE)
    sc = ctx.upload(cf.frame_1080_scene())
    r = v.Renderer(ctx, sc, 1920, 1080, bounces=2, temporal=True, denoise=True, fclamp=48.0, view_id=1)
    ctx.set_tuning(pt_kernel=args.pt_kernel, pt_waves=args.pt_waves)
    for i in range(3):
        r.set_pose(*cf.world_orbit(desc, i))
        r.frame()
    ctx.stats()
    print("frames done", flush=True)
    sc2 = ctx.upload(cf.traverse_scene(0.25))
    n = 1 << 24
    d_rays = torch.from_numpy(cf.traverse_rays(n, seed=77)).to("cuda:0")
    d_hits = torch.empty(n * 4, dtype=torch.int32, device="cuda:0")
    torch.cuda.synchronize()
    ctx.set_tuning(trace_kernel=args.trace_kernel, trace_blocks=1024, trace_steps=4, trace_min_refill=1)
    for _ in range(2):
        ctx.trace_rays_device(sc2, d_rays.data_ptr(), n, d_hits.data_ptr())
    print("traverse done", flush=True)
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
