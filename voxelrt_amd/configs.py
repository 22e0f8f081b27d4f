"""Seeded workloads of the frame pipeline.

  frame_128   128^3 terrain + 32^3 Cornell room + 2 boxes, 640x360, 1 spp,
              2 bounces, sun cone 0.5 deg at elevation 40 deg, firefly clamp
              10, fixed pose, temporal and denoise off
  traverse    256^3 random bricks (occupancy p in {0.05, 0.25, 0.6}, voxel
              fill 0.5), rays with origins uniform in the grid, directions
              uniform on the sphere (traversal only)
  frame_1080  256^3 terrain + Cornell room + 3 lanterns + 40 trees,
              1920x1080, 1 spp, 2 bounces, 64-frame orbit, temporal (rgba16f,
              history cap 64) + 3 a-trous passes
  batch       8 scenes of {64,128,256,512}^3, 64 views of uneven size and
              bounce count, rendered together
  serve       Poisson stream of render requests over 16 scenes
"""
from __future__ import annotations

import functools
import math

import numpy as np

from .lib import SCENE_RANDOM_BRICKS, SCENE_WORLD, generate_scene, scene_spawn

VOXEL_SIZE = 0.0625
SUN_40DEG = dict(sun_azimuth=0.9, sun_elevation=math.radians(40.0),
                 sun_cos_radius=math.cos(math.radians(0.5)))


# ---------------------------------------------------------------------------
# frame_128
# ---------------------------------------------------------------------------
FRAME_128_SCENE = dict(kind=SCENE_WORLD, n=128, seed=1, h_min=8, h_max=40, room=32, room_h=32,
                       boxes=2, trees=0, lanterns=0)
FRAME_128 = dict(width=640, height=360, bounces=2, fclamp=10.0, temporal=False, denoise=False,
                 spp=1, **SUN_40DEG)


def frame_128_scene():
    return generate_scene(**FRAME_128_SCENE)


def frame_128_pose():
    pos, yaw, pitch = scene_spawn(**FRAME_128_SCENE)
    return tuple(pos), yaw, pitch


# ---------------------------------------------------------------------------
# traverse
# ---------------------------------------------------------------------------
def traverse_scene(p: float, n: int = 256, seed: int = 2):
    return generate_scene(kind=SCENE_RANDOM_BRICKS, n=n, seed=seed, p_brick=p, p_voxel=0.5)


def traverse_rays(count: int, n: int = 256, seed: int = 5, voxel_size: float = VOXEL_SIZE):
    """Origins uniform in the grid's box (metres), directions uniform on the
    sphere via z ~ U(-1,1), phi ~ U(0, 2pi) (VoxelBench harness.cpp:215-217)."""
    rng = np.random.default_rng(seed)
    ext = n * voxel_size
    o = rng.uniform(0.0, ext, (count, 3)).astype(np.float32)
    z = rng.uniform(-1.0, 1.0, count)
    ph = rng.uniform(0.0, 2.0 * np.pi, count)
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    d = np.stack([r * np.cos(ph), z, r * np.sin(ph)], 1).astype(np.float32)
    return np.ascontiguousarray(np.concatenate([o, d], 1), np.float32)


# ---------------------------------------------------------------------------
# frame_1080
# ---------------------------------------------------------------------------
FRAME_1080_SCENE = dict(kind=SCENE_WORLD, n=256, seed=3, h_min=16, h_max=80, room=80, room_h=51,
                        boxes=2, trees=40, lanterns=3)
FRAME_1080 = dict(width=1920, height=1080, bounces=2, fclamp=48.0, temporal=True, denoise=True,
                  spp=1, max_history=64.0)
ORBIT_FRAMES = 64


def frame_1080_scene():
    return generate_scene(**FRAME_1080_SCENE)


@functools.lru_cache(maxsize=64)
def _spawn(items: tuple):
    pos, _, _ = scene_spawn(**dict(items))
    return tuple(float(x) for x in pos)


def world_orbit(desc: dict, i: int, n: int = ORBIT_FRAMES):
    """Pose i of an n-frame orbit around the room of a WORLD scene."""
    from .renderer import orbit_pose
    spawn = _spawn(tuple(sorted(desc.items())))
    N = desc["n"]
    room = desc.get("room", 0) or N // 4
    center = (spawn[0], spawn[1], (N * 0.55) * VOXEL_SIZE)
    radius = max(room * 1.6, N * 0.2) * VOXEL_SIZE
    return orbit_pose(i, n, center, radius, room * 0.5 * VOXEL_SIZE)


# ---------------------------------------------------------------------------
# batch / serve
# ---------------------------------------------------------------------------
def world_desc(n: int, seed: int) -> dict:
    return dict(kind=SCENE_WORLD, n=n, seed=seed, h_min=n * 0.0625, h_max=n * 0.3125,
                room=max(16, int(n * 0.3)), room_h=max(16, int(n * 0.2)), boxes=2,
                trees=max(2, n // 16), lanterns=3)


BATCH_GRIDS = (64, 128, 256, 512)


def batch_views(count: int = 64, seed: int = 11):
    """`count` views over 8 scenes: sizes 256x256 .. 1920x1080, aspect 1:1 to
    21:9, bounces 1..6.  Returns (scene descs, views as dicts)."""
    rng = np.random.default_rng(seed)
    descs = [world_desc(BATCH_GRIDS[i % 4], 100 + i) for i in range(8)]
    aspects = [1.0, 4 / 3, 16 / 10, 16 / 9, 2.0, 21 / 9]
    views = []
    for k in range(count):
        s = int(rng.integers(0, 8))
        aspect = aspects[int(rng.integers(0, len(aspects)))]
        h = int(rng.integers(256, 1081))
        w = min(1920, int(round(h * aspect)))
        w -= w % 2
        views.append(dict(scene=s, width=w, height=h, bounces=int(rng.integers(1, 7)),
                          orbit=int(rng.integers(0, ORBIT_FRAMES)), seed=int(k * 7717)))
    return descs, views


SERVE_GRIDS = (64, 128, 256, 512, 64, 128, 256, 512, 64, 128, 256, 512, 128, 256, 512, 1024)
SERVE_HEIGHTS = (360, 540, 720, 1080, 1440)


def serve_workload(requests: int = 256, streams: int = 32, seed: int = 21, grids=SERVE_GRIDS):
    """Serving workload: 16 scenes (64^3 .. 1024^3), `streams` camera streams
    (each a fixed scene, 16:9 resolution 360p..1440p, spp 1..4, bounces 1..4,
    temporal + denoise, orbiting one pose per frame) and `requests` frames
    drawn from the streams in a seeded order.  Arrival times are assigned by
    the caller (Poisson at a chosen rate).  Returns (scene descs, streams,
    request list of stream indices)."""
    rng = np.random.default_rng(seed)
    descs = [world_desc(g, 300 + i) for i, g in enumerate(grids)]
    cams = []
    for k in range(streams):
        h = SERVE_HEIGHTS[int(rng.integers(0, len(SERVE_HEIGHTS)))]
        cams.append(dict(scene=int(rng.integers(0, len(descs))), width=(h * 16) // 9, height=h,
                         spp=int(rng.integers(1, 5)), bounces=int(rng.integers(1, 5)),
                         orbit=int(rng.integers(0, ORBIT_FRAMES)), seed=int(k * 131)))
    order = [int(x) for x in rng.integers(0, streams, requests)]
    return descs, cams, order
