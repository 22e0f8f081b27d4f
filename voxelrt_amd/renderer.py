"""Renderer: the reference's per-frame host flow on libvoxelrt_hip.

Mirrors src/main.js: ``writeUniforms`` (399-426, previous-frame view-proj
and camera position carried over), ``tick`` (680-803: pathtrace -> temporal ->
a-trous x3 with accum/gbuf ping-pong), and ``capture({hdr})`` (571-661).
"""
from __future__ import annotations

import math

import numpy as np

from .lib import (ACCUM_RGBA16F, FLAG_DENOISE, FLAG_TEMPORAL, Context, Scene, View,
                  make_uniforms)


def orbit_pose(i: int, n: int, center, radius: float, height: float):
    """Camera pose i of an n-frame orbit around `center` (metres), looking at
    it: (pos, yaw, pitch).  yaw=0 faces +Z (math.js:95)."""
    a = 2.0 * math.pi * i / max(n, 1)
    cx, cy, cz = center
    pos = (cx - radius * math.sin(a), cy + height, cz - radius * math.cos(a))
    dx, dy, dz = cx - pos[0], cy - pos[1], cz - pos[2]
    yaw = math.atan2(dx, dz)
    pitch = math.atan2(dy, math.hypot(dx, dz))
    return pos, yaw, pitch


class Renderer:
    """One view of one scene, rendered frame by frame."""

    _next_id = 1

    def __init__(self, ctx: Context, scene: Scene, width: int, height: int, bounces: int = 2,
                 temporal: bool = True, denoise: bool = True, fclamp: float = 48.0,
                 max_history: float = 64.0, accum_format: int = ACCUM_RGBA16F, seed: int = 0,
                 spp: int = 1, sun_azimuth: float = 0.9, sun_elevation: float = 0.85,
                 sun_cos_radius=None, view_id=None):
        self.ctx, self.scene = ctx, scene
        self.width, self.height = width, height
        self.bounces, self.fclamp, self.max_history = bounces, fclamp, max_history
        self.temporal, self.denoise = temporal, denoise
        self.accum_format, self.seed, self.spp = accum_format, seed, spp
        self.sun = (sun_azimuth, sun_elevation, sun_cos_radius)
        if view_id is None:
            view_id = Renderer._next_id
            Renderer._next_id += 1
        self.view_id = view_id
        self.frame_index = 0
        self.prev_view_proj = None
        self.prev_cam_pos = None
        self.pose = ((0.0, 0.0, 0.0), 0.0, 0.0)
        self.last_view = None

    def set_pose(self, pos, yaw, pitch):
        self.pose = (tuple(pos), float(yaw), float(pitch))

    def view(self) -> View:
        pos, yaw, pitch = self.pose
        flags = (FLAG_TEMPORAL if self.temporal else 0) | (FLAG_DENOISE if self.denoise else 0)
        u, vp = make_uniforms(pos, yaw, pitch, self.width, self.height,
                              frame=self.frame_index + self.seed, bounces=self.bounces,
                              flags=flags, prev_view_proj=self.prev_view_proj,
                              prev_cam_pos=self.prev_cam_pos, sun_azimuth=self.sun[0],
                              sun_elevation=self.sun[1], sun_cos_radius=self.sun[2],
                              max_history=self.max_history)
        self._pending = (vp, pos)
        return View(self.scene.handle, u, self.fclamp, self.view_id, self.spp, self.accum_format)

    def advance(self):
        vp, pos = self._pending
        self.prev_view_proj, self.prev_cam_pos = vp, pos
        self.frame_index += 1

    def frame(self, outputs=()):
        """Render one frame; `outputs` names the targets to read back."""
        v = self.view()
        self.last_view = v
        outs = None
        arrays = {}
        if outputs:
            arrays = self._alloc(outputs)
            outs = [arrays]
        self.ctx.render([v], outs)
        self.advance()
        return arrays

    def capture(self, hdr: bool = True):
        """Render a frame and return what main.js's capture() returns
        (main.js:571-659): the presented rgb bytes (present.wgsl:40-55), the
        mean / std of their per-pixel (r+g+b)/3, and with hdr=True the linear
        HDR image."""
        out = self.frame(("ldr", "hdr") if hdr else ("ldr",))
        rgb = out["ldr"][..., :3]
        lum = rgb.astype(np.float64).sum(-1) / 3.0
        res = {"rgb": rgb, "mean": float(lum.mean()), "std": float(lum.std()), "w": self.width, "h": self.height}
        if hdr:
            res["hdr"] = out["hdr"]
        return res

    def _alloc(self, names):
        h, w = self.height, self.width
        shapes = {"radiance": ((h, w, 4), np.uint16), "albedo": ((h, w, 4), np.uint8),
                  "gbuf": ((h, w, 4), np.float32),
                  "accum": ((h, w, 4), np.uint16 if self.accum_format == 0 else np.float32),
                  "denoised": ((h, w, 4), np.uint16), "hdr": ((h, w, 3), np.float32),
                  "ldr": ((h, w, 4), np.uint8)}
        return {n: np.zeros(*shapes[n]) for n in names}
