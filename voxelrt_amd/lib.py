"""ctypes binding of the libvoxelrt_hip C ABI (include/voxelrt.h)."""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libvoxelrt_hip.so")
_lib = None

VOXEL_SIZE = 0.0625
FACE_NONE = 6
FACE_MISS = 255
VOXEL_MISS = 0xFFFFFFFF
NOALLOC_MAX_RAYS = 0x7FFFFFFF  # VRT_NOALLOC_MAX_RAYS
FLAG_TEMPORAL = 1
FLAG_DENOISE = 2
ACCUM_RGBA16F = 0
ACCUM_RGBA32F = 1
SCENE_RANDOM_BRICKS = 1
SCENE_WORLD = 2
SCENE_VOXELRT = 3
SCENE_EDITABLE = 1
SCENE_EVICTABLE = 2

ERRORS = {-1: "VRT_ERR_INVALID", -2: "VRT_ERR_DEVICE", -3: "VRT_ERR_NO_MEMORY",
          -4: "VRT_ERR_BUDGET", -5: "VRT_ERR_NOT_READY", -6: "VRT_ERR_CANCELLED",
          -7: "VRT_ERR_NO_DEVICE"}

HIT_DTYPE = np.dtype([("t", "<f4"), ("voxel", "<u4"), ("brick_steps", "<u2"),
                      ("voxel_steps", "<u2"), ("face", "u1"), ("_pad", "u1", (3,))])


class VrtError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Uniforms(C.Structure):
    """The 304-byte Uniforms block of common.wgsl:13-35."""
    _fields_ = [("inv_view_proj", C.c_float * 16), ("prev_view_proj", C.c_float * 16),
                ("cam_pos", C.c_float * 4), ("prev_cam_pos", C.c_float * 4),
                ("sun_dir", C.c_float * 4), ("sun_radiance", C.c_float * 4),
                ("params0", C.c_uint32 * 4), ("params1", C.c_float * 4),
                ("params2", C.c_uint32 * 4), ("params3", C.c_float * 4),
                ("params4", C.c_float * 4), ("params5", C.c_float * 4),
                ("params6", C.c_float * 4)]


assert C.sizeof(Uniforms) == 304


class CreateInfo(C.Structure):
    _fields_ = [("device", C.c_int32), ("flags", C.c_uint32), ("hbm_budget", C.c_uint64)]


class View(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("u", Uniforms), ("firefly_clamp", C.c_float),
                ("view_id", C.c_uint32), ("spp", C.c_uint32), ("accum_format", C.c_uint32)]


class Frame(C.Structure):
    _fields_ = [("radiance", C.c_void_p), ("albedo", C.c_void_p), ("gbuf", C.c_void_p),
                ("accum", C.c_void_p), ("denoised", C.c_void_p), ("hdr", C.c_void_p),
                ("rays", C.c_uint64), ("brick_steps", C.c_uint64), ("voxel_steps", C.c_uint64),
                ("ldr", C.c_void_p)]


class Stats(C.Structure):
    _fields_ = [("rays", C.c_uint64), ("brick_steps", C.c_uint64), ("voxel_steps", C.c_uint64),
                ("hits", C.c_uint64), ("frames", C.c_uint64), ("gpu_ms_pathtrace", C.c_double),
                ("gpu_ms_temporal", C.c_double), ("gpu_ms_atrous", C.c_double),
                ("gpu_ms_frame", C.c_double), ("hbm_resident", C.c_uint64),
                ("hbm_budget", C.c_uint64), ("timed_calls", C.c_uint64),
                ("sum_ms_pathtrace", C.c_double), ("sum_ms_temporal", C.c_double),
                ("sum_ms_atrous", C.c_double), ("sum_ms_frame", C.c_double),
                ("scene_evictions", C.c_uint64), ("scene_restores", C.c_uint64)]


class Tuning(C.Structure):
    _fields_ = [("trace_kernel", C.c_uint32), ("trace_blocks", C.c_uint32),
                ("trace_steps", C.c_uint32), ("trace_min_refill", C.c_uint32),
                ("pt_kernel", C.c_uint32), ("pt_blocks", C.c_uint32), ("unused0", C.c_uint32),
                ("pt_min_refill", C.c_uint32), ("unused1", C.c_uint32), ("pt_waves", C.c_uint32),
                ("batch_mode", C.c_uint32), ("batch_blocks", C.c_uint32), ("dda_batch", C.c_uint32),
                ("empty_skip", C.c_uint32), ("pt_order", C.c_uint32), ("pt_tile", C.c_uint32),
                ("trace_waves", C.c_uint32), ("pt_shade_min", C.c_uint32), ("batch_tile", C.c_uint32),
                ("reserved", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k not in ("reserved", "unused0", "unused1")}


class Request(C.Structure):
    _fields_ = [("view", View), ("flags", C.c_uint32), ("_pad", C.c_uint32)]


class ServeConfig(C.Structure):
    _fields_ = [("streams", C.c_uint32), ("max_batch_views", C.c_uint32), ("max_batch_pixels", C.c_uint64)]


class Timing(C.Structure):
    _fields_ = [("submit_ms", C.c_double), ("launch_ms", C.c_double), ("done_ms", C.c_double),
                ("latency_ms", C.c_double), ("gpu_ms", C.c_double), ("batch_views", C.c_uint32),
                ("stream", C.c_uint32), ("batch_id", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServeStats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("submitted", "completed", "cancelled", "failed", "batches",
                                          "batched_requests", "evictions", "queued", "running",
                                          "max_running_batches")]


class SceneDesc(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("nx", C.c_uint32), ("ny", C.c_uint32), ("nz", C.c_uint32),
                ("seed", C.c_uint64), ("p_brick", C.c_float), ("p_voxel", C.c_float),
                ("h_min", C.c_float), ("h_max", C.c_float), ("room", C.c_uint32),
                ("room_h", C.c_uint32), ("boxes", C.c_uint32), ("trees", C.c_uint32),
                ("lanterns", C.c_uint32), ("variant", C.c_uint32)]


def lib_path() -> str:
    return _LIB


def build(quiet: bool = True) -> None:
    """Compile libvoxelrt_hip.so for gfx950 (hipcc cross-compiles without a GPU)."""
    subprocess.run(["make", "-s", "-j8", "-C", os.path.join(_HERE, "csrc")], check=True,
                   stdout=subprocess.DEVNULL if quiet else None)


def load():
    """Load the HIP library; raises if it is not built (no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.isfile(_LIB):
            raise ImportError(f"{_LIB} is not built: run voxelrt_amd.build() "
                              f"(make -C voxelrt_amd/csrc)")
        L = C.CDLL(_LIB)
        _declare(L)
        _lib = L
    return _lib


def _declare(L):
    P = C.c_void_p
    I = C.c_int
    L.vrt_create.argtypes = [C.POINTER(CreateInfo), C.POINTER(P)]
    L.vrt_destroy.argtypes = [P]
    L.vrt_destroy.restype = None
    L.vrt_last_error.argtypes = [P]
    L.vrt_last_error.restype = C.c_char_p
    L.vrt_device_count.argtypes = [C.POINTER(C.c_int32)]
    L.vrt_scene_upload.argtypes = [P, P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64,
                                   C.POINTER(P)]
    L.vrt_scene_release.argtypes = [P, P]
    L.vrt_scene_upload_ex.argtypes = [P, P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32,
                                      C.POINTER(P)]
    L.vrt_scene_edit.argtypes = [P, P, P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                 C.c_uint32]
    L.vrt_scene_upload_ex.restype = I
    L.vrt_scene_edit.restype = I
    L.vrt_scene_edit_counts.argtypes = [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.vrt_scene_edit_counts.restype = I
    L.vrt_scene_restore.argtypes = [P, P]
    L.vrt_scene_restore.restype = I
    L.vrt_scene_resident.argtypes = [P, C.POINTER(C.c_int32)]
    L.vrt_scene_resident.restype = I
    L.vrt_scene_bricks.argtypes = [P, P, P, P]
    L.vrt_scene_dims.argtypes = [P, P, C.POINTER(C.c_uint64)]
    L.vrt_trace_rays.argtypes = [P, P, P, C.c_uint64, C.c_float, P]
    L.vrt_trace_rays_device.argtypes = [P, P, P, C.c_uint64, C.c_float, P, C.POINTER(C.c_float)]
    L.vrt_trace_rays_device_noalloc.argtypes = [P, P, P, C.c_uint64, C.c_float, P]
    L.vrt_device_alloc.argtypes = [P, C.c_uint64, C.POINTER(P)]
    L.vrt_device_free.argtypes = [P, P]
    L.vrt_copy.argtypes = [P, P, P, C.c_uint64, C.c_uint32]
    L.vrt_render.argtypes = [P, C.POINTER(View), C.c_uint32, C.POINTER(Frame)]
    L.vrt_view_release.argtypes = [P, C.c_uint32]
    L.vrt_host_alloc.argtypes = [P, C.c_uint64, C.POINTER(P)]
    L.vrt_host_free.argtypes = [P, P]
    L.vrt_temporal_pass.argtypes = [P, C.POINTER(Uniforms), P, P, P, P, C.c_uint32, P]
    L.vrt_atrous_pass.argtypes = [P, C.POINTER(Uniforms), P, C.c_uint32, P, C.c_int32, P]
    L.vrt_tuning_get.argtypes = [P, C.POINTER(Tuning)]
    L.vrt_tuning_set.argtypes = [P, C.POINTER(Tuning)]
    L.vrt_autotune.argtypes = [P, P, C.c_uint64, C.POINTER(Uniforms), C.POINTER(Tuning)]
    L.vrt_autotune_batch.argtypes = [P, C.POINTER(View), C.c_uint32, C.c_uint32, C.POINTER(Tuning)]
    L.vrt_stats_get.argtypes = [P, C.POINTER(Stats)]
    L.vrt_stats_reset.argtypes = [P]
    L.vrt_camera_matrices.argtypes = [P, C.c_double, C.c_double, C.c_double, C.c_double,
                                      C.c_double, C.c_double, P, P]
    L.vrt_camera_matrices.restype = None
    L.vrt_uniforms_init.argtypes = [C.POINTER(Uniforms), P, C.c_double, C.c_double, P, P,
                                    C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                    C.c_double, C.c_double, C.c_double, C.c_float]
    L.vrt_uniforms_init.restype = None
    L.vrt_generate_scene.argtypes = [C.POINTER(SceneDesc), P]
    L.vrt_scene_spawn.argtypes = [C.POINTER(SceneDesc), P, C.POINTER(C.c_double),
                                  C.POINTER(C.c_double)]
    L.vrt_serve_configure.argtypes = [P, C.POINTER(ServeConfig)]
    L.vrt_submit.argtypes = [P, C.POINTER(Request), C.POINTER(C.c_uint64)]
    L.vrt_poll.argtypes = [P, C.c_uint64, C.POINTER(Frame), C.POINTER(Timing)]
    L.vrt_wait.argtypes = [P, C.c_uint64, C.c_uint32, C.POINTER(Frame), C.POINTER(Timing)]
    L.vrt_cancel.argtypes = [P, C.c_uint64]
    L.vrt_serve_stats_get.argtypes = [P, C.POINTER(ServeStats)]
    for name in ("vrt_serve_configure", "vrt_submit", "vrt_poll", "vrt_wait", "vrt_cancel",
                 "vrt_serve_stats_get"):
        getattr(L, name).restype = I
    for name in ("vrt_create", "vrt_device_count", "vrt_scene_upload", "vrt_scene_release",
                 "vrt_scene_bricks", "vrt_scene_dims", "vrt_trace_rays", "vrt_trace_rays_device",
                 "vrt_trace_rays_device_noalloc", "vrt_device_alloc", "vrt_device_free", "vrt_copy",
                 "vrt_render", "vrt_view_release", "vrt_host_alloc", "vrt_host_free",
                 "vrt_temporal_pass", "vrt_atrous_pass",
                 "vrt_stats_get", "vrt_stats_reset", "vrt_generate_scene", "vrt_scene_spawn",
                 "vrt_tuning_get", "vrt_tuning_set", "vrt_autotune", "vrt_autotune_batch"):
        getattr(L, name).restype = I


def _ptr(a: np.ndarray):
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data_as(C.c_void_p)


def device_count() -> int:
    n = C.c_int32(0)
    load().vrt_device_count(C.byref(n))
    return n.value


# ---------------------------------------------------------------------------
# host helpers
# ---------------------------------------------------------------------------
def camera_matrices(pos, yaw, pitch, fov_y=70 * np.pi / 180, aspect=16 / 9, near=0.05,
                    far=400.0):
    p = np.asarray(pos, np.float64)
    vp = np.zeros(16, np.float32)
    ivp = np.zeros(16, np.float32)
    load().vrt_camera_matrices(_ptr(p), yaw, pitch, fov_y, aspect, near, far, _ptr(vp), _ptr(ivp))
    return vp, ivp


def make_uniforms(pos, yaw, pitch, width, height, frame=0, bounces=2, flags=0,
                  prev_view_proj=None, prev_cam_pos=None, sun_azimuth=0.9, sun_elevation=0.85,
                  sun_cos_radius=None, max_history=64.0):
    """Uniforms as main.js:399-426 writes them for the base preset.  Returns
    (uniforms, view_proj) so the caller can pass view_proj as the next frame's
    prev_view_proj."""
    u = Uniforms()
    p = np.asarray(pos, np.float64)
    pvp = None if prev_view_proj is None else np.ascontiguousarray(prev_view_proj, np.float32)
    ppos = None if prev_cam_pos is None else np.asarray(prev_cam_pos, np.float64)
    cr = float(np.cos(0.03)) if sun_cos_radius is None else float(sun_cos_radius)
    load().vrt_uniforms_init(C.byref(u), _ptr(p), yaw, pitch,
                             None if pvp is None else _ptr(pvp),
                             None if ppos is None else _ptr(ppos), int(width), int(height),
                             int(frame) & 0xFFFFFFFF, int(bounces), int(flags), sun_azimuth,
                             sun_elevation, cr, max_history)
    vp, _ = camera_matrices(pos, yaw, pitch, 70 * np.pi / 180, width / height)
    return u, vp


def generate_scene(kind=SCENE_WORLD, n=128, ny=None, nz=None, seed=0, p_brick=0.25,
                   p_voxel=0.5, h_min=0.0, h_max=0.0, room=0, room_h=0, boxes=0, trees=0,
                   lanterns=0, variant=0) -> np.ndarray:
    d = SceneDesc(kind, n, ny or n, nz or n, seed, p_brick, p_voxel, h_min, h_max, room, room_h,
                  boxes, trees, lanterns, variant)
    vox = np.zeros((d.nz, d.ny, d.nx), np.uint32)
    rc = load().vrt_generate_scene(C.byref(d), _ptr(vox))
    if rc != 0:
        raise VrtError(rc, "vrt_generate_scene rejected the description")
    return vox


def scene_spawn(kind=SCENE_WORLD, n=128, ny=None, nz=None, seed=0, **kw):
    d = SceneDesc(kind, n, ny or n, nz or n, seed, kw.get("p_brick", 0.25), kw.get("p_voxel", 0.5),
                  kw.get("h_min", 0.0), kw.get("h_max", 0.0), kw.get("room", 0),
                  kw.get("room_h", 0), kw.get("boxes", 0), kw.get("trees", 0),
                  kw.get("lanterns", 0), kw.get("variant", 0))
    pos = np.zeros(3, np.float64)
    yaw = C.c_double()
    pitch = C.c_double()
    rc = load().vrt_scene_spawn(C.byref(d), _ptr(pos), C.byref(yaw), C.byref(pitch))
    if rc != 0:
        raise VrtError(rc, "vrt_scene_spawn rejected the description")
    return pos, yaw.value, pitch.value


# ---------------------------------------------------------------------------
# context / scenes
# ---------------------------------------------------------------------------
class Scene:
    def __init__(self, ctx: "Context", handle: int, dims):
        self.ctx = ctx
        self.handle = handle
        self.dims = dims

    def bricks(self):
        nx, ny, nz = self.dims
        nb = (nx // 8) * (ny // 8) * (nz // 8)
        words = np.zeros(nb, np.uint32)
        masks = np.zeros(nb * 16, np.uint32)
        self.ctx._check(self.ctx.L.vrt_scene_bricks(self.ctx.h, self.handle, _ptr(words),
                                                    _ptr(masks)))
        return words, masks

    def nbytes(self) -> int:
        dims = np.zeros(3, np.uint32)
        b = C.c_uint64()
        self.ctx._check(self.ctx.L.vrt_scene_dims(self.handle, _ptr(dims), C.byref(b)))
        return b.value

    def edit(self, box: np.ndarray, origin):
        """Write box (sz, sy, sx) uint32 at voxel origin (x0, y0, z0) of an
        editable scene and update it on the device (the touched bricks in
        place, or a rebuild when their occupancy changes)."""
        box = np.ascontiguousarray(box, np.uint32)
        sz, sy, sx = box.shape
        x0, y0, z0 = (int(c) for c in origin)
        self.ctx._check(self.ctx.L.vrt_scene_edit(self.ctx.h, self.handle, _ptr(box), x0, y0, z0, sx, sy, sz))

    def restore(self):
        """Bring the scene back to the device if it was evicted."""
        self.ctx._check(self.ctx.L.vrt_scene_restore(self.ctx.h, self.handle))

    def resident(self) -> bool:
        r = C.c_int32()
        self.ctx._check(self.ctx.L.vrt_scene_resident(self.handle, C.byref(r)))
        return bool(r.value)

    def edit_counts(self):
        """(edits applied in place to the touched bricks, edits that rebuilt
        the scene)."""
        a, b = C.c_uint64(), C.c_uint64()
        self.ctx._check(self.ctx.L.vrt_scene_edit_counts(self.handle, C.byref(a), C.byref(b)))
        return a.value, b.value

    def release(self):
        if self.handle:
            self.ctx._check(self.ctx.L.vrt_scene_release(self.ctx.h, self.handle))
            self.handle = None


class Context:
    """One libvoxelrt_hip context on one HIP device."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        self.L = load()
        h = C.c_void_p()
        rc = self.L.vrt_create(C.byref(CreateInfo(device, 0, hbm_budget)), C.byref(h))
        if rc != 0:
            raise VrtError(rc, "vrt_create failed (no HIP device?)")
        self.h = h

    def close(self):
        if self.h:
            self.L.vrt_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int):
        if rc != 0:
            raise VrtError(rc, self.L.vrt_last_error(self.h).decode())

    def upload(self, vox: np.ndarray, key: int = 0, editable: bool = False, evictable: bool = False) -> Scene:
        vox = np.ascontiguousarray(vox, np.uint32)
        nz, ny, nx = vox.shape
        out = C.c_void_p()
        flags = (SCENE_EDITABLE if editable else 0) | (SCENE_EVICTABLE if evictable else 0)
        self._check(self.L.vrt_scene_upload_ex(self.h, _ptr(vox), nx, ny, nz, key, flags, C.byref(out)))
        return Scene(self, out.value, (nx, ny, nz))

    def trace_rays(self, scene: Scene, rays: np.ndarray, tmax: float = 1e4) -> np.ndarray:
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        hits = np.zeros(len(rays), HIT_DTYPE)
        self._check(self.L.vrt_trace_rays(self.h, scene.handle, _ptr(rays), len(rays), tmax,
                                          _ptr(hits)))
        return hits

    def trace_rays_device(self, scene: Scene, d_rays: int, n: int, d_hits: int,
                          tmax: float = 1e4) -> float:
        ms = C.c_float()
        self._check(self.L.vrt_trace_rays_device(self.h, scene.handle, C.c_void_p(d_rays), n,
                                                 tmax, C.c_void_p(d_hits), C.byref(ms)))
        return ms.value

    def trace_rays_device_noalloc(self, scene: Scene, d_rays: int, n: int, d_hits: int,
                                  tmax: float = 1e4) -> None:
        """vrt_trace_rays_device_noalloc: enqueued on the context's stream,
        returns without waiting."""
        self._check(self.L.vrt_trace_rays_device_noalloc(self.h, scene.handle, C.c_void_p(d_rays), n,
                                                         tmax, C.c_void_p(d_hits)))

    def copy_device(self, dst: "DeviceArray", src: "DeviceArray", nbytes: int = 0):
        """Device-to-device copy (vrt_copy), waits for it."""
        n = nbytes or min(dst.nbytes, src.nbytes)
        self._check(self.L.vrt_copy(self.h, C.c_void_p(dst.ptr), C.c_void_p(src.ptr), n, 3))

    def device_array(self, nbytes: int) -> "DeviceArray":
        """A device buffer owned by this context (vrt_device_alloc)."""
        p = C.c_void_p()
        self._check(self.L.vrt_device_alloc(self.h, max(int(nbytes), 1), C.byref(p)))
        return DeviceArray(self, p.value, int(nbytes))

    def render(self, views, outs=None):
        """views: list of View; outs: list of dicts of numpy destinations."""
        nv = len(views)
        varr = (View * nv)(*views)
        farr = None
        keep = []
        if outs is not None:
            farr = (Frame * nv)()
            for k, o in enumerate(outs):
                for name in ("radiance", "albedo", "gbuf", "accum", "denoised", "hdr", "ldr"):
                    a = o.get(name) if o else None
                    if a is not None:
                        keep.append(a)
                        setattr(farr[k], name, a.ctypes.data)
        self._check(self.L.vrt_render(self.h, varr, nv, farr))
        return farr

    def release_view(self, view_id: int):
        self._check(self.L.vrt_view_release(self.h, view_id))

    def host_array(self, shape, dtype=np.float32) -> np.ndarray:
        """A zeroed numpy array in page-locked host memory (vrt_host_alloc), for
        frame read-backs at full link speed; freed when the array and its views
        are collected."""
        dt = np.dtype(dtype)
        n = int(np.prod(shape)) * dt.itemsize
        p = C.c_void_p()
        self._check(self.L.vrt_host_alloc(self.h, max(n, 1), C.byref(p)))
        buf = (C.c_uint8 * max(n, 1)).from_address(p.value)
        a = np.frombuffer(buf, np.uint8, n).view(dt).reshape(shape)
        a[...] = 0
        weakref.finalize(buf, self.L.vrt_host_free, None, p.value)
        return a

    def temporal_pass(self, u: Uniforms, radiance, gbuf_cur, accum_prev, gbuf_prev, fmt=0):
        w, h = int(u.params1[0]), int(u.params1[1])
        out = np.zeros((h, w, 4), np.uint16 if fmt == 0 else np.float32)
        self._check(self.L.vrt_temporal_pass(self.h, C.byref(u), _ptr(radiance), _ptr(gbuf_cur),
                                             _ptr(accum_prev), _ptr(gbuf_prev), fmt, _ptr(out)))
        return out

    def atrous_pass(self, u: Uniforms, src, gbuf, step: int, src_fmt: int = 0):
        w, h = int(u.params1[0]), int(u.params1[1])
        out = np.zeros((h, w, 4), np.uint16)
        self._check(self.L.vrt_atrous_pass(self.h, C.byref(u), _ptr(src), src_fmt, _ptr(gbuf),
                                           step, _ptr(out)))
        return out

    def tuning(self) -> Tuning:
        t = Tuning()
        self._check(self.L.vrt_tuning_get(self.h, C.byref(t)))
        return t

    def set_tuning(self, **kw) -> Tuning:
        t = self.tuning()
        for k, val in kw.items():
            setattr(t, k, int(val))
        self._check(self.L.vrt_tuning_set(self.h, C.byref(t)))
        return t

    def autotune(self, scene: Scene, nrays: int = 1 << 22, u: Uniforms = None) -> Tuning:
        t = Tuning()
        self._check(self.L.vrt_autotune(self.h, scene.handle, nrays,
                                        None if u is None else C.byref(u), C.byref(t)))
        return t

    def autotune_batch(self, views, modes=(0, 1, 2)) -> Tuning:
        """Batched-form launch shapes (batch_mode among `modes`, pt_tile,
        batch_blocks) timed on `views` rendered as one call; the views'
        histories are untouched."""
        nv = len(views)
        t = Tuning()
        mask = sum(1 << m for m in modes)
        self._check(self.L.vrt_autotune_batch(self.h, (View * nv)(*views), nv, mask, C.byref(t)))
        return t

    def stats(self) -> dict:
        s = Stats()
        self._check(self.L.vrt_stats_get(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in Stats._fields_}

    def reset_stats(self):
        self._check(self.L.vrt_stats_reset(self.h))

    # ---- serving (continuous batching) ----
    def serve_configure(self, streams: int = 4, max_batch_views: int = 8,
                        max_batch_pixels: int = 4 << 20):
        cfg = ServeConfig(streams, max_batch_views, max_batch_pixels)
        self._check(self.L.vrt_serve_configure(self.h, C.byref(cfg)))

    def submit(self, view: View) -> int:
        req = Request(view, 0, 0)
        t = C.c_uint64()
        self._check(self.L.vrt_submit(self.h, C.byref(req), C.byref(t)))
        return t.value

    def _take(self, rc, fr, tm, hdr):
        if rc == -5:
            return None
        self._check(rc)
        return {"hdr": hdr, "rays": fr.rays, "brick_steps": fr.brick_steps,
                "voxel_steps": fr.voxel_steps, "timing": tm.as_dict()}

    def poll(self, ticket: int, hdr: np.ndarray = None):
        """None while the request is queued or running; else a dict with the
        hdr array (when given), counters and timing.  The ticket is retired."""
        fr, tm = Frame(), Timing()
        if hdr is not None:
            fr.hdr = _ptr(hdr).value
        return self._take(self.L.vrt_poll(self.h, ticket, C.byref(fr), C.byref(tm)), fr, tm, hdr)

    def wait(self, ticket: int, hdr: np.ndarray = None, timeout_ms: int = 0):
        fr, tm = Frame(), Timing()
        if hdr is not None:
            fr.hdr = _ptr(hdr).value
        return self._take(self.L.vrt_wait(self.h, ticket, timeout_ms, C.byref(fr), C.byref(tm)), fr, tm, hdr)

    def cancel(self, ticket: int):
        self._check(self.L.vrt_cancel(self.h, ticket))

    def serve_stats(self) -> dict:
        s = ServeStats()
        self._check(self.L.vrt_serve_stats_get(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in ServeStats._fields_}


class DeviceArray:
    """Device memory from vrt_device_alloc: ``ptr`` for the device-resident
    entry points, ``upload``/``download`` through vrt_copy (synchronous)."""

    def __init__(self, ctx: Context, ptr: int, nbytes: int):
        self.ctx, self.ptr, self.nbytes = ctx, ptr, nbytes

    def upload(self, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError("array larger than the device buffer")
        self.ctx._check(self.ctx.L.vrt_copy(self.ctx.h, C.c_void_p(self.ptr), _ptr(a), a.nbytes, 1))
        return self

    def download(self, a: np.ndarray) -> np.ndarray:
        if not a.flags["C_CONTIGUOUS"] or a.nbytes > self.nbytes:
            raise ValueError("destination must be C-contiguous and fit the device buffer")
        self.ctx._check(self.ctx.L.vrt_copy(self.ctx.h, _ptr(a), C.c_void_p(self.ptr), a.nbytes, 2))
        return a

    def free(self):
        if self.ptr and self.ctx.h:
            self.ctx._check(self.ctx.L.vrt_device_free(self.ctx.h, C.c_void_p(self.ptr)))
        self.ptr = 0
