"""TEST INFRASTRUCTURE ONLY: ctypes binding of the CPU parity oracle.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg import this package, and only as the checker (or the timed CPU baseline).
The product (``voxelrt_amd``) never imports it.

The C sources (``vrt_oracle.cpp``) restate VoxelRT's WGSL base frame path
(reference: src/shaders/voxel.wgsl, common.wgsl, pathtrace.wgsl:248-330,
temporal.wgsl, atrous.wgsl; src/camera.js, src/math.js, src/scene.js:411-433).
``vb_*`` functions wrap the reference's own VoxelBench traversal
(VoxelBench/pipeline/core.cpp:23-130), compiled from the reference sources
into ``oracle/_ref`` where the reference tree exists.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libvrt_oracle.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libvoxelbench_core.so")
_lib = None
_ref = None

VOXEL_SIZE = 0.0625
FACE_NONE = 6
FACE_MISS = 255
VOXEL_MISS = 0xFFFFFFFF

HIT_DTYPE = np.dtype([("t", "<f4"), ("voxel", "<u4"), ("brick_steps", "<u2"),
                      ("voxel_steps", "<u2"), ("face", "u1"), ("_pad", "u1", (3,))])
assert HIT_DTYPE.itemsize == 16


class Grid(C.Structure):
    _fields_ = [("vox", C.c_void_p), ("bricks", C.c_void_p), ("masks", C.c_void_p),
                ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("voxel_size", C.c_float)]


class Counters(C.Structure):
    _fields_ = [("rays", C.c_uint64), ("brick_steps", C.c_uint64),
                ("voxel_steps", C.c_uint64), ("hits", C.c_uint64)]


class Uniforms(C.Structure):
    """304-byte Uniforms block (common.wgsl:13-35)."""
    _fields_ = [("inv_view_proj", C.c_float * 16), ("prev_view_proj", C.c_float * 16),
                ("cam_pos", C.c_float * 4), ("prev_cam_pos", C.c_float * 4),
                ("sun_dir", C.c_float * 4), ("sun_radiance", C.c_float * 4),
                ("params0", C.c_uint32 * 4), ("params1", C.c_float * 4),
                ("params2", C.c_uint32 * 4), ("params3", C.c_float * 4),
                ("params4", C.c_float * 4), ("params5", C.c_float * 4),
                ("params6", C.c_float * 4)]


assert C.sizeof(Uniforms) == 304


def build(quiet: bool = True) -> None:
    """Compile the oracle (and oracle/_ref where the reference tree exists)."""
    targets = ["all"]
    if os.path.isfile("/root/reference/VoxelBench/pipeline/core.cpp"):
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", _HERE] + targets, check=True,
                   stdout=subprocess.DEVNULL if quiet else None)


def lib():
    global _lib
    if _lib is None:
        if not os.path.isfile(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


def ref_lib():
    """The reference's VoxelBench core (None when it was not built)."""
    global _ref
    if _ref is None and os.path.isfile(_REF_PATH):
        _ref = C.CDLL(_REF_PATH)
        _ref.vb_set_grid.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        _ref.vb_set_grid.restype = C.c_int
        _ref.vb_trace_rays.argtypes = [C.c_void_p, C.c_long, C.c_float, C.c_int, C.c_int,
                                       C.c_void_p, C.c_void_p]
        _ref.vb_trace_rays.restype = None
    return _ref


def _declare(L):
    P = C.c_void_p
    L.orc_build_bricks.argtypes = [P, C.c_int, C.c_int, C.c_int, P, P]
    L.orc_build_bricks.restype = C.c_int
    L.orc_trace.argtypes = [C.POINTER(Grid), P, P, C.c_float, P, P, P]
    L.orc_trace_rays.argtypes = [C.POINTER(Grid), P, C.c_uint64, C.c_float, P, C.c_int]
    L.orc_pcg_hash.argtypes = [C.c_uint32]
    L.orc_pcg_hash.restype = C.c_uint32
    L.orc_rand_stream.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, P, P]
    for f in ("orc_sin", "orc_cos", "orc_exp2", "orc_log2", "orc_exp", "orc_exp_w"):
        getattr(L, f).argtypes = [C.c_float]
        getattr(L, f).restype = C.c_float
    L.orc_sub_texel.argtypes = [C.c_float, P, P]
    L.orc_exp_w_table.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, P]
    L.orc_pow.argtypes = [C.c_float, C.c_float]
    L.orc_pow.restype = C.c_float
    L.orc_f32_to_f16.argtypes = [C.c_float]
    L.orc_f32_to_f16.restype = C.c_uint16
    L.orc_f16_to_f32.argtypes = [C.c_uint16]
    L.orc_f16_to_f32.restype = C.c_float
    L.orc_camera_matrices.argtypes = [P, C.c_double, C.c_double, C.c_double, C.c_double,
                                      C.c_double, C.c_double, P, P]
    L.orc_camera_ray.argtypes = [C.POINTER(Uniforms), C.c_float, C.c_float, P]
    L.orc_sky_color.argtypes = [C.POINTER(Uniforms), P, C.c_int, P]
    L.orc_sample_sun_dir.argtypes = [C.POINTER(Uniforms), P, P]
    L.orc_cosine_hemisphere.argtypes = [P, P, P]
    L.orc_pathtrace.argtypes = [C.POINTER(Grid), C.POINTER(Uniforms), C.c_float, C.c_uint32,
                                C.c_int, C.c_int, P, P, P, C.POINTER(Counters), C.c_int]
    L.orc_pathtrace_pixels.argtypes = [C.POINTER(Grid), C.POINTER(Uniforms), C.c_float,
                                       C.c_uint32, P, C.c_uint64, P, P, P,
                                       C.POINTER(Counters), C.c_int]
    L.orc_temporal.argtypes = [C.POINTER(Uniforms), P, P, P, P, C.c_uint32, P, C.c_int]
    L.orc_atrous.argtypes = [C.POINTER(Uniforms), P, C.c_uint32, P, C.c_int, P, C.c_int]
    L.orc_hdr.argtypes = [C.c_int, C.c_int, P, C.c_uint32, P, P]
    L.orc_present.argtypes = [C.c_int, C.c_int, P, C.c_uint32, P, C.c_float, P]


def exp_w(x: float) -> float:
    """e^x as the a-trous weights compute it (atrous.wgsl:74-75)."""
    return lib().orc_exp_w(x)


def exp_w_table(start: int, stride: int, n: int) -> np.ndarray:
    """exp_w over the float bit patterns start + i*stride (mod 2^32)."""
    out = np.empty(n, np.float32)
    lib().orc_exp_w_table(start, stride, n, _ptr(out))
    return out


def sub_texel(tc: float):
    """(floor, fraction) of a bilinear texel coordinate at 8 fractional bits."""
    fl, fr = C.c_float(), C.c_float()
    lib().orc_sub_texel(tc, C.byref(fl), C.byref(fr))
    return fl.value, fr.value


def _ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------------------
# grid
# ---------------------------------------------------------------------------
class OracleGrid:
    """Dense packed-u32 grid (x fastest) plus its oracle-built bricks/masks."""

    def __init__(self, vox: np.ndarray, voxel_size: float = VOXEL_SIZE):
        assert vox.ndim == 3 and vox.dtype == np.uint32
        nz, ny, nx = vox.shape
        self.vox = np.ascontiguousarray(vox)
        self.dims = (nx, ny, nz)
        self.bricks, self.masks = build_bricks(self.vox)
        self.voxel_size = voxel_size
        self.c = Grid(self.vox.ctypes.data, self.bricks.ctypes.data, self.masks.ctypes.data,
                      nx, ny, nz, voxel_size)


def build_bricks(vox: np.ndarray):
    nz, ny, nx = vox.shape
    nb = (nx // 8) * (ny // 8) * (nz // 8)
    bricks = np.zeros(nb, np.uint32)
    masks = np.zeros(nb * 16, np.uint32)
    rc = lib().orc_build_bricks(_ptr(vox), nx, ny, nz, _ptr(bricks), _ptr(masks))
    if rc != 0:
        raise ValueError("grid dims must be positive multiples of 8")
    return bricks, masks


def trace_rays(grid: OracleGrid, rays: np.ndarray, tmax: float = 1e4, threads: int = 0):
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
    hits = np.zeros(len(rays), HIT_DTYPE)
    lib().orc_trace_rays(C.byref(grid.c), _ptr(rays), len(rays), tmax, _ptr(hits), threads)
    return hits


def trace_one(grid: OracleGrid, ro, rd, tmax: float = 1e4):
    ro = np.asarray(ro, np.float32)
    rd = np.asarray(rd, np.float32)
    hit = np.zeros(1, HIT_DTYPE)
    n = np.zeros(3, np.float32)
    mat = np.zeros(1, np.uint32)
    lib().orc_trace(C.byref(grid.c), _ptr(ro), _ptr(rd), tmax, _ptr(hit), _ptr(n), _ptr(mat))
    return hit[0], n, int(mat[0])


# ---------------------------------------------------------------------------
# RNG / math
# ---------------------------------------------------------------------------
def pcg_hash(v: int) -> int:
    return lib().orc_pcg_hash(v & 0xFFFFFFFF)


def rand_stream(px: int, py: int, frame: int, n: int):
    states = np.zeros(n, np.uint32)
    vals = np.zeros(n, np.float32)
    lib().orc_rand_stream(px, py, frame, n, _ptr(states), _ptr(vals))
    return states, vals


def f32_to_f16(x: float) -> int:
    return lib().orc_f32_to_f16(float(x))


def f16_to_f32(h: int) -> float:
    return lib().orc_f16_to_f32(h)


def camera_matrices(pos, yaw, pitch, fov_y=70 * np.pi / 180, aspect=16 / 9, near=0.05, far=400.0):
    p = np.asarray(pos, np.float64)
    vp = np.zeros(16, np.float32)
    ivp = np.zeros(16, np.float32)
    lib().orc_camera_matrices(_ptr(p), yaw, pitch, fov_y, aspect, near, far, _ptr(vp), _ptr(ivp))
    return vp, ivp


def make_uniforms(pos, yaw, pitch, width, height, frame=0, bounces=2, flags=0,
                  prev_view_proj=None, prev_cam_pos=None, sun_azimuth=0.9,
                  sun_elevation=0.85, sun_cos_radius=None, max_history=64.0,
                  sun_radiance=(5.0, 4.5, 3.8, 1.0)) -> Uniforms:
    """Oracle-side restatement of main.js:399-426 for the base preset."""
    u = Uniforms()
    vp, ivp = camera_matrices(pos, yaw, pitch, 70 * np.pi / 180, width / height, 0.05, 400.0)
    u.inv_view_proj[:] = ivp.tolist()
    u.prev_view_proj[:] = (vp if prev_view_proj is None else np.asarray(prev_view_proj, np.float32)).tolist()
    u.cam_pos[:] = [float(np.float32(p)) for p in pos] + [0.0]
    pp = pos if prev_cam_pos is None else prev_cam_pos
    u.prev_cam_pos[:] = [float(np.float32(p)) for p in pp] + [0.0]
    d = np.array([np.cos(sun_elevation) * np.cos(sun_azimuth), np.sin(sun_elevation),
                  np.cos(sun_elevation) * np.sin(sun_azimuth)])
    d = d / np.sqrt((d * d).sum())
    cr = np.cos(0.03) if sun_cos_radius is None else sun_cos_radius
    u.sun_dir[:] = [float(np.float32(x)) for x in d] + [float(np.float32(cr))]
    u.sun_radiance[:] = list(sun_radiance)
    u.params0[:] = [frame & 0xFFFFFFFF, bounces, flags, 0]
    u.params1[:] = [float(width), float(height), 1.0, 0.0]
    u.params4[:] = [30.0, max_history, 24.0, 0.0]
    return u, vp


def camera_ray(u: Uniforms, uvx: float, uvy: float):
    out = np.zeros(3, np.float32)
    lib().orc_camera_ray(C.byref(u), uvx, uvy, _ptr(out))
    return out


def sky_color(u: Uniforms, rd, disc: bool):
    rd = np.asarray(rd, np.float32)
    out = np.zeros(3, np.float32)
    lib().orc_sky_color(C.byref(u), _ptr(rd), int(disc), _ptr(out))
    return out


def sample_sun_dir(u: Uniforms, state: int):
    s = np.array([state], np.uint32)
    out = np.zeros(3, np.float32)
    lib().orc_sample_sun_dir(C.byref(u), _ptr(s), _ptr(out))
    return out, int(s[0])


def cosine_hemisphere(n, state: int):
    n = np.asarray(n, np.float32)
    s = np.array([state], np.uint32)
    out = np.zeros(3, np.float32)
    lib().orc_cosine_hemisphere(_ptr(n), _ptr(s), _ptr(out))
    return out, int(s[0])


# ---------------------------------------------------------------------------
# frame passes
# ---------------------------------------------------------------------------
def pathtrace(grid: OracleGrid, u: Uniforms, fclamp: float = 48.0, spp: int = 1,
              rows=None, threads: int = 0):
    w, h = int(u.params1[0]), int(u.params1[1])
    rad = np.zeros((h, w, 4), np.uint16)
    alb = np.zeros((h, w, 4), np.uint8)
    gb = np.zeros((h, w, 4), np.float32)
    cnt = Counters()
    y0, y1 = (0, h) if rows is None else rows
    lib().orc_pathtrace(C.byref(grid.c), C.byref(u), fclamp, spp, y0, y1, _ptr(rad), _ptr(alb),
                        _ptr(gb), C.byref(cnt), threads)
    return rad, alb, gb, cnt


def pathtrace_pixels(grid: OracleGrid, u: Uniforms, pix: np.ndarray, fclamp: float = 48.0,
                     spp: int = 1, threads: int = 0):
    w, h = int(u.params1[0]), int(u.params1[1])
    pix = np.ascontiguousarray(pix, np.uint32)
    rad = np.zeros((h, w, 4), np.uint16)
    alb = np.zeros((h, w, 4), np.uint8)
    gb = np.zeros((h, w, 4), np.float32)
    cnt = Counters()
    lib().orc_pathtrace_pixels(C.byref(grid.c), C.byref(u), fclamp, spp, _ptr(pix), len(pix),
                               _ptr(rad), _ptr(alb), _ptr(gb), C.byref(cnt), threads)
    return rad, alb, gb, cnt


def temporal(u: Uniforms, radiance, gbuf_cur, accum_prev, gbuf_prev, fmt: int = 0,
             threads: int = 0):
    w, h = int(u.params1[0]), int(u.params1[1])
    out = np.zeros((h, w, 4), np.uint16 if fmt == 0 else np.float32)
    lib().orc_temporal(C.byref(u), _ptr(radiance), _ptr(gbuf_cur), _ptr(accum_prev),
                       _ptr(gbuf_prev), fmt, _ptr(out), threads)
    return out


def atrous(u: Uniforms, src, gbuf, step: int, src_fmt: int = 0, threads: int = 0):
    w, h = int(u.params1[0]), int(u.params1[1])
    out = np.zeros((h, w, 4), np.uint16)
    lib().orc_atrous(C.byref(u), _ptr(src), src_fmt, _ptr(gbuf), step, _ptr(out), threads)
    return out


def hdr(illum, illum_fmt, albedo):
    h, w = albedo.shape[:2]
    out = np.zeros((h, w, 3), np.float32)
    lib().orc_hdr(w, h, _ptr(illum), illum_fmt, _ptr(albedo), _ptr(out))
    return out


def present(illum, illum_fmt, albedo, exposure=1.0):
    """present.wgsl:40-55 -> (h, w, 4) uint8 (alpha 255)."""
    h, w = albedo.shape[:2]
    out = np.zeros((h, w, 4), np.uint8)
    lib().orc_present(w, h, _ptr(illum), illum_fmt, _ptr(albedo), C.c_float(exposure), _ptr(out))
    return out


# ---------------------------------------------------------------------------
# VoxelBench (reference sources) traversal
# ---------------------------------------------------------------------------
_vb_grid_key = None


def vb_trace(occ: np.ndarray, rays_vox: np.ndarray, naive: bool, tmax: float = 1e9,
             threads: int = 0):
    """Trace rays (voxel units) with VoxelBench's trace/traceNaive on an
    occupancy grid shaped (nz, ny, nx).  Returns (vox_xyz_face[n,4], t[n])."""
    global _vb_grid_key
    R = ref_lib()
    if R is None:
        raise RuntimeError("oracle/_ref/libvoxelbench_core.so not built")
    # the VoxelBench grid + brickmap is rebuilt only for a new occupancy (keyed
    # by content: a freed array's id and address are reused by the next one)
    occ8 = np.ascontiguousarray(occ != 0, np.uint8)
    key = (occ8.shape, hashlib.blake2b(occ8.tobytes(), digest_size=16).digest())
    if key != _vb_grid_key:
        nz, ny, nx = occ8.shape
        if R.vb_set_grid(_ptr(occ8), nx, ny, nz) != 0:
            raise ValueError("VoxelBench dims must be multiples of 4")
        _vb_grid_key = key
    rays_vox = np.ascontiguousarray(rays_vox, np.float32).reshape(-1, 6)
    out = np.zeros((len(rays_vox), 4), np.int32)
    t = np.zeros(len(rays_vox), np.float32)
    R.vb_trace_rays(_ptr(rays_vox), len(rays_vox), tmax, int(naive), threads, _ptr(out), _ptr(t))
    return out, t
