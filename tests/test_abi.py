"""The C-ABI library without a GPU: it loads, exports every entry point that
include/voxelrt.h declares, reports the missing device cleanly, and its host
helpers (camera, uniforms, scene generators) behave as specified."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle
import voxelrt_amd as v
from voxelrt_amd import configs as cf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for hdr in ("voxelrt.h",):
        src = open(os.path.join(ROOT, "include", hdr)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(vrt_[a-z0-9_]+)\s*\(", src))
    return names


def test_library_exports_every_declared_symbol():
    L = v.load()
    names = _declared()
    assert len(names) >= 20
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_struct_layouts():
    assert C.sizeof(v.Uniforms) == 304  # common.wgsl:13-35
    assert v.HIT_DTYPE.itemsize == 16
    assert C.sizeof(v.lib.View) == 8 + 304 + 4 * 4
    assert C.sizeof(v.lib.Request) == C.sizeof(v.lib.View) + 8
    assert C.sizeof(v.lib.ServeConfig) == 16
    assert C.sizeof(v.lib.Timing) == 5 * 8 + 2 * 4 + 8
    assert C.sizeof(v.lib.ServeStats) == 10 * 8


def test_serving_entry_points_reject_null_context():
    L = v.load()
    t = C.c_uint64(7)
    assert L.vrt_submit(None, None, C.byref(t)) == -1
    assert L.vrt_poll(None, 1, None, None) == -1
    assert L.vrt_wait(None, 1, 1, None, None) == -1
    assert L.vrt_cancel(None, 1) == -1
    assert L.vrt_serve_stats_get(None, None) == -1
    assert L.vrt_serve_configure(None, None) == -1


def test_device_entry_points_reject_null_context():
    L = v.load()
    p = C.c_void_p()
    assert L.vrt_trace_rays_device_noalloc(None, None, None, 0, 1.0, None) == -1
    assert L.vrt_device_alloc(None, 16, C.byref(p)) == -1
    assert L.vrt_device_free(None, None) == -1
    assert L.vrt_copy(None, None, None, 0, 1) == -1


def test_error_paths_without_device():
    L = v.load()
    n = C.c_int32(-1)
    L.vrt_device_count(C.byref(n))
    if n.value > 0:
        pytest.skip("a HIP device is visible")
    h = C.c_void_p()
    assert L.vrt_create(None, C.byref(h)) == -7  # VRT_ERR_NO_DEVICE
    assert not h.value
    assert L.vrt_scene_release(None, None) == -1
    assert L.vrt_render(None, None, 0, None) == -1
    assert L.vrt_last_error(None) == b"null context"
    with pytest.raises(v.VrtError):
        v.Context(0)


def test_camera_and_uniforms_match_oracle():
    for pos, yaw, pitch, w, h in [((8.8, 4.7, 7.0), -0.45, 0.02, 1920, 1080),
                                  ((1.0, 2.0, 3.0), 2.5, -0.3, 640, 360)]:
        vp, ivp = v.camera_matrices(pos, yaw, pitch, aspect=w / h)
        vpo, ivpo = oracle.camera_matrices(pos, yaw, pitch, aspect=w / h)
        assert np.array_equal(vp, vpo) and np.array_equal(ivp, ivpo)
        # inverse relation (double check on the math itself)
        m = vp.reshape(4, 4).T.astype(np.float64) @ ivp.reshape(4, 4).T.astype(np.float64)
        assert np.allclose(m, np.eye(4), atol=2e-4)
        u, _ = v.make_uniforms(pos, yaw, pitch, w, h, frame=5, bounces=3, flags=3)
        uo, _ = oracle.make_uniforms(pos, yaw, pitch, w, h, frame=5, bounces=3, flags=3)
        assert bytes(u) == bytes(uo)


def test_scene_generators_deterministic_and_valid():
    a = v.generate_scene(v.SCENE_WORLD, n=64, seed=3, room=16, room_h=16, boxes=2, trees=3, lanterns=2)
    b = v.generate_scene(v.SCENE_WORLD, n=64, seed=3, room=16, room_h=16, boxes=2, trees=3, lanterns=2)
    c = v.generate_scene(v.SCENE_WORLD, n=64, seed=4, room=16, room_h=16, boxes=2, trees=3, lanterns=2)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    emissive = (a >> 24) != 0
    assert emissive.any()
    r = cf.traverse_scene(0.25, n=64)
    bricks, _ = oracle.build_bricks(r)
    assert 0.15 < bricks.mean() < 0.35
    fill = (r != 0).reshape(8, 8, 8, 8, 8, 8).transpose(0, 2, 4, 1, 3, 5).reshape(-1, 512)
    occ = fill.any(1)
    assert 0.45 < fill[occ].mean() < 0.55
    assert np.all((r[r != 0] >> 24) == 0) and np.all(r[r != 0] & 1)
    with pytest.raises(v.VrtError):
        v.generate_scene(v.SCENE_WORLD, n=60)


def test_frame128_scene_matches_its_description():
    vox = cf.frame_128_scene()
    assert vox.shape == (128, 128, 128)
    # terrain heights 8..40 voxels (columns away from the room)
    col = vox[:, :, 5] != 0  # x = 5, shape (z, y)
    tops = np.array([np.max(np.nonzero(col[z])[0]) for z in range(128)])
    assert tops.min() >= 8 and tops.max() <= 40
    # one emissive panel of intensity byte 200
    e = (vox >> 24) & 0xFF
    assert set(np.unique(e[e != 0]).tolist()) == {200}


def test_voxelrt_default_scene_layout():
    """Restatement of generateScene(256, 'default') (scene.js:232-409):
    layout facts readable from the source."""
    vox = v.generate_scene(v.SCENE_VOXELRT, n=256)
    N = 256
    roomW, roomH = 80, 51  # min(V(5), floor(0.45N)), min(V(3.2), floor(0.3N))
    cx, cz = 128, 140
    e = (vox >> 24) & 0xFF
    lights = set(int(x) for x in np.unique(vox[e != 0]))
    # ceilLight (210), the three lanterns and the warm room lamp (255)
    M = lambda r, g, b, em=0: r | (g << 8) | (b << 16) | (em << 24)
    assert lights == {M(255, 240, 220, 210), M(255, 190, 110, 255), M(90, 220, 255, 255), M(255, 90, 200, 255)}
    # red (-x) and green (+x) walls
    assert M(200, 45, 40) in set(np.unique(vox[:, :, cx - roomW // 2 - 1]).tolist())
    assert M(55, 175, 60) in set(np.unique(vox[:, :, cx + roomW // 2 + 1]).tolist())
    pos, yaw, pitch = v.scene_spawn(v.SCENE_VOXELRT, n=N)
    assert pos[0] == cx * 0.0625 and yaw == 0.0 and pitch == -0.05
    assert pos[2] == (cz - roomW // 2 - 56) * 0.0625  # z0 - V(3.5)
