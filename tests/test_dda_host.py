"""The library's DDA variants run on the CPU (a host build of the device
headers, tests/dda_host.cpp) against the oracle: nested loops and the
single-loop form on LDS and on global memory (step by step, with the bulk
skip of empty brick rows, its wave-batched form, and the step_ev stepping
with caller-side level switches of the step-level refill path tracer) must
return the
same hit voxel, face, both step counts and the bits of t for every ray.  This
checks the variants' logic without a GPU (the GPU tests check the same on
the device).  Sparse scenes (isolated blocks, a floor, an empty grid, grids
whose padded bitmap words cover several rows) make the skip take long runs;
a test checks that it does (fewer loop iterations) so that the equality
checks are not vacuous."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle
import voxelrt_amd as v
from voxelrt_amd import configs as cf
from tests.sparse import sparse_rays as _sparse_rays, sparse_scenes as _sparse_scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
VARIANTS = {"nested": 0, "fast": 3, "fast_skip": 4, "fast_batched_skip": 5, "fast_global": 6,
            "fast_global_batched": 7, "fast_global_km": 8, "fast_km": 9, "step_ev_lds": 10, "step_ev_global": 11,
            "step_ev_peek_lds": 12, "step_ev_peek_global": 13}


@pytest.fixture(scope="module")
def dda(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("dda") / "dda_host.so")
    # -mfma: the explicit fused multiply-adds (exp2_r, exp_w) as one
    # instruction each instead of libm calls (same correctly rounded result)
    cmd = ["/opt/rocm/bin/hipcc", "-x", "hip", "--cuda-host-only", "-O2", "-std=c++17", "-ffp-contract=off",
           "-mfma", "-fPIC", "-shared", os.path.join(ROOT, "tests", "dda_host.cpp"), "-o", out]
    if not os.path.exists(cmd[0]):
        pytest.skip("hipcc not available")
    subprocess.run(cmd, check=True, capture_output=True)
    L = C.CDLL(out)
    L.dda_host_trace.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_uint64, C.c_float,
                                 C.c_int, C.c_void_p, C.c_void_p]
    L.dda_host_trace.restype = C.c_int

    def run(vox, rays, variant, tmax=1e4, iters=None):
        nz, ny, nx = vox.shape
        vox = np.ascontiguousarray(vox, np.uint32)
        rays = np.ascontiguousarray(rays, np.float32)
        hits = np.zeros(len(rays), v.HIT_DTYPE)
        it = np.zeros(1, np.uint64)
        rc = L.dda_host_trace(vox.ctypes.data, nx, ny, nz, rays.ctypes.data, len(rays), tmax, variant,
                              hits.ctypes.data, it.ctypes.data)
        assert rc == 0
        if iters is not None:
            iters.append(int(it[0]))
        return hits
    run.lib = L
    return run


def _equal(got, want, name):
    for f in ("voxel", "face", "brick_steps", "voxel_steps"):
        bad = np.flatnonzero(got[f] != want[f])
        assert len(bad) == 0, f"{name} {f}: {len(bad)} mismatches, first {bad[:5]}"
    tb = np.flatnonzero(got["t"].view(np.uint32) != want["t"].view(np.uint32))
    assert len(tb) == 0, f"{name} t: {len(tb)} mismatches, first {tb[:5]}"


def _edge_rays(n):
    """Axis-aligned and diagonal rays from grid faces, edges and corners,
    including zero direction components."""
    ext = n * v.VOXEL_SIZE
    rays = []
    for o in [(0, 0, 0), (ext, ext, ext), (ext / 2, 0, ext / 2), (0, ext / 3, ext / 7), (ext, ext / 2, 0),
              (ext / 2, ext / 2, ext / 2), (-1.0, ext / 2, ext / 2), (ext / 4, ext + 2.0, ext / 4)]:
        for d in [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 1, 0), (1, -1, 1),
                  (-1, -1, -1), (0.6, 0.0, -0.8), (1e-9, 1.0, 0.0), (0.0, 0.0, 0.0)]:
            dd = np.asarray(d, np.float64)
            nrm = np.linalg.norm(dd)
            rays.append(list(o) + list(dd / nrm if nrm > 0 else dd))
    return np.asarray(rays, np.float32)


@pytest.mark.parametrize("variant", list(VARIANTS))
@pytest.mark.parametrize("shape,p", [((64, 64, 64), 0.05), ((64, 64, 64), 0.25), ((64, 64, 64), 0.6),
                                     ((32, 48, 64), 0.3), ((16, 8, 24), 1.0)])
def test_variants_match_oracle(dda, variant, shape, p):
    nz, ny, nx = shape
    vox = v.generate_scene(v.SCENE_RANDOM_BRICKS, n=nx, ny=ny, nz=nz, seed=int(p * 100) + nx, p_brick=p,
                           p_voxel=0.5)
    n = max(shape)
    rays = np.concatenate([cf.traverse_rays(20000, n=n, seed=3), _edge_rays(n)])
    # scale origins to the (possibly non-cubic) box
    rays[:, :3] *= np.array([nx, ny, nz], np.float32) / n
    want = oracle.trace_rays(oracle.OracleGrid(vox), rays)
    got = dda(vox, rays, VARIANTS[variant])
    _equal(got, want, f"{variant} {shape} p={p}")


@pytest.mark.parametrize("variant", list(VARIANTS))
def test_variants_on_world_scene_with_short_tmax(dda, variant):
    vox = v.generate_scene(**cf.world_desc(64, 4))
    rays = cf.traverse_rays(20000, n=64, seed=8)
    og = oracle.OracleGrid(vox)
    for tmax in (1e4, 0.75):
        _equal(dda(vox, rays, VARIANTS[variant], tmax), oracle.trace_rays(og, rays, tmax=tmax),
               f"{variant} tmax={tmax}")


@pytest.mark.parametrize("variant", ["fast", "fast_skip", "fast_batched_skip", "fast_global"])
@pytest.mark.parametrize("name", [s[0] for s in _sparse_scenes()])
def test_empty_row_skip_matches_oracle(dda, variant, name):
    vox = dict(_sparse_scenes())[name]
    rays = np.concatenate([_sparse_rays(vox, 6000, 11), _edge_rays(max(vox.shape))])
    og = oracle.OracleGrid(vox)
    for tmax in (1e4, 1.3):
        _equal(dda(vox, rays, VARIANTS[variant], tmax), oracle.trace_rays(og, rays, tmax=tmax),
               f"{variant} {name} tmax={tmax}")


def test_empty_row_skip_takes_runs(dda):
    """The skip really replaces coarse iterations on sparse scenes (and never
    adds any; none on a full grid, where no row is empty)."""
    most = {"blocks": 0.7, "floor": 0.8, "empty": 0.7, "wide": 0.97}
    for name, vox in _sparse_scenes():
        rays = _sparse_rays(vox, 3000, 5)
        a, b = [], []
        dda(vox, rays, VARIANTS["fast"], iters=a)
        dda(vox, rays, VARIANTS["fast_skip"], iters=b)
        assert b[0] <= a[0], (name, a, b)
        if name == "full":
            assert b[0] == a[0], (name, a, b)
        if name in most:
            assert b[0] < most[name] * a[0], (name, a, b)


def test_exp_w_device_equals_oracle(dda):
    """The a-trous weights' exponential (vrt_device.h exp_w, built here from
    the device header) equals the oracle's exp_w bit for bit on every 61st
    float bit pattern with the sign bit set (35M values: all negative
    magnitudes, -0, the negative NaNs) and on +0 and the positive NaNs."""
    import oracle
    L = dda.lib
    L.exp_w_table.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_void_p]
    n = (1 << 31) // 61
    for start, stride, cnt in ((0x80000000, 61, n), (0x7fc00000, 1, 1 << 16), (0, 1, 1)):
        got = np.empty(cnt, np.float32)
        L.exp_w_table(start, stride, cnt, got.ctypes.data_as(C.c_void_p))
        want = oracle.exp_w_table(start, stride, cnt)
        bad = np.flatnonzero(got.view(np.uint32) != want.view(np.uint32))
        assert len(bad) == 0, (hex(start), bad[:5], got[bad[:5]], want[bad[:5]])


BOX_VARIANTS = {"step_ev_peek_box_lds": 14, "step_ev_peek_box_global": 15}


def _box_scene(kind):
    if kind == "world":
        return v.generate_scene(**cf.world_desc(64, 4))
    vox = v.generate_scene(v.SCENE_RANDOM_BRICKS, n=64, seed=31, p_brick=0.3, p_voxel=0.5)
    if kind == "subbox":  # occupied bricks only in bricks 2..4 (x), 1..5 (y), 3..6 (z)
        keep = np.zeros_like(vox, bool)
        keep[24:56, 8:48, 16:40] = True
        vox[~keep] = 0
    elif kind == "empty":
        vox[:] = 0
    return vox


@pytest.mark.parametrize("variant", list(BOX_VARIANTS))
@pytest.mark.parametrize("kind", ["world", "subbox", "empty", "full"])
def test_box_variants_end_misses_early(dda, variant, kind):
    """The path tracer's DDA (BOX = true: rays end at the box of occupied
    bricks) against the oracle's grid walk: every record's voxel, face, t
    bits and voxel steps are equal; a hit's brick steps are equal; a miss
    takes at most as many brick steps, and fewer wherever the box is smaller
    than the grid."""
    vox = _box_scene(kind)
    rays = np.concatenate([cf.traverse_rays(20000, n=64, seed=5), _edge_rays(64)])
    want = oracle.trace_rays(oracle.OracleGrid(vox), rays)
    got = dda(vox, rays, BOX_VARIANTS[variant])
    for f in ("voxel", "face", "voxel_steps"):
        bad = np.flatnonzero(got[f] != want[f])
        assert len(bad) == 0, f"{kind} {f}: {len(bad)} mismatches, first {bad[:5]}"
    assert np.array_equal(got["t"].view(np.uint32), want["t"].view(np.uint32))
    hit = want["t"] >= 0
    assert np.array_equal(got["brick_steps"][hit], want["brick_steps"][hit])
    assert np.all(got["brick_steps"][~hit] <= want["brick_steps"][~hit])
    saved = int(want["brick_steps"][~hit].sum()) - int(got["brick_steps"][~hit].sum())
    if kind == "full":
        assert saved == 0  # the box is the grid
    else:
        assert saved > 0 and (~hit).sum() > 0
    if kind == "empty":
        assert not hit.any() and np.all(got["brick_steps"] == 0)


FBOX_VARIANTS = {"step_ev_peek_box_fbox_lds": 16, "step_ev_peek_box_fbox_global": 17}


@pytest.mark.parametrize("variant", list(FBOX_VARIANTS))
@pytest.mark.parametrize("kind", ["world", "subbox", "full"])
def test_fbox_variants_end_fine_walks_early(dda, variant, kind):
    """The path tracer's DDA with both boxes (rays end at the box of occupied
    bricks; a fine walk ends at the entered brick's box of solid voxels, and
    a brick whose box the ray cannot reach is left at once) against the
    oracle's grid walk: every record's voxel, face and t bits are equal, a
    hit's brick steps are equal, a miss's at most equal, and the voxel steps
    at most equal (fewer in the world scene, whose bricks leave empty
    margins; random bricks at half fill span their whole brick)."""
    vox = _box_scene(kind)
    rays = np.concatenate([cf.traverse_rays(20000, n=64, seed=5), _edge_rays(64)])
    want = oracle.trace_rays(oracle.OracleGrid(vox), rays)
    got = dda(vox, rays, FBOX_VARIANTS[variant])
    for f in ("voxel", "face"):
        bad = np.flatnonzero(got[f] != want[f])
        assert len(bad) == 0, f"{kind} {f}: {len(bad)} mismatches, first {bad[:5]}"
    assert np.array_equal(got["t"].view(np.uint32), want["t"].view(np.uint32))
    hit = want["t"] >= 0
    assert np.array_equal(got["brick_steps"][hit], want["brick_steps"][hit])
    assert np.all(got["brick_steps"][~hit] <= want["brick_steps"][~hit])
    assert np.all(got["voxel_steps"] <= want["voxel_steps"])
    if kind == "world":
        assert got["voxel_steps"].sum() < want["voxel_steps"].sum()


def test_fbox_on_sparse_bricks_skips_whole_bricks(dda):
    """Bricks holding one solid voxel in a corner: most rays that enter such a
    brick cannot reach that voxel and leave it at once (no voxel step)."""
    vox = np.zeros((64, 64, 64), np.uint32)
    rng = np.random.default_rng(2)
    for bz, by, bx in rng.integers(0, 8, (120, 3)):
        vox[bz * 8 + 7, by * 8 + 7, bx * 8 + 7] = 0xFF8080FF
    rays = cf.traverse_rays(20000, n=64, seed=6)
    want = oracle.trace_rays(oracle.OracleGrid(vox), rays)
    for variant in FBOX_VARIANTS.values():
        got = dda(vox, rays, variant)
        assert np.array_equal(got["voxel"], want["voxel"]) and np.array_equal(got["face"], want["face"])
        assert np.array_equal(got["t"].view(np.uint32), want["t"].view(np.uint32))
        assert got["voxel_steps"].sum() * 4 < want["voxel_steps"].sum()


def test_noiter_degenerate_rays_end_like_the_budgeted_walk(dda):
    """The path tracer's DDA has no iteration budget (NOITER): degenerate
    walks end by the per-axis counts instead.  On zero directions (from solid
    and from empty voxels), directions with tiny or infinite components and
    infinite origins, every record's voxel, face and t bits equal the
    oracle's budgeted walk (step counts may be fewer).  Rays with a NaN
    component (which the oracle cannot walk: its cell indices would come from
    NaN) end at setup as a miss in every variant."""
    vox = _box_scene("world")
    ext = 64 * v.VOXEL_SIZE
    rays = []
    for z, y, x in np.concatenate([np.argwhere(vox != 0)[:64], np.argwhere(vox == 0)[::5000][:64]]):
        o = (np.array([x, y, z], np.float64) + 0.5) * v.VOXEL_SIZE
        rays.append([*o, 0.0, 0.0, 0.0])
        rays.append([*o, -0.0, 0.0, -0.0])
    inf = float("inf")
    for o in ([ext / 2, ext / 2, ext / 2], [ext / 3, 0.1, ext / 4], [inf, ext / 2, ext / 2], [-inf, 1.0, 1.0]):
        for d in ([0.6, 0.0, -0.8], [0.0, -1.0, 0.0], [1e-30, 1e-30, 1.0], [1e-30, -1.0, 1e-38]):
            rays.append([*o, *d])
    for o in ([ext / 2, ext / 2, ext / 2], [ext / 3, 0.1, ext / 4]):
        rays.append([*o, inf, 0.0, 0.0])
    rays = np.asarray(rays, np.float32)
    want = oracle.trace_rays(oracle.OracleGrid(vox), rays)
    for variant in FBOX_VARIANTS.values():
        got = dda(vox, rays, variant)
        for f in ("voxel", "face"):
            bad = np.flatnonzero(got[f] != want[f])
            assert len(bad) == 0, (variant, f, bad[:5], rays[bad[:5]])
        assert np.array_equal(got["t"].view(np.uint32), want["t"].view(np.uint32))
    assert (want["t"][:128:2] >= 0).all()  # zero directions from solid voxels hit at once
    nan = float("nan")
    bad_rays = np.asarray([[nan, 1.0, 1.0, 0.0, 1.0, 0.0], [1.0, 1.0, 1.0, nan, 0.0, 0.0],
                           [1.0, 1.0, 1.0, 0.0, nan, 1.0], [1.0, nan, nan, nan, nan, nan],
                           [inf, 1.0, 1.0, inf, 0.0, 0.0], [1.0, 1.0, 1.0, nan, nan, nan]], np.float32)
    for variant in [*VARIANTS.values(), *BOX_VARIANTS.values(), *FBOX_VARIANTS.values()]:
        got = dda(vox, bad_rays, variant)
        assert (got["t"] == -1.0).all() and (got["voxel"] == v.VOXEL_MISS).all(), variant
        assert (got["brick_steps"] == 0).all() and (got["voxel_steps"] == 0).all(), variant
