"""Pins for the oracle's grid build and two-level DDA (CPU only).

Independent checks: numpy re-derivation of the brick occupancy/masks,
axis-aligned closed forms (hit voxel, face, t and step counts), a
double-precision brute-force voxel walk on small random grids, and the
reference's own VoxelBench traversal (VoxelBench/pipeline/core.cpp:23-130,
compiled from the reference sources into oracle/_ref) -- live when that
library exists, and through committed vectors otherwise.
"""
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _rand_grid(n, p, seed, shape=None):
    rng = np.random.default_rng(seed)
    shape = shape or (n, n, n)
    vox = (rng.random(shape) < p).astype(np.uint32) * rng.integers(1, 2**24, shape, dtype=np.uint32)
    return vox


# ---------------------------------------------------------------------------
# brick occupancy + masks, scene.js:411-433
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("shape", [(16, 16, 16), (8, 24, 40), (32, 16, 8)])
def test_brick_build_matches_numpy(shape):
    vox = _rand_grid(0, 0.02, 3, shape)
    bricks, masks = oracle.build_bricks(vox)
    nz, ny, nx = shape
    b = (vox != 0).reshape(nz // 8, 8, ny // 8, 8, nx // 8, 8)  # z, lz, y, ly, x, lx
    occ = b.any(axis=(1, 3, 5)).reshape(-1).astype(np.uint32)
    assert np.array_equal(bricks, occ)
    # bit = lx + 8 ly + 64 lz  ->  order (lz, ly, lx) little-endian in 16 words
    bits = b.transpose(0, 2, 4, 1, 3, 5).reshape(-1, 512).astype(np.uint64)
    words = np.zeros((bits.shape[0], 16), np.uint64)
    for k in range(512):
        words[:, k >> 5] |= bits[:, k] << np.uint64(k & 31)
    assert np.array_equal(masks.reshape(-1, 16).astype(np.uint64), words)


def test_brick_build_rejects_bad_dims():
    with pytest.raises(ValueError):
        oracle.build_bricks(np.zeros((8, 8, 12), np.uint32))


# ---------------------------------------------------------------------------
# closed forms
# ---------------------------------------------------------------------------
FACES = {(1, 0, 0): 1, (-1, 0, 0): 0, (0, 1, 0): 3, (0, -1, 0): 2, (0, 0, 1): 5, (0, 0, -1): 4}


@pytest.mark.parametrize("d", list(FACES.keys()))
def test_axis_aligned_hit(d):
    n = 32
    vox = np.zeros((n, n, n), np.uint32)
    target = np.array([20, 11, 9])
    vox[target[2], target[1], target[0]] = 0x123456
    d = np.array(d, np.float32)
    o = (target + 0.5 - d * 13.25).astype(np.float32)  # 13.25 voxels before the centre
    g = oracle.OracleGrid(vox, voxel_size=1.0)
    hit, nrm, mat = oracle.trace_one(g, o, d)
    assert hit["voxel"] == target[0] + target[1] * n + target[2] * n * n
    assert hit["face"] == FACES[tuple(int(x) for x in d)]
    assert hit["t"] == np.float32(12.75)  # to the near face, exact in binary
    assert mat == 0x123456
    assert np.array_equal(nrm, -d)


def test_empty_grid_step_counts():
    # +x ray through an empty 64^3 grid crosses all 8 bricks of its row
    n = 64
    g = oracle.OracleGrid(np.zeros((n, n, n), np.uint32), voxel_size=1.0)
    hit, _, _ = oracle.trace_one(g, (0.5, 20.5, 30.5), (1, 0, 0))
    assert hit["t"] == -1 and hit["voxel"] == oracle.VOXEL_MISS and hit["face"] == oracle.FACE_MISS
    assert hit["brick_steps"] == 8 and hit["voxel_steps"] == 0
    # a diagonal through the main diagonal of bricks
    hit, _, _ = oracle.trace_one(g, (0.25, 0.5, 0.75), (1, 1, 1))
    assert hit["t"] == -1 and hit["voxel_steps"] == 0 and 8 <= hit["brick_steps"] <= 3 * 8 + 2


def test_solid_brick_is_hit_on_first_voxel_step():
    n = 32
    vox = np.zeros((n, n, n), np.uint32)
    vox[8:16, 8:16, 8:16] = 7
    g = oracle.OracleGrid(vox, voxel_size=1.0)
    hit, _, _ = oracle.trace_one(g, (0.5, 12.5, 12.5), (1, 0, 0))
    assert hit["voxel"] == 8 + 12 * n + 12 * n * n and hit["t"] == np.float32(7.5)
    assert hit["brick_steps"] == 2 and hit["voxel_steps"] == 1


def test_world_units_scale_t():
    n = 32
    vox = np.zeros((n, n, n), np.uint32)
    vox[5, 5, 20] = 1
    g = oracle.OracleGrid(vox)  # 1/16 m voxels
    hit, _, _ = oracle.trace_one(g, (10.5 / 16, 5.5 / 16, 5.5 / 16), (1, 0, 0))
    assert hit["t"] == np.float32(9.5 / 16)


def test_tmax_cutoff():
    n = 32
    vox = np.zeros((n, n, n), np.uint32)
    vox[5, 5, 20] = 1
    g = oracle.OracleGrid(vox, voxel_size=1.0)
    assert oracle.trace_one(g, (10.5, 5.5, 5.5), (1, 0, 0), tmax=9.4)[0]["t"] == -1
    assert oracle.trace_one(g, (10.5, 5.5, 5.5), (1, 0, 0), tmax=9.6)[0]["t"] == np.float32(9.5)


# ---------------------------------------------------------------------------
# brute force: exact voxel walk in double precision
# ---------------------------------------------------------------------------
def _brute(vox, o, d):
    """Walk voxels with double-precision DDA from o (inside the grid) and
    return (voxel index, face, t, margin) where margin is the smallest gap
    between the winning and runner-up boundary distances seen (ties make the
    answer ambiguous)."""
    nz, ny, nx = vox.shape
    o = o.astype(np.float64)
    d = d.astype(np.float64)
    p = np.floor(o).astype(int)
    if vox[p[2], p[1], p[0]]:
        return None
    step = np.sign(d).astype(int)
    inv = np.where(d != 0, 1.0 / np.where(d != 0, d, 1), np.inf)
    tmax = np.where(step > 0, (p + 1 - o) * inv, np.where(step < 0, (p - o) * inv, np.inf))
    tdel = np.abs(inv)
    margin = np.inf
    while True:
        order = np.argsort(tmax)
        ax = order[0]
        margin = min(margin, tmax[order[1]] - tmax[ax])
        t = tmax[ax]
        p[ax] += step[ax]
        tmax[ax] += tdel[ax]
        if not (0 <= p[0] < nx and 0 <= p[1] < ny and 0 <= p[2] < nz):
            return (oracle.VOXEL_MISS, oracle.FACE_MISS, -1.0, margin)
        if vox[p[2], p[1], p[0]]:
            face = 2 * ax + (1 if step[ax] > 0 else 0)
            return (p[0] + p[1] * nx + p[2] * nx * ny, face, t, margin)


@pytest.mark.parametrize("seed,p", [(0, 0.02), (1, 0.1), (2, 0.3)])
def test_dda_matches_brute_force(seed, p):
    n = 24
    vox = _rand_grid(n, p, seed)
    g = oracle.OracleGrid(vox, voxel_size=1.0)
    rng = np.random.default_rng(seed + 10)
    checked = 0
    for _ in range(1500):
        o = rng.uniform(0.01, n - 0.01, 3).astype(np.float32)
        d = rng.normal(size=3)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        want = _brute(vox, o, d)
        if want is None or want[3] < 1e-3:
            continue  # origin inside a solid voxel, or an ambiguous tie
        hit, _, _ = oracle.trace_one(g, o, d, tmax=1e9)
        assert (hit["voxel"], hit["face"]) == want[:2], (o, d, hit, want)
        if want[2] >= 0:
            assert abs(float(hit["t"]) - want[2]) < 1e-3
        checked += 1
    assert checked > 1000


# ---------------------------------------------------------------------------
# the reference's VoxelBench traversal (core.cpp:23-130)
# ---------------------------------------------------------------------------
def _vb_rays(n, count, seed):
    # harness.cpp:215-217 ray distribution (origins kept off the boundary)
    rng = np.random.default_rng(seed)
    o = rng.uniform(1.0, n - 1.0, (count, 3)).astype(np.float32)
    z = rng.uniform(-1, 1, count)
    ph = rng.uniform(0, 2 * np.pi, count)
    r = np.sqrt(1 - z * z)
    d = np.stack([r * np.cos(ph), z, r * np.sin(ph)], 1).astype(np.float32)
    return np.concatenate([o, d], 1).astype(np.float32)


def _compare_with_vb(vox, rays, vb_out, vb_t):
    """Oracle (voxel_size 1) vs VoxelBench on rays whose origin voxel is empty
    (VoxelBench skips the origin voxel, core.cpp:42,114): voxel + face equal
    and |dt| <= 1e-3 voxels (bench_accel.cpp:23-30)."""
    n = vox.shape[2]
    g = oracle.OracleGrid(vox, voxel_size=1.0)
    hits = oracle.trace_rays(g, rays, 1e9)
    oi = np.floor(rays[:, :3]).astype(int)
    empty = vox[oi[:, 2], oi[:, 1], oi[:, 0]] == 0
    vh = vb_out[:, 0] >= 0
    vidx = np.where(vh, vb_out[:, 0] + vb_out[:, 1] * n + vb_out[:, 2] * n * vox.shape[1], 0xFFFFFFFF).astype(np.uint32)
    vface = np.where(vh, vb_out[:, 3], 255).astype(np.uint8)
    m = empty
    assert np.array_equal(hits["voxel"][m], vidx[m])
    assert np.array_equal(hits["face"][m], vface[m])
    both = m & vh
    assert np.max(np.abs(hits["t"][both] - vb_t[both])) <= 1e-3
    return int(m.sum())


@pytest.mark.skipif(oracle.ref_lib() is None, reason="oracle/_ref not built (no reference tree)")
@pytest.mark.parametrize("naive", [True, False])
@pytest.mark.parametrize("p", [0.01, 0.1])
def test_dda_matches_voxelbench_live(naive, p):
    n = 64
    vox = _rand_grid(n, p, 7)
    rays = _vb_rays(n, 40000, 8)
    out, t = oracle.vb_trace(vox, rays, naive=naive)
    assert _compare_with_vb(vox, rays, out, t) > 30000


def test_dda_matches_voxelbench_golden():
    """Vectors written by tests/golden/make_golden.py from the reference's
    VoxelBench trace()/traceNaive() (available without the reference tree)."""
    path = os.path.join(GOLDEN, "voxelbench_trace.npz")
    z = np.load(path)
    vox = z["vox"]
    assert _compare_with_vb(vox, z["rays"], z["naive_out"], z["naive_t"]) > 3000
    assert _compare_with_vb(vox, z["rays"], z["brick_out"], z["brick_t"]) > 3000
