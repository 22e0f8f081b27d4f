"""Sparse test scenes for the empty-row skip of the single-loop DDA (shared
by the CPU host-build tests and the GPU parity tests)."""
import numpy as np

import voxelrt_amd as v


def sparse_scenes():
    """Scenes where whole brick rows are empty: isolated blocks in a large
    grid, a floor, an empty grid, a full grid, and non-cubic grids whose
    padded bitmap words cover several rows (bgx < 32) or part of one
    (bgx > 32)."""
    out = []
    g = np.zeros((96, 64, 128), np.uint32)  # nz, ny, nx
    g[40:44, 20:23, 60:70] = 0x00808080 | 1
    g[10:12, 50:52, 5:7] = 0x00FF0000 | 1
    g[90, 0, 127] = 0x0000FF00 | 1
    out.append(("blocks", g))
    f = np.zeros((128, 128, 128), np.uint32)
    f[:, 0:9, :] = 0x00406080 | 1
    f[60:64, 9:50, 30:34] = 0x00804020 | 1  # a pillar
    out.append(("floor", f))
    out.append(("empty", np.zeros((32, 40, 48), np.uint32)))
    out.append(("full", np.full((16, 16, 16), 0x00101010 | 1, np.uint32)))
    r = np.zeros((40, 24, 24), np.uint32)  # bgx 3: four rows per bitmap word
    r[5:7, 3:5, 20:22] = 0x00202020 | 1
    r[30, 20, 1] = 0x00303030 | 1
    out.append(("narrow", r))
    w = np.zeros((16, 32, 512), np.uint32)  # bgx 64: two bitmap words per row
    w[8, 16, 300:310] = 0x00505050 | 1
    out.append(("wide", w))
    return out


def sparse_rays(vox, n, seed):
    nz, ny, nx = vox.shape
    rng = np.random.default_rng(seed)
    o = rng.random((n, 3)) * np.array([nx, ny, nz]) * v.VOXEL_SIZE
    # a quarter of the origins outside the grid, a tenth of the directions
    # along an axis or a diagonal of two axes (ties between the axes' tMax)
    o[: n // 4] = (rng.random((n // 4, 3)) * 1.6 - 0.3) * np.array([nx, ny, nz]) * v.VOXEL_SIZE
    d = rng.normal(size=(n, 3))
    k = n // 10
    d[:k] = np.eye(3)[rng.integers(0, 3, k)] * rng.choice([-1.0, 1.0], (k, 1))
    d[k:2 * k, rng.integers(0, 3)] = 0.0
    d /= np.maximum(np.linalg.norm(d, axis=1, keepdims=True), 1e-30)
    return np.concatenate([o, d], 1).astype(np.float32)
