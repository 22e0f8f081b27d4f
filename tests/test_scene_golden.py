"""The default-scene restatement (vrt_generate_scene, VRT_SCENE_VOXELRT)
against the reference's own generateScene (VoxelRT src/scene.js:232-476):
SHA-256 digests of its packed voxel grid, brick occupancy words and brick
masks, its solid-voxel count and spawn pose, for N = 64, 128, 256 and both
variants, written by tests/golden/make_scene_golden.mjs from the reference
module.  The brick words and masks are the oracle's a3 build over the
restated grid (scene.js:411-433)."""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle
import voxelrt_amd as v

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "voxelrt_scene.json")
CASES = json.load(open(GOLDEN))["scenes"]


def _sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<u4").tobytes()).hexdigest()


@pytest.mark.parametrize("case", CASES, ids=[f"{c['variant']}-{c['n']}" for c in CASES])
def test_default_scene_matches_reference_bytes(case):
    variant = {"default": 0, "lamps": 1}[case["variant"]]
    n = case["n"]
    vox = v.generate_scene(v.SCENE_VOXELRT, n=n, variant=variant)
    assert int((vox != 0).sum()) == case["solid"]
    assert _sha(vox) == case["vox_sha256"]
    bricks, masks = oracle.build_bricks(vox)
    assert _sha(bricks != 0) == case["bricks_sha256"]
    assert _sha(masks) == case["masks_sha256"]
    pos, yaw, pitch = v.scene_spawn(v.SCENE_VOXELRT, n=n, variant=variant)
    assert np.allclose(pos, case["spawn"]["pos"], rtol=0, atol=1e-12)
    assert yaw == case["spawn"]["yaw"] and pitch == case["spawn"]["pitch"]
