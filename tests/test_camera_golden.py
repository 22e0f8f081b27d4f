"""vrt_camera_matrices against the reference's own Camera.matrices
(VoxelRT src/camera.js:47-57 over src/math.js): viewProj and invViewProj as
the float32 words main.js:407 stores, for 16 poses x 5 aspect ratios
(fov 70 deg, near 0.05, far 400), written by
tests/golden/make_camera_golden.mjs from the reference modules."""
import json
import os

import numpy as np
import pytest

import voxelrt_amd as v

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "voxelrt_camera.json")
CASES = json.load(open(GOLDEN))["cases"]


@pytest.mark.parametrize("k", range(0, len(CASES), 5), ids=lambda k: f"pose{k // 5}")
def test_camera_matrices_match_reference_bits(k):
    for c in CASES[k:k + 5]:  # the pose at each aspect ratio
        vp, ivp = v.camera_matrices(c["pos"], c["yaw"], c["pitch"], aspect=c["aspect"])
        assert vp.view(np.uint32).tolist() == c["view_proj"], (c["pos"], c["aspect"])
        assert ivp.view(np.uint32).tolist() == c["inv_view_proj"], (c["pos"], c["aspect"])
