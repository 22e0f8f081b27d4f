"""Write the committed golden fixtures (run in the build container, where the
reference tree is mounted; the fixtures then travel with the repository).

voxelbench_trace.npz
    Rays traced by the reference's own VoxelBench traversal -- traceNaive and
    the brick trace of VoxelBench/pipeline/core.cpp:23-130, compiled from the
    reference sources into oracle/_ref by oracle/Makefile -- on a seeded 32^3
    occupancy grid, with the harness.cpp:215-217 ray distribution.  Inputs
    and outputs only.

reference_convergence.json
    The converged channel means of the reference's base preset (800 and 1600
    frames, 1920x1080, 12 bounces, firefly clamp off, default 256^3 scene),
    copied as numbers from test/eval/reference-convergence.v1.json, with the
    scenario poses from test/eval/claim-manifest.v1.json:93-123.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

REF = "/root/reference"


def voxelbench_vectors():
    n = 32
    rng = np.random.default_rng(2024)
    vox = ((rng.random((n, n, n)) < 0.08).astype(np.uint32) *
           rng.integers(1, 2**24, (n, n, n), dtype=np.uint32))
    count = 4000
    o = rng.uniform(1.0, n - 1.0, (count, 3)).astype(np.float32)
    z = rng.uniform(-1, 1, count)
    ph = rng.uniform(0, 2 * np.pi, count)
    r = np.sqrt(1 - z * z)
    d = np.stack([r * np.cos(ph), z, r * np.sin(ph)], 1).astype(np.float32)
    rays = np.concatenate([o, d], 1).astype(np.float32)
    naive_out, naive_t = oracle.vb_trace(vox, rays, naive=True)
    brick_out, brick_t = oracle.vb_trace(vox, rays, naive=False)
    np.savez_compressed(os.path.join(HERE, "voxelbench_trace.npz"), vox=vox, rays=rays,
                        naive_out=naive_out, naive_t=naive_t, brick_out=brick_out, brick_t=brick_t)


def convergence_means():
    conv = json.load(open(os.path.join(REF, "test/eval/reference-convergence.v1.json")))
    claim = json.load(open(os.path.join(REF, "test/eval/claim-manifest.v1.json")))
    out = {"source": "test/eval/reference-convergence.v1.json; poses test/eval/claim-manifest.v1.json",
           "resolution": conv["resolution"], "reference": claim["reference"],
           "scenarios": {}}
    for row in conv["rows"]:
        sc = row["scenario"]
        m = row["primary"]["metrics"]
        out["scenarios"][sc] = {
            "pose": claim["scenarios"][sc]["pose"],
            "scene": claim["scenarios"][sc].get("scene", "default"),
            "motion": claim["scenarios"][sc].get("motion"),
            "evaluationFrame": row["evaluationFrame"],
            "mean_rgb_1600": m["referenceMeanRgb"],
            "mean_rgb_800": m["testMeanRgb"],
        }
    with open(os.path.join(HERE, "reference_convergence.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    oracle.build()
    voxelbench_vectors()
    convergence_means()
    print("fixtures written to", HERE)
