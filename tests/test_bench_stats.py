"""bench.py's reference-shaped performance figures (test/bench.mjs:480-553)
on inputs whose answers are worked out by hand."""
import math
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402


def test_sample_stats_hand_values():
    s = bench._sample_stats([4.0, 1.0, 3.0, 2.0])
    assert s["n"] == 4
    assert s["mean"] == 2.5
    assert s["median"] == 2.5  # halfway between the 2nd and 3rd order statistics
    # p95: position (n - 1) * 0.95 = 2.85 -> 3 + 0.85 * (4 - 3)
    assert s["p95"] == 3.85
    # sample (n - 1) standard deviation: sqrt(5 / 3)
    assert s["stddev"] == round(math.sqrt(5.0 / 3.0), 4)
    assert s["cv"] == round(math.sqrt(5.0 / 3.0) / 2.5, 4)
    assert (s["min"], s["max"]) == (1.0, 4.0)


def test_sample_stats_single_and_nonfinite():
    s = bench._sample_stats([7.0, float("nan"), float("inf")])
    assert s["n"] == 1 and s["stddev"] == 0.0 and s["cv"] == 0.0
    assert s["median"] == s["p95"] == s["min"] == s["max"] == 7.0
    assert bench._sample_stats([]) is None


def test_perf_stats_shares_and_dominant_pass():
    p = bench._perf_stats(10, 2000, 1000, 4.0, {"pathtrace": 3.0, "temporal": 0.5, "atrous": 0.5})
    assert p["megapixels"] == 2.0
    assert p["gpu_ms_per_frame"] == 4.0 and p["gpu_fps"] == 250.0
    assert p["gpu_ms_per_megapixel"] == 2.0
    assert p["gpu_run_ms"] == 40.0 and p["wall_run_ms"] == 40.0 and p["wall_fps"] == 250.0
    assert p["gpu_pass_percent"] == {"pathtrace": 75.0, "temporal": 12.5, "atrous": 12.5}
    assert p["dominant_pass"] == "pathtrace" and p["dominant_pass_percent"] == 75.0
