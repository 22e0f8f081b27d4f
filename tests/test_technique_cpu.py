"""VoxelBench's Technique interface driving the adapter MethodHIP
(include/voxelrt_technique.h) on the CPU, with every host access checked:
oracle/_ref/vb_technique_driver_asan is the GPU test's driver built with
AddressSanitizer (leak checking on) and UBSan, its vrt_* calls answered by
the oracle's frame chain (oracle/vrt_cpu_stub.cpp).  It runs VoxelBench's
own scene code (scenes.h) and harness order -- build, init, frame, event,
onEvent, frame, orientation pass, exit with ~MethodHIP -- for bunker and
cavern (geometry edits) and courtyard (a light change), and must end with
status 0 and no sanitizer report; each frame's `out` must equal the Python
oracle chain on the grid the adapter uploaded, as in the GPU test.

Built here from the reference sources (make -C oracle asan); skipped where
they are absent."""
import fcntl
import os
import subprocess

import numpy as np
import pytest

import oracle
from tests.chain import OracleChain, assert_bits_equal

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "vb_technique_driver_asan")
REF = "/root/reference/VoxelBench/pipeline/core.cpp"
W, H = 240, 135  # VoxelBench core.h:138


@pytest.fixture(scope="module")
def driver():
    if not os.path.exists(REF):
        pytest.skip("VoxelBench sources absent")
    os.makedirs(os.path.dirname(DRIVER), exist_ok=True)
    with open(DRIVER + ".lock", "w") as lk:  # parallel test workers build it once
        fcntl.flock(lk, fcntl.LOCK_EX)
        r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "asan"], capture_output=True,
                           text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    return DRIVER


def _grid(path):
    raw = np.fromfile(path, np.uint32)
    px, py, pz = (int(x) for x in raw[:3])
    return raw[3:].reshape(pz, py, px).copy()


@pytest.mark.parametrize("scene,frames,event", [("bunker", 4, 2), ("courtyard", 3, 1), ("cavern", 3, 1)])
def test_technique_adapter_clean_under_sanitizers(driver, tmp_path, scene, frames, event):
    env = dict(os.environ, OMP_NUM_THREADS="4",
               ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([driver, scene, str(frames), str(event), str(tmp_path)], capture_output=True, text=True,
                       timeout=300, env=env)
    log = r.stdout + r.stderr
    assert r.returncode == 0, log
    assert "runtime error" not in log and "Sanitizer" not in log, log
    assert log.count("frame ") == frames and "done" in r.stdout, log
    chain = OracleChain(oracle.OracleGrid(_grid(tmp_path / "grid0.bin")), W, H, 1)
    g1 = tmp_path / "grid1.bin"
    assert g1.exists() == (scene != "courtyard")  # courtyard's event changes the light only
    for f in range(frames):
        if f == event and g1.exists():
            chain.g = oracle.OracleGrid(_grid(g1))
        raw = (tmp_path / f"frame{f:03d}.bin").read_bytes()
        u = oracle.Uniforms.from_buffer_copy(raw[:304])
        rays = int(np.frombuffer(raw[304:312], np.int64)[0])
        out = np.frombuffer(raw[312:], np.float32).reshape(H, W, 3)
        want = chain.frame(u, 48.0)
        assert_bits_equal(f"{scene} frame {f} out", out, want["hdr"])
        assert rays == want["counters"].rays


def _sun_after(u):
    d = np.array(u.sun_dir[:3], np.float64)
    return d / np.linalg.norm(d), np.array(u.sun_radiance[:4], np.float64)


def test_technique_light_change_under_sanitizers(driver, tmp_path):
    """onEvent(lightChange) after a geometry event: cavern's collapse at frame
    1, then the sun moved (the driver's light_frame, frame 2).  Every frame
    equals the oracle chain, and the frames from the light change on carry
    the new sun in their uniforms (the adapter reads SKY every frame)."""
    frames, event, light = 4, 1, 2
    env = dict(os.environ, OMP_NUM_THREADS="4",
               ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([driver, "cavern", str(frames), str(event), str(tmp_path), str(light)], capture_output=True,
                       text=True, timeout=300, env=env)
    log = r.stdout + r.stderr
    assert r.returncode == 0, log
    assert "runtime error" not in log and "Sanitizer" not in log, log
    assert f"light change before render {light}" in r.stdout, log
    chain = OracleChain(oracle.OracleGrid(_grid(tmp_path / "grid0.bin")), W, H, 1)
    want_dir = np.array([-0.3, 0.9, 0.25]) / np.linalg.norm([-0.3, 0.9, 0.25])
    for f in range(frames):
        if f == event:
            chain.g = oracle.OracleGrid(_grid(tmp_path / "grid1.bin"))
        raw = (tmp_path / f"frame{f:03d}.bin").read_bytes()
        u = oracle.Uniforms.from_buffer_copy(raw[:304])
        d, e = _sun_after(u)
        moved = np.allclose(d, want_dir, atol=1e-5) and np.allclose(e, [14.0, 9.0, 4.5, 1.0])
        assert moved == (f >= light), (f, d, e)
        out = np.frombuffer(raw[312:], np.float32).reshape(H, W, 3)
        assert_bits_equal(f"cavern light-change frame {f} out", out, chain.frame(u, 48.0)["hdr"])
