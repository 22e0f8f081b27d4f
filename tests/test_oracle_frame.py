"""Closed-form pins of the oracle's frame passes (CPU only):

  path tracer (pathtrace.wgsl:248-330)
    - an empty grid renders the analytic sky with the sun disc;
    - a floor lit by the sun with 0 bounces is Lambertian: demodulated
      radiance = E (n.l) / pi with l inside the sun cone;
    - inside a closed box of one emissive material with 2 bounces every pixel
      is the geometric series e*24*(1 + a + a^2) (no randomness survives);
  temporal (temporal.wgsl:17-77)
    - off -> copy; static camera -> running mean with alpha = 1/history up to
      the cap;
  a-trous (atrous.wgsl:26-83)
    - a constant image with constant normal/depth is a fixed point;
    - sky pixels are copied.
"""
import math

import numpy as np
import pytest

import oracle
from tests.chain import f16


def _mat(r, g, b, e=0):
    return np.uint32(r | (g << 8) | (b << 16) | (e << 24))


def test_empty_grid_is_sky():
    g = oracle.OracleGrid(np.zeros((16, 16, 16), np.uint32))
    u, _ = oracle.make_uniforms((0.5, 0.5, -3.0), 0.9, 0.6, 48, 32, bounces=2)
    rad, alb, gb, cnt = oracle.pathtrace(g, u, 48.0, 1)
    assert cnt.rays == 48 * 32 and cnt.hits == 0
    assert np.all(alb == 255) and np.all(gb[..., 3] == -1)
    for y, x in [(0, 0), (10, 20), (31, 47)]:
        rd = -gb[y, x, :3]
        want = oracle.sky_color(u, rd, True)
        assert np.allclose(f16(rad[y, x, :3]), want, rtol=1e-3)
    assert np.all(f16(rad[..., 3]) == 1.0)


def test_sunlit_floor_is_lambertian():
    n = 64
    vox = np.zeros((n, n, n), np.uint32)
    vox[:, 0:4, :] = _mat(180, 180, 180)
    g = oracle.OracleGrid(vox)
    el = math.radians(60)
    cos_r = math.cos(math.radians(0.5))
    # camera above the floor looking straight down-ish: every pixel hits it
    u, _ = oracle.make_uniforms((2.0, 1.5, 2.0), 0.0, -1.45, 32, 32, bounces=0, sun_elevation=el,
                                sun_cos_radius=cos_r)
    rad, alb, gb, _ = oracle.pathtrace(g, u, 0.0, 1)
    assert np.all(gb[..., 3] > 0) and np.all(gb[..., 1] == 1.0)
    E = np.array([5.0, 4.5, 3.8])
    lo = E * math.cos(math.pi / 2 - el + math.radians(0.5)) / math.pi
    hi = E * math.cos(math.pi / 2 - el - math.radians(0.5)) / math.pi
    r = f16(rad)[..., :3]
    assert np.all(r >= lo * (1 - 2e-3)) and np.all(r <= hi * (1 + 2e-3))
    assert np.allclose(r.mean((0, 1)), E * math.sin(el) / math.pi, rtol=3e-3)
    assert np.all(alb[..., :3] == 180)


@pytest.mark.parametrize("bounces", [0, 1, 2])
def test_closed_emissive_box_geometric_series(bounces):
    n = 32
    m = _mat(200, 150, 100, 20)
    vox = np.full((n, n, n), m, np.uint32)
    vox[8:24, 8:24, 8:24] = 0
    g = oracle.OracleGrid(vox)
    u, _ = oracle.make_uniforms((1.0, 1.0, 1.0), 0.7, 0.2, 24, 16, bounces=bounces)
    rad, alb, gb, _ = oracle.pathtrace(g, u, 0.0, 1)
    a = (np.array([200, 150, 100], np.float64) / 255) ** 2
    e = 20 / 255 * 24
    want = e * sum(a ** k for k in range(bounces + 1))
    assert np.allclose(f16(rad)[..., :3], want, rtol=1e-3)


def _uniform_gbuf(h, w, t=3.0):
    gb = np.zeros((h, w, 4), np.float32)
    gb[..., 1] = 1.0
    gb[..., 3] = t
    return gb


def _f16img(vals, h, w):
    a = np.zeros((h, w, 4), np.float16)
    a[..., :3] = vals
    a[..., 3] = 1.0
    return a.view(np.uint16)


def test_temporal_off_copies():
    h, w = 8, 12
    u, _ = oracle.make_uniforms((0, 2, 0), 0, -1.2, w, h, flags=0)
    rad = _f16img([0.3, 0.6, 0.9], h, w)
    out = oracle.temporal(u, rad, _uniform_gbuf(h, w), np.zeros((h, w, 4), np.uint16), _uniform_gbuf(h, w))
    assert np.array_equal(out[..., :3], rad[..., :3]) and np.all(f16(out)[..., 3] == 1)


def test_temporal_static_camera_running_mean():
    """A static camera over a flat floor: history grows by one per frame and
    the accumulation is the running mean of the inputs until the cap."""
    n = 32
    vox = np.zeros((n, n, n), np.uint32)
    vox[:, 0:2, :] = _mat(128, 128, 128)
    g = oracle.OracleGrid(vox)
    h, w = 16, 16
    pos = (1.0, 1.2, 1.0)
    u0, vp = oracle.make_uniforms(pos, 0.0, -1.5, w, h, flags=1, max_history=4.0)
    _, _, gb, _ = oracle.pathtrace(g, u0, 0.0, 1)
    u, _ = oracle.make_uniforms(pos, 0.0, -1.5, w, h, flags=1, prev_view_proj=vp, prev_cam_pos=pos,
                                max_history=4.0)
    acc = np.zeros((h, w, 4), np.float32)
    vals = [1.0, 2.0, 4.0, 8.0, 16.0, 32.0, 64.0]
    for k, val in enumerate(vals):
        acc = oracle.temporal(u, _f16img([val] * 3, h, w), gb, acc, gb, 1)
        inner = acc[2:-2, 2:-2]
        hist = min(k, 4) + 1 if k > 0 else 1
        assert np.all(inner[..., 3] == hist), (k, np.unique(inner[..., 3]))
    # with cap 4 the last frames blend with alpha = 1/5: exponential tail
    ref = vals[0]
    for k, val in enumerate(vals[1:], start=1):
        ref = ref + (val - ref) / (min(k, 4) + 1)
    assert np.allclose(acc[2:-2, 2:-2, 0], ref, rtol=1e-5)


@pytest.mark.parametrize("step", [1, 2, 4])
def test_atrous_constant_image_fixed_point(step):
    h, w = 20, 28
    u, _ = oracle.make_uniforms((0, 0, 0), 0, 0, w, h)
    src = _f16img([0.25, 1.5, 3.0], h, w)
    gb = _uniform_gbuf(h, w)
    gb[:3, :5, 3] = -1.0  # a sky patch is copied, and skipped as a tap
    out = oracle.atrous(u, src, gb, step, 0)
    assert np.array_equal(out, src)


def _present_f64(illum, alb8, exposure):
    """present.wgsl:40-55 in float64 with numpy's pow (independent of the
    oracle's FP32 code and own pow)."""
    s = alb8[..., :3].astype(np.float64) / 255.0
    x = illum[..., :3].astype(np.float64) * s * s * exposure
    t = np.clip((x * (2.51 * x + 0.03)) / (x * (2.43 * x + 0.59) + 0.14), 0.0, 1.0)
    return np.floor(np.power(t, 1.0 / 2.2) * 255.0 + 0.5)


@pytest.mark.parametrize("exposure", [1.0, 0.35, 4.0])
def test_present_matches_float64_formula(exposure):
    rng = np.random.default_rng(7)
    h, w = 64, 96
    illum = np.exp(rng.uniform(-8, 5, (h, w, 4))).astype(np.float32)
    illum[0, :8, :3] = 0.0                   # black
    illum[1, :8, :3] = 1e6                   # saturates to 1 -> 255
    alb = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    alb[:2, :8, :3] = 255
    got = oracle.present(illum, 1, alb, exposure)
    want = _present_f64(illum, alb, exposure)
    assert (got[..., 3] == 255).all()
    assert np.abs(got[..., :3].astype(np.int64) - want.astype(np.int64)).max() <= 1
    assert (got[0, :8, :3] == 0).all() and (got[1, :8, :3] == 255).all()
    # monotone in illumination for a fixed albedo
    ramp = np.linspace(0, 20, 4096, dtype=np.float32)
    il = np.zeros((1, 4096, 4), np.float32)
    il[0, :, 0] = ramp
    a = np.full((1, 4096, 4), 200, np.uint8)
    r = oracle.present(il, 1, a, exposure)[0, :, 0].astype(np.int64)
    assert (np.diff(r) >= 0).all()


def test_present_reads_f16_history():
    rng = np.random.default_rng(3)
    h, w = 16, 16
    f = rng.uniform(0, 3, (h, w, 4)).astype(np.float16)
    alb = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    a = oracle.present(f.view(np.uint16), 0, alb, 1.0)
    b = oracle.present(f.astype(np.float32), 1, alb, 1.0)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("fmt", [0, 1], ids=["f16", "f32"])
def test_hdr_capture_matches_float64_formula(fmt):
    """HDR capture (main.js:628-659): each channel is ir * a * a in double
    with a = albedo byte / 255, stored to a Float32Array -- restated here
    with numpy float64 from the raw texels (f16 history read as halves);
    includes bytes 0 and 255, zero, huge and denormal-range illumination."""
    rng = np.random.default_rng(77 + fmt)
    h, w = 9, 13
    alb = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    alb[0, 0, :3] = (0, 255, 1)
    if fmt == 0:
        bits = rng.integers(0, 0x7BFF, (h, w, 4), dtype=np.uint16)  # finite non-negative halves
        bits[0, 0, :3] = (0, 0x7BFF, 0x0001)                          # 0, max, smallest denormal
        illum = bits
        ir = bits.view(np.float16).astype(np.float64)
    else:
        illum = (rng.random((h, w, 4)) * 50.0).astype(np.float32)
        illum[0, 0, :3] = (0.0, 3.0e38, 1.0e-40)
        ir = illum.astype(np.float64)
    a = alb[..., :3].astype(np.float64) / 255.0
    want = (ir[..., :3] * a * a).astype(np.float32)
    got = oracle.hdr(np.ascontiguousarray(illum), fmt, alb)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
