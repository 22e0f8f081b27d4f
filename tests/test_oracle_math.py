"""Pins for the oracle's scalar building blocks (CPU only).

Each check compares against something other than the oracle's own formula:
libm (Python math, double precision), numpy's IEEE float16 conversion, or a
mathematical property (bijectivity, uniformity, the hemisphere and cone
moments).
"""
import math

import numpy as np
import pytest

import oracle

L = None


def setup_module():
    global L
    L = oracle.lib()


# ---------------------------------------------------------------------------
# PCG hash, common.wgsl:69-73
# ---------------------------------------------------------------------------
def _pcg_py(v):
    # an independent typing of the WGSL source with Python integers
    v = (v * 747796405 + 2891336453) & 0xFFFFFFFF
    w = (((v >> ((v >> 28) + 4)) ^ v) * 277803737) & 0xFFFFFFFF
    return ((w >> 22) ^ w) & 0xFFFFFFFF


def test_pcg_matches_independent_typing():
    rng = np.random.default_rng(0)
    xs = list(range(4096)) + [int(x) for x in rng.integers(0, 2**32, 4096, dtype=np.uint64)] + [2**32 - 1]
    for x in xs:
        assert oracle.pcg_hash(x) == _pcg_py(x), x


def test_pcg_is_injective_on_a_block():
    # every stage of pcgHash is a bijection of u32 (odd multiplier, xorshift
    # by >= 4 bits, odd multiplier, xorshift 22); a dropped or wrong stage
    # typically breaks injectivity on a contiguous block
    vals = np.array([oracle.pcg_hash(x) for x in range(1 << 16)], np.uint64)
    assert len(np.unique(vals)) == 1 << 16


def test_rand_stream_seeding_and_uniformity():
    # initRng(pix, frame) = pcg(px + pcg(py + pcg(frame))), common.wgsl:75-77
    px, py, fr = 17, 5, 3
    s0 = _pcg_py((px + _pcg_py((py + _pcg_py(fr)) & 0xFFFFFFFF)) & 0xFFFFFFFF)
    states, vals = oracle.rand_stream(px, py, fr, 8)
    assert states[0] == _pcg_py(s0)
    assert vals[0] == np.float32(np.float32(states[0]) * np.float32(2.0**-32))
    # uniformity of rand() over many pixels: chi-square on 16 bins
    allv = np.concatenate([oracle.rand_stream(x, y, 0, 4)[1] for x in range(64) for y in range(16)])
    hist = np.histogram(allv, bins=16, range=(0, 1))[0]
    exp = len(allv) / 16
    chi2 = ((hist - exp) ** 2 / exp).sum()
    assert chi2 < 40, chi2  # 15 dof, p ~ 5e-4


# ---------------------------------------------------------------------------
# own transcendentals vs libm
# ---------------------------------------------------------------------------
def test_sin_cos_accuracy_over_phi_range():
    xs = np.linspace(0.0, 2 * math.pi, 20001, dtype=np.float32)
    es = max(abs(L.orc_sin(float(x)) - math.sin(float(x))) for x in xs)
    ec = max(abs(L.orc_cos(float(x)) - math.cos(float(x))) for x in xs)
    assert es < 1.2e-7 and ec < 1.2e-7, (es, ec)
    # quadrant boundaries and the identity sin^2 + cos^2 = 1
    for x in [0.0, math.pi / 2, math.pi, 1.5 * math.pi, 2 * math.pi, -1.0, -3.0]:
        s, c = L.orc_sin(x), L.orc_cos(x)
        assert abs(s - math.sin(x)) < 2e-7 and abs(c - math.cos(x)) < 2e-7
        assert abs(s * s + c * c - 1) < 5e-7


def test_exp2_log2_pow_exp_accuracy():
    rng = np.random.default_rng(1)
    for y in np.concatenate([np.linspace(-125.9, 127.9, 5001), rng.uniform(-20, 20, 2000)]).astype(np.float32):
        r = L.orc_exp2(float(y))
        assert abs(r / 2.0 ** float(y) - 1) < 2.5e-7, y
    for x in np.exp(rng.uniform(-87, 88, 5000)).astype(np.float32):
        assert abs(L.orc_log2(float(x)) - math.log2(float(x))) < 2e-7 * max(1.0, abs(math.log2(float(x)))) + 1e-7
    assert L.orc_exp2(-200.0) == 0.0 and math.isinf(L.orc_exp2(200.0))
    assert math.isinf(L.orc_log2(0.0)) and L.orc_log2(0.0) < 0
    assert L.orc_log2(1e-42) == pytest.approx(math.log2(np.float32(1e-42)), abs=1e-5)  # denormal
    for x in np.linspace(0.001, 1.0, 500, dtype=np.float32):
        assert abs(L.orc_pow(float(x), 0.55) / float(x) ** 0.55 - 1) < 1e-6
        # exp(x) = exp2(x log2 e): rounding the product costs |x| * 2^-24 * ln 2
        assert abs(L.orc_exp(-float(x) * 30) / math.exp(-float(x) * 30) - 1) < 3e-7 + float(x) * 30 * 1.2e-7
    assert L.orc_pow(0.0, 0.55) == 0.0


# ---------------------------------------------------------------------------
# f16 storage (rgba16float textures): IEEE round-to-nearest-even
# ---------------------------------------------------------------------------
def test_f16_conversion_matches_numpy():
    rng = np.random.default_rng(2)
    halfs = np.arange(65536, dtype=np.uint16)
    h32 = halfs.view(np.float16).astype(np.float32)
    finite = np.isfinite(h32)
    # midpoints between consecutive halfs exercise ties-to-even
    mids = ((h32[:-1].astype(np.float64) + h32[1:].astype(np.float64)) / 2)
    mids = mids[np.isfinite(mids)].astype(np.float32)
    xs = np.concatenate([h32[finite], mids, np.nextafter(mids, np.float32(np.inf)),
                         np.nextafter(mids, np.float32(-np.inf)),
                         rng.normal(0, 1000, 20000).astype(np.float32),
                         np.exp(rng.uniform(-30, 12, 20000)).astype(np.float32),
                         np.array([65504, 65519.99, 65520, 1e6, -1e6, 6e-8, 2.98e-8, 1e-45], np.float32)])
    want = xs.astype(np.float16).view(np.uint16)
    got = np.array([oracle.f32_to_f16(float(x)) for x in xs], np.uint16)
    bad = np.flatnonzero(got != want)
    assert len(bad) == 0, [(xs[i], got[i], want[i]) for i in bad[:5]]
    back = np.array([oracle.f16_to_f32(int(h)) for h in halfs[finite]], np.float32)
    assert np.array_equal(back, h32[finite])


# ---------------------------------------------------------------------------
# sampling (common.wgsl:118-139): cone and cosine-hemisphere moments
# ---------------------------------------------------------------------------
def test_sun_cone_samples():
    cos_r = math.cos(math.radians(0.5))
    u, _ = oracle.make_uniforms((1, 2, 3), 0.0, 0.0, 64, 64, sun_elevation=math.radians(40),
                                sun_cos_radius=cos_r)
    w = np.array(u.sun_dir[:3], np.float64)
    cs = []
    for i in range(4000):
        st = oracle.pcg_hash(i * 7919 + 1)
        d, _ = oracle.sample_sun_dir(u, st)
        assert abs(np.linalg.norm(d) - 1) < 1e-6
        cs.append(float(np.dot(d, w)))
    cs = np.array(cs)
    assert cs.min() >= cos_r - 2e-7
    # uniform over the cap's solid angle: cos(theta) ~ U(cos_r, 1)
    assert abs(cs.mean() - (1 + cos_r) / 2) < 3e-6


@pytest.mark.parametrize("n", [(0, 1, 0), (1, 0, 0), (0, 0, -1), (0.6, 0.0, 0.8)])
def test_cosine_hemisphere_moments(n):
    n = np.array(n, np.float32)
    c = []
    for i in range(20000):
        d, _ = oracle.cosine_hemisphere(n, oracle.pcg_hash(i + 99))
        c.append(float(np.dot(d, n)))
    c = np.array(c)
    assert c.min() >= -1e-6
    assert abs(c.mean() - 2 / 3) < 0.01   # E[cos] = 2/3 for a cosine lobe
    assert abs((c * c).mean() - 0.5) < 0.01  # E[cos^2] = 1/2


# ---------------------------------------------------------------------------
# camera (camera.js:47-56, math.js, common.wgsl:90-95)
# ---------------------------------------------------------------------------
def test_camera_center_ray_is_forward():
    for pos, yaw, pitch in [((8.8, 4.7, 7.0), -0.45, 0.02), ((1, 2, 3), 2.0, -0.7), ((0, 0, 0), 0.0, 0.0)]:
        u, _ = oracle.make_uniforms(pos, yaw, pitch, 1920, 1080)
        d = oracle.camera_ray(u, 0.5, 0.5)
        fw = np.array([math.sin(yaw) * math.cos(pitch), math.sin(pitch), math.cos(yaw) * math.cos(pitch)])
        assert np.allclose(d, fw, atol=2e-5), (d, fw)


def test_camera_frustum_corners():
    pos, yaw, pitch = (3.0, 1.0, -2.0), 0.3, 0.1
    w, h = 640, 360
    u, _ = oracle.make_uniforms(pos, yaw, pitch, w, h)
    fw = np.array([math.sin(yaw) * math.cos(pitch), math.sin(pitch), math.cos(yaw) * math.cos(pitch)])
    rt = np.array([-math.cos(yaw), 0, math.sin(yaw)])
    up = np.cross(rt, fw)
    th = math.tan(math.radians(70) / 2)
    for ux, uy in [(0, 0), (1, 0), (0, 1), (1, 1), (0.25, 0.75)]:
        ndc = (ux * 2 - 1, 1 - uy * 2)
        want = fw + rt * ndc[0] * th * (w / h) + up * ndc[1] * th
        want /= np.linalg.norm(want)
        assert np.allclose(oracle.camera_ray(u, ux, uy), want, atol=3e-5)


def test_sky_closed_forms():
    u, _ = oracle.make_uniforms((0, 0, 0), 0, 0, 8, 8)
    s = np.array(u.sun_dir[:3], np.float64)
    for rd, base in [((0, 1, 0), (0.15, 0.32, 0.62)), ((1, 0, 0), (0.55, 0.68, 0.85)),
                     ((0, -1, 0), (0.28, 0.26, 0.24))]:
        rd = np.array(rd, np.float64)
        amt = max(float(rd @ s), 0.0)
        want = np.array(base) + np.array([1, 0.75, 0.5]) * amt ** 8 * 0.2
        assert np.allclose(oracle.sky_color(u, rd, False), want, atol=2e-6)
    # the sun disc adds E / solid angle
    got = oracle.sky_color(u, s.astype(np.float32), True) - oracle.sky_color(u, s.astype(np.float32), False)
    omega = 2 * math.pi * (1 - u.sun_dir[3])
    assert np.allclose(got, np.array([5.0, 4.5, 3.8]) / omega, rtol=2e-4)
