"""Closed-form pins of the oracle's filter weights, reprojection and
estimator (CPU only).  Each case is built so that one factor of the
reference formula decides the result, and the expected value comes from the
formula in float64 (numpy), not from the oracle's code:

  a-trous (atrous.wgsl:26-83) on images where the centre has exactly one
  valid tap: the B3 kernel product K[|dx|] K[|dy|] for every tap offset and
  step, the normal weight max(n.nq, 0)^32, the depth weight
  exp(-|dt| / (0.05 t + 0.01)), the luminance weight exp(-|dl| / sigmaL)
  with sigmaL = max(0.08, 2 / sqrt(history)), and the skipping of sky taps;

  temporal (temporal.wgsl:17-77): a previous view-projection shifted by half
  a pixel, so the bilinear history fetch is the mean of two texels; the
  depth gate |pg.w - eT| < 0.08 eT + 0.02 and the normal gate dot > 0.85 on
  both sides of their thresholds; reprojections off screen or onto sky;

  path tracer (pathtrace.wgsl:303-319): Russian roulette keeps the estimate
  unbiased (a closed emissive box at 6 bounces has mean e*24*sum a^k while
  most paths end early), and a bounce that misses adds the sky without the
  sun disc (an open sunlit floor at 1 bounce has mean
  E n.l / pi + (1/pi) * integral of sky(w) cos(theta) over the hemisphere,
  the integral by float64 quadrature);

  the a-trous weights' exponential exp_w against float64 exp (error bounded
  by the rounding of x log2 e and the polynomial), its special values, and
  the bilinear filter position sub_texel (8 fractional bits, nearest-even).
"""
import math

import numpy as np
import pytest

import oracle
from tests.chain import f16

K = (0.375, 0.25, 0.0625)
LW = np.array([0.2126, 0.7152, 0.0722])


# ---------------------------------------------------------------------------
# a-trous
# ---------------------------------------------------------------------------
def one_tap_case(c0, c1, n0, n1, t0, t1, hist0, dx, dy, step):
    """Image of (4 step + 1)^2 pixels: the centre and the one tap at
    (dx, dy) * step are surfaces, every other pixel is sky with a bright
    colour (which must not leak in).  Returns (u, src rgba32f, gbuf, centre)."""
    s = 4 * step + 1
    src = np.full((s, s, 4), 5.0, np.float32)
    src[..., 3] = 1.0
    gb = np.zeros((s, s, 4), np.float32)
    gb[..., 1] = 1.0
    gb[..., 3] = -1.0
    cx = cy = 2 * step
    src[cy, cx] = [*c0, hist0]
    gb[cy, cx] = [*n0, t0]
    tx, ty = cx + dx * step, cy + dy * step
    src[ty, tx] = [*c1, 1.0]
    gb[ty, tx] = [*n1, t1]
    u, _ = oracle.make_uniforms((0, 0, 0), 0, 0, s, s)
    return u, src, gb, (cy, cx)


def one_tap_expected(c0, c1, n0, n1, t0, t1, hist0, dx, dy):
    c0, c1, n0, n1 = (np.asarray(a, np.float64) for a in (c0, c1, n0, n1))
    sigma = max(0.08, 2.0 / math.sqrt(max(hist0, 1.0)))
    w = (K[abs(dx)] * K[abs(dy)] * max(float(n0 @ n1), 0.0) ** 32
         * math.exp(-abs(t0 - t1) / (0.05 * t0 + 0.01))
         * math.exp(-abs(float(c0 @ LW) - float(c1 @ LW)) / sigma))
    return (c0 * K[0] ** 2 + c1 * w) / (K[0] ** 2 + w)


def _check(got, want, tag):
    # f16 output: within one half-ulp of rounding plus float32 arithmetic
    assert np.allclose(got, want, rtol=1.2e-3, atol=1e-6), (tag, got, want)


@pytest.mark.parametrize("step", [1, 2, 4])
def test_atrous_kernel_table_every_tap(step):
    c0, c1 = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    n = (0.0, 1.0, 0.0)
    for dy in range(-2, 3):
        for dx in range(-2, 3):
            if dx == 0 and dy == 0:
                continue
            u, src, gb, (cy, cx) = one_tap_case(c0, c1, n, n, 2.0, 2.0, 1.0, dx, dy, step)
            got = f16(oracle.atrous(u, src, gb, step, 1))[cy, cx, :3]
            _check(got, one_tap_expected(c0, c1, n, n, 2.0, 2.0, 1.0, dx, dy), (step, dx, dy))


@pytest.mark.parametrize("cos_t", [1.0, 0.995, 0.98, 0.9, 0.3, 0.0, -0.5])
def test_atrous_normal_weight_pow32(cos_t):
    sin_t = math.sqrt(1 - cos_t * cos_t)
    n0, n1 = (0.0, 1.0, 0.0), (0.0, cos_t, sin_t)
    c0, c1 = (0.1, 0.2, 0.3), (2.0, 1.0, 0.5)
    u, src, gb, (cy, cx) = one_tap_case(c0, c1, n0, n1, 2.0, 2.0, 4.0, 1, 0, 1)
    got = f16(oracle.atrous(u, src, gb, 1, 1))[cy, cx, :3]
    _check(got, one_tap_expected(c0, c1, n0, n1, 2.0, 2.0, 4.0, 1, 0), cos_t)


@pytest.mark.parametrize("t0,t1", [(2.0, 2.11), (2.0, 1.89), (2.0, 2.33), (0.5, 0.53), (9.0, 9.45), (2.0, 2.0)])
def test_atrous_depth_weight(t0, t1):
    n = (1.0, 0.0, 0.0)
    c0, c1 = (0.5, 0.5, 0.5), (0.6, 0.6, 0.6)
    u, src, gb, (cy, cx) = one_tap_case(c0, c1, n, n, t0, t1, 1.0, 0, -1, 1)
    got = f16(oracle.atrous(u, src, gb, 1, 1))[cy, cx, :3]
    _check(got, one_tap_expected(c0, c1, n, n, t0, t1, 1.0, 0, -1), (t0, t1))


@pytest.mark.parametrize("hist0", [0.5, 1.0, 16.0, 400.0, 1000.0])
@pytest.mark.parametrize("l1", [0.21, 0.5, 1.3])
def test_atrous_luminance_weight(hist0, l1):
    """sigmaL = max(0.08, 2/sqrt(max(history, 1))): history 0.5 -> 2 (the
    max with 1), 16 -> 0.5, 400 -> 0.1, 1000 -> 0.08 (the floor)."""
    n = (0.0, 0.0, -1.0)
    c0, c1 = (0.2, 0.2, 0.2), (l1, l1, l1)
    u, src, gb, (cy, cx) = one_tap_case(c0, c1, n, n, 3.0, 3.0, hist0, -2, 1, 1)
    got = f16(oracle.atrous(u, src, gb, 1, 1))[cy, cx, :3]
    _check(got, one_tap_expected(c0, c1, n, n, 3.0, 3.0, hist0, -2, 1), (hist0, l1))
    # coloured tap: luminance, not a channel, sets the weight
    c1 = (l1 * 2.0, l1 * 0.5, l1 * 0.1)
    u, src, gb, (cy, cx) = one_tap_case(c0, c1, n, n, 3.0, 3.0, hist0, -2, 1, 1)
    got = f16(oracle.atrous(u, src, gb, 1, 1))[cy, cx, :3]
    _check(got, one_tap_expected(c0, c1, n, n, 3.0, 3.0, hist0, -2, 1), (hist0, c1))


def test_atrous_skips_sky_taps():
    """Sky neighbours (t < 0) are skipped whatever their weight would be:
    here a near-zero centre depth makes the depth weight of a t = -0.01 sky
    texel large (exp(-0.02/0.0105)), and the sky colour 5 would show."""
    n = (0.0, 1.0, 0.0)
    c0, c1 = (0.2, 0.4, 0.6), (1.0, 1.0, 1.0)
    u, src, gb, (cy, cx) = one_tap_case(c0, c1, n, n, 0.01, 0.012, 1.0, 1, 1, 1)
    gb[gb[..., 3] < 0, 3] = -0.01
    got = f16(oracle.atrous(u, src, gb, 1, 1))[cy, cx, :3]
    _check(got, one_tap_expected(c0, c1, n, n, 0.01, 0.012, 1.0, 1, 1), "sky skip")


def test_atrous_all_factors_together_and_history_kept():
    n0, n1 = (0.0, 1.0, 0.0), (0.0, 0.99, math.sqrt(1 - 0.99 ** 2))
    c0, c1 = (0.3, 0.9, 0.1), (1.2, 0.4, 0.8)
    u, src, gb, (cy, cx) = one_tap_case(c0, c1, n0, n1, 4.0, 4.1, 9.0, 2, 2, 2)
    out = f16(oracle.atrous(u, src, gb, 2, 1))
    _check(out[cy, cx, :3], one_tap_expected(c0, c1, n0, n1, 4.0, 4.1, 9.0, 2, 2), "all")
    assert out[cy, cx, 3] == 9.0
    # sky pixels are copied as they are
    assert np.all(out[0, 0] == src[0, 0])


# ---------------------------------------------------------------------------
# temporal
# ---------------------------------------------------------------------------
W, H = 16, 12
POS = (2.0, 2.0, -3.0)


def _wall(t=3.0, n=(0.0, 0.0, -1.0)):
    gb = np.zeros((H, W, 4), np.float32)
    gb[..., :3] = n
    gb[..., 3] = t
    return gb


def _shifted(vp, dndc_x):
    """prevViewProj = T * viewProj with T adding dndc_x * clip.w to clip.x
    (column-major: element (row, col) at col * 4 + row)."""
    p = np.array(vp, np.float64).copy()
    for c in range(4):
        p[c * 4 + 0] += dndc_x * p[c * 4 + 3]
    return p.astype(np.float32)


def _ramp_history(h_alpha=1.0):
    a = np.zeros((H, W, 4), np.float32)
    xs = np.arange(W, dtype=np.float32)[None, :]
    ys = np.arange(H, dtype=np.float32)[:, None]
    a[..., 0] = xs
    a[..., 1] = 2.0 * xs
    a[..., 2] = 10.0 * ys + xs
    a[..., 3] = h_alpha
    return a


def _cur(val=(1.0, 1.0, 1.0)):
    r = np.zeros((H, W, 4), np.float16)
    r[..., :3] = val
    r[..., 3] = 1.0
    return r.view(np.uint16)


def test_temporal_half_pixel_shift_is_the_mean_of_two_texels():
    u0, vp = oracle.make_uniforms(POS, 0.0, 0.0, W, H, flags=1, max_history=64.0)
    u, _ = oracle.make_uniforms(POS, 0.0, 0.0, W, H, flags=1, prev_view_proj=_shifted(vp, 1.0 / W),
                                prev_cam_pos=POS, max_history=64.0)
    hist = _ramp_history(1.0)
    out = oracle.temporal(u, _cur(), _wall(), hist, _wall(), 1)
    xs = np.broadcast_to(np.arange(W - 1, dtype=np.float64)[None, :], (H, W - 1))
    ys = np.broadcast_to(np.arange(H, dtype=np.float64)[:, None], (H, W - 1))
    # prevUv.x = uv.x + 0.5/W: texel coordinate x + 0.5 -> (x, x+1) half and
    # half; y unchanged.  history 1 -> 2, alpha = 1/2.
    prev = np.stack([xs + 0.5, 2.0 * xs + 1.0, 10.0 * ys + xs + 0.5], -1)
    want = prev * 0.5 + 1.0 * 0.5
    assert np.array_equal(out[:, :W - 1, :3], want.astype(np.float32)), out[0, :4, :3]
    assert np.all(out[:, :W - 1, 3] == 2.0)
    # (the last column reprojects onto prevUv.x = 1 up to rounding: the
    # screen edge, where either side of the test may hold)


def test_temporal_depth_gate_threshold():
    u, _ = oracle.make_uniforms(POS, 0.0, 0.0, W, H, flags=1, max_history=64.0)
    gprev = _wall()
    thr = 0.08 * 3.0 + 0.02
    sgn = np.where(np.arange(W) % 2 == 0, -1.0, 1.0)  # even columns inside, odd outside
    side = np.where(np.arange(W) % 4 < 2, 1.0, -1.0)  # above and below eT
    gprev[..., 3] = (3.0 + side * thr * (1.0 + sgn * 1e-3))[None, :]
    gprev[0, :, 3] = -1.0                             # sky in the previous frame: rejected
    out = oracle.temporal(u, _cur(), _wall(), _ramp_history(3.0), gprev, 1)
    ok = (np.arange(W) % 2 == 0)
    assert np.all(out[1:, ok, 3] == 4.0) and np.all(out[1:, ~ok, 3] == 1.0), out[1, :, 3]
    assert np.all(out[0, :, 3] == 1.0)


def test_temporal_normal_gate_threshold():
    u, _ = oracle.make_uniforms(POS, 0.0, 0.0, W, H, flags=1, max_history=64.0)
    gprev = _wall()
    c = np.where(np.arange(W) % 2 == 0, 0.851, 0.849)
    gprev[..., 1] = np.sqrt(1.0 - c * c)[None, :]
    gprev[..., 2] = -c[None, :]
    out = oracle.temporal(u, _cur(), _wall(), _ramp_history(2.0), gprev, 1)
    assert np.all(out[:, 0::2, 3] == 3.0) and np.all(out[:, 1::2, 3] == 1.0)


def test_temporal_off_screen_and_cap():
    u0, vp = oracle.make_uniforms(POS, 0.0, 0.0, W, H, flags=1, max_history=5.0)
    u, _ = oracle.make_uniforms(POS, 0.0, 0.0, W, H, flags=1, prev_view_proj=_shifted(vp, 2.5),
                                prev_cam_pos=POS, max_history=5.0)
    out = oracle.temporal(u, _cur(), _wall(), _ramp_history(50.0), _wall(), 1)
    assert np.all(out[..., 3] == 1.0) and np.all(out[..., :3] == 1.0)
    # same view: history 50 capped at 5 -> 6, colour = mix(prev, cur, 1/6)
    u, _ = oracle.make_uniforms(POS, 0.0, 0.0, W, H, flags=1, max_history=5.0)
    hist = _ramp_history(50.0)
    out = oracle.temporal(u, _cur(), _wall(), hist, _wall(), 1)
    a = np.float32(1.0) / np.float32(6.0)
    want = hist[..., :3] * (np.float32(1.0) - a) + np.float32(1.0) * a
    assert np.all(out[..., 3] == 6.0) and np.array_equal(out[..., :3], want)


# ---------------------------------------------------------------------------
# path tracer: Russian roulette and the bounce-miss sky
# ---------------------------------------------------------------------------
def _mat(r, g, b, e=0):
    return np.uint32(r | (g << 8) | (b << 16) | (e << 24))


def test_russian_roulette_is_unbiased():
    n = 32
    vox = np.full((n, n, n), _mat(200, 150, 100, 20), np.uint32)
    vox[8:24, 8:24, 8:24] = 0
    g = oracle.OracleGrid(vox)
    w, h, spp, nb = 24, 16, 256, 6
    u, _ = oracle.make_uniforms((1.0, 1.0, 1.0), 0.7, 0.2, w, h, bounces=nb, frame=3)
    rad, _, _, cnt = oracle.pathtrace(g, u, 0.0, spp)
    a = (np.array([200, 150, 100], np.float64) / 255) ** 2
    want = 20 / 255 * 24 * sum(a ** k for k in range(nb + 1))
    px = f16(rad)[..., :3].reshape(-1, 3).astype(np.float64)
    mean, se = px.mean(0), px.std(0) / math.sqrt(len(px))
    assert np.all(np.abs(mean - want) < 5 * se + 1e-3 * want), (mean, want, se)
    # roulette fired: far fewer rays than 1 + 6 bounce + 7 shadow rays per sample
    assert cnt.rays < 0.6 * w * h * spp * (1 + nb + nb + 1)
    # without the 1/p reweighting the mean would sit far below (sanity of the
    # test's resolution): the surviving tail is a sizeable part of the sum
    tail = 20 / 255 * 24 * sum(a ** k for k in range(3, nb + 1))
    assert np.all(tail > 20 * se)


def _sky_integral(u):
    """(1/pi) * integral over the upper hemisphere of skyColor(w, no disc)
    cos(theta), skyColor (common.wgsl:101-115) in float64, midpoint rule."""
    nt, npf = 1200, 1600
    th = (np.arange(nt) + 0.5) * (math.pi / 2) / nt
    ph = (np.arange(npf) + 0.5) * (2 * math.pi) / npf
    T, P = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], -1)
    up = d[..., 1]
    tt = up ** 0.55
    c1, c2 = np.array([0.55, 0.68, 0.85]), np.array([0.15, 0.32, 0.62])
    sky = c1 * (1 - tt[..., None]) + c2 * tt[..., None]
    sd = np.array(u.sun_dir[:3], np.float64)
    amt = np.maximum(d @ sd, 0.0)
    sky = sky + np.array([1.0, 0.75, 0.5]) * (amt ** 8)[..., None] * 0.2
    sky = sky * u.sun_radiance[3]
    wgt = (np.cos(T) * np.sin(T))[..., None] * (math.pi / 2 / nt) * (2 * math.pi / npf)
    return (sky * wgt).sum((0, 1)) / math.pi


def test_bounce_miss_adds_sky_without_disc():
    n = 64
    vox = np.zeros((n, n, n), np.uint32)
    vox[:, 0:4, :] = _mat(180, 180, 180)
    g = oracle.OracleGrid(vox)
    el, cos_r = math.radians(40.0), math.cos(math.radians(0.5))
    w, h, spp = 32, 32, 128
    u, _ = oracle.make_uniforms((2.0, 1.5, 2.0), 0.0, -1.45, w, h, bounces=1, sun_elevation=el,
                                sun_cos_radius=cos_r, frame=9)
    rad, _, gb, _ = oracle.pathtrace(g, u, 0.0, spp)
    assert np.all(gb[..., 3] > 0) and np.all(gb[..., 1] == 1.0)
    E = np.array(u.sun_radiance[:3], np.float64)
    # uniform over the cone: E[l] = sun_dir (1 + cos r) / 2, n = +y
    direct = E * u.sun_dir[1] * (1 + cos_r) / 2 / math.pi
    want = direct + _sky_integral(u)
    px = f16(rad)[..., :3].reshape(-1, 3).astype(np.float64)
    mean, se = px.mean(0), px.std(0) / math.sqrt(len(px))
    assert np.all(np.abs(mean - want) < 5 * se + 2e-4 * want), (mean, want, se)
    # a disc on the bounce ray would add about the direct term once more
    assert np.all(direct > 100 * se)


# ---------------------------------------------------------------------------
# the weights' exponential and the bilinear filter position
# ---------------------------------------------------------------------------
def test_exp_w_against_float64_exp():
    """exp_w (e^x for the a-trous weights, x <= 0) against float64 exp on
    every 97th float from -0 down to -87.3 (~11 M arguments) and around the
    polynomial's range ends: the degree-5 polynomial for 2^f is within 2.1
    ulps on [-1/2, 1/2] and the rounding of x log2(e) costs up to |x| 2^-24
    relative, so |rel| <= 2e-7 for |x| < 1 and 1.5e-7 (|x| + 1) up to the
    smallest normal result -- inside the 3 + 2|x| ulps WGSL allows exp.  A
    dropped last polynomial term (4e-5 at f = 1/2), a wrong coefficient or a
    wrong log2(e) fails it."""
    lo = np.float32(-87.3).view(np.uint32)
    neg0 = np.float32(-0.0).view(np.uint32)
    n = (int(lo) - int(neg0)) // 97
    xs = np.arange(n, dtype=np.uint64) * 97 + int(neg0)
    got = oracle.exp_w_table(int(neg0), 97, n).astype(np.float64)
    x = xs.astype(np.uint32).view(np.float32).astype(np.float64)
    extra = np.array([-0.34657359, -0.5, -0.6931472, -1.0, -0.72134752, -0.72134757], np.float32)
    got = np.concatenate([got, [oracle.exp_w(float(v)) for v in extra]])
    x = np.concatenate([x, extra.astype(np.float64)])
    want = np.exp(x)
    rel = np.abs(got - want) / want
    small = np.abs(x) < 1.0
    assert rel[small].max() <= 2e-7, rel[small].max()
    assert (rel <= 1.5e-7 * (np.abs(x) + 1.0)).all(), (rel / (np.abs(x) + 1.0)).max()


def test_exp_w_special_values():
    """e^0 = 1 exactly; NaN, -inf and arguments whose power of two is below
    2^-160 give 0; results below the smallest normal are within a few
    denormal steps."""
    assert oracle.exp_w(0.0) == 1.0 and oracle.exp_w(-0.0) == 1.0
    assert oracle.exp_w(float("nan")) == 0.0
    assert oracle.exp_w(-float("inf")) == 0.0
    assert oracle.exp_w(-120.0) == 0.0
    for x in (-90.0, -95.0, -100.0):
        assert abs(oracle.exp_w(x) - math.exp(x)) <= 32 * 2.0 ** -149


def test_sub_texel_is_nearest_eighth_bit_position():
    """The bilinear filter position at 8 fractional bits: texel coordinates
    on the 1/256 grid come back exactly as (floor, fraction); midpoints
    between grid points round to the even grid point; the rounded position
    is within 1/512 of the coordinate and never decreases."""
    ks = np.arange(-1024, 1024 * 4)
    for k in ks[::7]:
        tc = float(k) / 256.0
        fl, fr = oracle.sub_texel(tc)
        assert fl == math.floor(tc) and fr == tc - math.floor(tc), (tc, fl, fr)
    for k in (-3, -2, 1, 2, 5, 6, 255, 256, 257):
        tc = (k + 0.5) / 256.0
        q = k if k % 2 == 0 else k + 1
        fl, fr = oracle.sub_texel(tc)
        assert fl + fr == q / 256.0, (tc, fl, fr)
    rng = np.random.default_rng(3)
    tcs = np.sort(rng.uniform(-4.0, 1920.0, 5000).astype(np.float32))
    pos = np.array([sum(oracle.sub_texel(float(t))) for t in tcs])
    assert (np.abs(pos - tcs.astype(np.float64)) <= 1.0 / 512 + 1e-4).all()
    assert (np.diff(pos) >= 0).all()
    fr = np.array([oracle.sub_texel(float(t))[1] for t in tcs])
    assert ((fr >= 0) & (fr < 1)).all() and np.allclose(fr * 256, np.round(fr * 256))
