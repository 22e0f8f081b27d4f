"""GPU parity of the filter passes on the closed-form cases that pin the
oracle (tests/test_oracle_pins.py): the device temporal and a-trous kernels,
called through the C ABI's single-pass entry points, must reproduce the
oracle bit for bit on every one of them -- including G-buffers whose normals
are not axis vectors (the a-trous kernel's general-normal path) and the
half-pixel bilinear history fetch."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
import voxelrt_amd as v
from tests.chain import assert_bits_equal
from tests import test_oracle_pins as pins

pytestmark = pytest.mark.gpu


def _product_uniforms(ou: oracle.Uniforms) -> v.Uniforms:
    u = v.Uniforms()
    C.memmove(C.byref(u), C.byref(ou), 304)
    return u


def _atrous_cases():
    n = (0.0, 1.0, 0.0)
    for step in (1, 2, 4):
        for dy in range(-2, 3):
            for dx in range(-2, 3):
                if dx or dy:
                    yield ((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), n, n, 2.0, 2.0, 1.0, dx, dy, step)
    for cos_t in (1.0, 0.995, 0.98, 0.9, 0.3, 0.0, -0.5):
        s = math.sqrt(1 - cos_t * cos_t)
        yield ((0.1, 0.2, 0.3), (2.0, 1.0, 0.5), (0.0, 1.0, 0.0), (0.0, cos_t, s), 2.0, 2.0, 4.0, 1, 0, 1)
    for t0, t1 in ((2.0, 2.11), (2.0, 1.89), (0.5, 0.53), (9.0, 9.45)):
        yield ((0.5,) * 3, (0.6,) * 3, (1.0, 0.0, 0.0), (1.0, 0.0, 0.0), t0, t1, 1.0, 0, -1, 1)
    for hist0 in (0.5, 16.0, 1000.0):
        yield ((0.2,) * 3, (1.3, 0.3, 0.05), (0.0, 0.0, -1.0), (0.0, 0.0, -1.0), 3.0, 3.0, hist0, -2, 1, 1)
    # axis normals of opposite sign and a zero normal at the centre
    yield ((0.3,) * 3, (0.9,) * 3, (0.0, 0.0, -1.0), (0.0, 0.0, 1.0), 3.0, 3.0, 2.0, 1, 1, 1)
    yield ((0.3,) * 3, (0.9,) * 3, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), 3.0, 3.0, 2.0, 1, 1, 1)


@pytest.mark.parametrize("fmt", [0, 1])
def test_atrous_one_tap_cases_bitexact(ctx, fmt):
    for case in _atrous_cases():
        *args, step = case
        u, src, gb, (cy, cx) = pins.one_tap_case(*args, step)
        if fmt == 0:
            src = src.astype(np.float16).view(np.uint16)
        want = oracle.atrous(u, src, gb, step, fmt)
        got = ctx.atrous_pass(_product_uniforms(u), src, gb, step, fmt)
        assert_bits_equal(f"atrous {case} fmt {fmt}", got, want)


def test_atrous_sky_taps_bitexact(ctx):
    n = (0.0, 1.0, 0.0)
    u, src, gb, _ = pins.one_tap_case((0.2, 0.4, 0.6), (1.0,) * 3, n, n, 0.01, 0.012, 1.0, 1, 1, 1)
    gb[gb[..., 3] < 0, 3] = -0.01
    assert_bits_equal("sky", ctx.atrous_pass(_product_uniforms(u), src, gb, 1, 1), oracle.atrous(u, src, gb, 1, 1))


def test_atrous_random_general_normals_bitexact(ctx):
    """Random unit normals (general path) and random axis normals (code
    path) with random depths, sky pixels and colours on a ragged image."""
    rng = np.random.default_rng(5)
    h, w = 45, 77
    u, _ = oracle.make_uniforms((0, 0, 0), 0, 0, w, h)
    src = rng.uniform(0, 4, (h, w, 4)).astype(np.float32)
    src[..., 3] = rng.integers(1, 70, (h, w))
    for general in (True, False):
        gb = np.zeros((h, w, 4), np.float32)
        if general:
            nrm = rng.normal(size=(h, w, 3))
            gb[..., :3] = nrm / np.linalg.norm(nrm, axis=-1, keepdims=True)
        else:
            ax = rng.integers(0, 3, (h, w))
            sg = rng.choice([-1.0, 1.0], (h, w))
            for a in range(3):
                gb[..., a] = np.where(ax == a, sg, 0.0)
            gb[rng.random((h, w)) < 0.03, :3] = 0.0
        gb[..., 3] = rng.uniform(0.5, 3.0, (h, w))
        gb[rng.random((h, w)) < 0.1, 3] = -1.0
        for step in (1, 2, 4):
            want = oracle.atrous(u, src, gb, step, 1)
            got = ctx.atrous_pass(_product_uniforms(u), src, gb, step, 1)
            assert_bits_equal(f"atrous general={general} step {step}", got, want)


@pytest.mark.parametrize("fmt", [0, 1])
def test_temporal_pin_cases_bitexact(ctx, fmt):
    W, H, POS = pins.W, pins.H, pins.POS
    u0, vp = oracle.make_uniforms(POS, 0.0, 0.0, W, H, flags=1, max_history=64.0)
    cases = []
    for dndc in (1.0 / W, 0.37 / W, -0.5 / W, 2.5):
        u, _ = oracle.make_uniforms(POS, 0.0, 0.0, W, H, flags=1, prev_view_proj=pins._shifted(vp, dndc),
                                    prev_cam_pos=POS, max_history=64.0)
        cases.append((u, pins._wall()))
    gdepth = pins._wall()
    thr = 0.08 * 3.0 + 0.02
    gdepth[..., 3] = (3.0 + np.where(np.arange(W) % 4 < 2, 1.0, -1.0) * thr
                      * (1.0 + np.where(np.arange(W) % 2 == 0, -1.0, 1.0) * 1e-3))[None, :]
    gdepth[0, :, 3] = -1.0
    cases.append((u0, gdepth))
    gnorm = pins._wall()
    c = np.where(np.arange(W) % 2 == 0, 0.851, 0.849)
    gnorm[..., 1] = np.sqrt(1.0 - c * c)[None, :]
    gnorm[..., 2] = -c[None, :]
    cases.append((u0, gnorm))
    hist = pins._ramp_history(3.0)
    if fmt == 0:
        hist = hist.astype(np.float16).view(np.uint16)
    for k, (u, gprev) in enumerate(cases):
        want = oracle.temporal(u, pins._cur(), pins._wall(), hist, gprev, fmt)
        got = ctx.temporal_pass(_product_uniforms(u), pins._cur(), pins._wall(), hist, gprev, fmt)
        assert_bits_equal(f"temporal case {k} fmt {fmt}", got, want)
