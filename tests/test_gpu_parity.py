"""GPU parity: libvoxelrt_hip (through the C ABI) against the CPU oracle.

Integer/index outputs (brick words, masks, voxel index, face, step counts)
and every stored texel (f16 radiance, unorm8 albedo, f32 G-buffer, f16/f32
accumulation, f16 a-trous output, f32 HDR) must match bit for bit: both sides
compute FP32 in the same operation order without fused multiply-add and with
the same own transcendental functions.
"""
import math

import numpy as np
import pytest

import oracle
import voxelrt_amd as v
from voxelrt_amd import configs as cf
from tests.chain import OracleChain, assert_bits_equal, f16, to_oracle_uniforms
from tests.sparse import sparse_rays, sparse_scenes

pytestmark = pytest.mark.gpu


def _hits_equal(got, want, name=""):
    for f in ("voxel", "face", "brick_steps", "voxel_steps"):
        bad = np.flatnonzero(got[f] != want[f])
        assert len(bad) == 0, f"{name} {f}: {len(bad)} mismatches, first {bad[:5]} got {got[f][bad[:5]]} want {want[f][bad[:5]]}"
    tb = np.flatnonzero(got["t"].view(np.uint32) != want["t"].view(np.uint32))
    assert len(tb) == 0, f"{name} t: {len(tb)} bit mismatches, first {tb[:5]} got {got['t'][tb[:5]]} want {want['t'][tb[:5]]}"


# ---------------------------------------------------------------------------
# scene build (scene.js:411-433 on the device)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("shape,p", [((64, 64, 64), 0.25), ((48, 32, 64), 0.6), ((32, 32, 32), 0.0),
                                     ((16, 24, 8), 1.0)])
def test_scene_bricks_match_oracle(ctx, shape, p):
    nz, ny, nx = shape
    vox = v.generate_scene(v.SCENE_RANDOM_BRICKS, n=nx, ny=ny, nz=nz, seed=9, p_brick=p, p_voxel=0.5)
    if p == 1.0:
        vox[:] = 0xFF00FF01  # every voxel solid
    sc = ctx.upload(vox)
    words, masks = sc.bricks()
    ob, om = oracle.build_bricks(vox)
    assert np.array_equal(words != 0, ob != 0)
    assert np.array_equal(masks, om)
    # slots are a permutation of 1..nocc in brick order
    occ = words[words != 0]
    assert np.array_equal(occ, np.arange(1, len(occ) + 1, dtype=np.uint32))
    sc.release()


def test_scene_build_multi_slab(ctx):
    """A grid larger than one 512 MiB staging slab (re-uploaded per slab)."""
    vox = v.generate_scene(v.SCENE_RANDOM_BRICKS, n=1024, ny=256, nz=1024, seed=4, p_brick=0.02,
                           p_voxel=0.5)
    sc = ctx.upload(vox)
    words, masks = sc.bricks()
    ob, om = oracle.build_bricks(vox)
    assert np.array_equal(words != 0, ob != 0)
    assert np.array_equal(masks, om)
    # a few rays through it
    rays = cf.traverse_rays(1 << 14, n=256, seed=3) * np.array([4, 1, 4, 1, 1, 1], np.float32)
    got = ctx.trace_rays(sc, rays)
    want = oracle.trace_rays(oracle.OracleGrid(vox), rays)
    _hits_equal(got, want, "1024x256x1024")
    sc.release()


# ---------------------------------------------------------------------------
# traversal (voxel.wgsl:73-153)
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module", params=[0.05, 0.25, 0.6])
def traverse(request, ctx):
    vox = cf.traverse_scene(request.param)
    sc = ctx.upload(vox)
    yield request.param, vox, sc, oracle.OracleGrid(vox)
    sc.release()


TRACE_VARIANTS = {
    "lane": dict(trace_kernel=0),
    "lane-fast": dict(trace_kernel=4),
    "lane-fast-b8": dict(trace_kernel=4, dda_batch=8),
    "lane-fast-b64": dict(trace_kernel=4, dda_batch=64),
    "lane-fast-skip": dict(trace_kernel=4, empty_skip=1),
    "lane-fast-b16-skip": dict(trace_kernel=4, dda_batch=16, empty_skip=1),
    "lane-fast-global": dict(trace_kernel=6),
    "lane-fast-global-b16": dict(trace_kernel=6, dda_batch=16),
    "persist-fast": dict(trace_kernel=5, trace_blocks=2048, trace_steps=4, trace_min_refill=1),
    "persist-fast-333": dict(trace_kernel=5, trace_blocks=333, trace_steps=1, trace_min_refill=64),
    "persist-fast-b16": dict(trace_kernel=5, trace_blocks=1024, trace_steps=8, trace_min_refill=16, dda_batch=16),
    "persist-fast-skip": dict(trace_kernel=5, trace_blocks=700, trace_steps=3, trace_min_refill=8, dda_batch=16,
                              empty_skip=1),
    "lane-fast-b16-w6": dict(trace_kernel=4, dda_batch=16, trace_waves=6),
    "lane-fast-global-w6": dict(trace_kernel=6, trace_waves=6),
    "persist-fast-b16-w6": dict(trace_kernel=5, trace_blocks=1024, trace_steps=8, trace_min_refill=16, dda_batch=16,
                                trace_waves=6),
    "wave-m16-f32": dict(trace_kernel=7, trace_blocks=1024, trace_min_refill=16, trace_steps=32),
    "wave-m1-f1-333": dict(trace_kernel=7, trace_blocks=333, trace_min_refill=1, trace_steps=1),
    "wave-m64-f64-w6": dict(trace_kernel=7, trace_blocks=2048, trace_min_refill=64, trace_steps=64, trace_waves=6),
}


@pytest.fixture(params=list(TRACE_VARIANTS))
def trace_variant(request, ctx):
    old = ctx.tuning().as_dict()
    ctx.set_tuning(**TRACE_VARIANTS[request.param])
    yield TRACE_VARIANTS[request.param]
    ctx.set_tuning(**old)


def test_trace_bitexact_traverse(ctx, traverse, trace_variant):
    p, vox, sc, og = traverse
    rays = cf.traverse_rays(1 << 20, seed=int(p * 100))
    got = ctx.trace_rays(sc, rays)
    want = oracle.trace_rays(og, rays)
    _hits_equal(got, want, f"p={p} {trace_variant}")
    assert (got["t"] >= 0).mean() > 0.5
    st = ctx.stats()
    assert st["rays"] > 0


def test_trace_full_size_properties(ctx, traverse):
    """2^24 rays: every hit lands on a solid voxel, misses carry the miss
    markers, and a 2^16 random subset matches the oracle exactly."""
    p, vox, sc, og = traverse
    rays = cf.traverse_rays(1 << 24, seed=77)
    got = ctx.trace_rays(sc, rays)
    hit = got["t"] >= 0
    flat = vox.reshape(-1)
    assert np.all(flat[got["voxel"][hit]] != 0)
    assert np.all(got["voxel"][~hit] == v.VOXEL_MISS) and np.all(got["face"][~hit] == v.FACE_MISS)
    assert np.all(got["face"][hit] <= v.FACE_NONE)
    idx = np.random.default_rng(1).choice(len(rays), 1 << 16, replace=False)
    _hits_equal(got[idx], oracle.trace_rays(og, rays[idx]), "subset")


def _edge_rays(n: int):
    ext = n * cf.VOXEL_SIZE
    rng = np.random.default_rng(3)
    rays = []
    axes = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], np.float32)
    for d in axes:  # axis-aligned, from inside and outside the grid
        for _ in range(64):
            o = rng.uniform(-0.5 * ext, 1.5 * ext, 3).astype(np.float32)
            rays.append(np.concatenate([o, d]))
    for _ in range(256):  # one or two zero components, signed zeros
        d = rng.normal(size=3).astype(np.float32)
        d[rng.integers(0, 3)] = 0.0
        if rng.random() < 0.5:
            d[rng.integers(0, 3)] = -0.0
        d /= np.linalg.norm(d) + 1e-30
        o = rng.uniform(0, ext, 3).astype(np.float32)
        rays.append(np.concatenate([o, d]))
    for _ in range(256):  # origins exactly on voxel / brick boundaries and grid faces
        o = (rng.integers(0, n + 1, 3) * cf.VOXEL_SIZE).astype(np.float32)
        d = rng.normal(size=3).astype(np.float32)
        d /= np.linalg.norm(d)
        rays.append(np.concatenate([o, d]))
    for _ in range(256):  # diagonal directions: exact tMax ties
        o = (rng.integers(0, n, 3) * cf.VOXEL_SIZE + cf.VOXEL_SIZE / 2).astype(np.float32)
        d = np.array([rng.choice([-1, 1]) for _ in range(3)], np.float32) / np.float32(np.sqrt(3))
        rays.append(np.concatenate([o, d]))
    return np.ascontiguousarray(np.stack(rays), np.float32)


@pytest.mark.parametrize("kind", ["random", "empty", "full", "noncubic"])
@pytest.mark.parametrize("tmax", [1e4, 1.5])
def test_trace_edge_cases(ctx, kind, tmax, trace_variant):
    n = 64
    if kind == "random":
        vox = v.generate_scene(v.SCENE_RANDOM_BRICKS, n=n, seed=5, p_brick=0.3, p_voxel=0.3)
    elif kind == "empty":
        vox = np.zeros((n, n, n), np.uint32)
    elif kind == "full":
        vox = np.full((n, n, n), 0x00808080, np.uint32)
    else:
        vox = v.generate_scene(v.SCENE_RANDOM_BRICKS, n=96, ny=32, nz=48, seed=6, p_brick=0.2, p_voxel=0.4)
    sc = ctx.upload(vox)
    rays = _edge_rays(n)
    _hits_equal(ctx.trace_rays(sc, rays, tmax), oracle.trace_rays(oracle.OracleGrid(vox), rays, tmax),
                f"{kind} tmax={tmax}")
    sc.release()


@pytest.mark.parametrize("name", [s[0] for s in sparse_scenes()])
def test_trace_sparse_scenes(ctx, name, trace_variant):
    """Scenes with long runs of empty brick rows (where the empty-row skip
    takes its bulk steps), with rays from inside and outside the grid."""
    vox = dict(sparse_scenes())[name]
    sc = ctx.upload(vox)
    rays = np.concatenate([sparse_rays(vox, 20000, 11), _edge_rays(max(vox.shape))])
    og = oracle.OracleGrid(vox)
    for tmax in (1e4, 1.3):
        _hits_equal(ctx.trace_rays(sc, rays, tmax), oracle.trace_rays(og, rays, tmax), f"{name} tmax={tmax}")
    sc.release()


def test_autotune_keeps_results(ctx, traverse):
    p, vox, sc, og = traverse
    if p != 0.25:
        pytest.skip("one occupancy is enough")
    old = ctx.tuning().as_dict()
    t = ctx.autotune(sc, 1 << 20)
    assert t.trace_kernel in (0, 4, 5, 6, 7) and t.trace_waves in (0, 6)
    rays = cf.traverse_rays(1 << 18, seed=5)
    _hits_equal(ctx.trace_rays(sc, rays), oracle.trace_rays(og, rays), f"autotuned {t.as_dict()}")
    ctx.set_tuning(**old)


# ---------------------------------------------------------------------------
# path tracer (pathtrace.wgsl:248-330)
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def frame128(ctx):
    vox = cf.frame_128_scene()
    sc = ctx.upload(vox)
    yield vox, sc, oracle.OracleGrid(vox)
    sc.release()


PT_VARIANTS = {
    "lane": dict(pt_kernel=0),
    "lane-w5": dict(pt_kernel=0, pt_waves=5),
    "lane-fast": dict(pt_kernel=4),
    "lane-fast-w5": dict(pt_kernel=4, pt_waves=5),
    "lane-fast-w5-xcd": dict(pt_kernel=4, pt_waves=5, pt_order=1),
    "lane-xcdrev": dict(pt_kernel=0, pt_order=3),
    "lane-fast-global": dict(pt_kernel=7),
    "lane-fast-global-w5": dict(pt_kernel=7, pt_waves=5),
    "lane-fast-global-8x8": dict(pt_kernel=7, pt_tile=1),
    "lane-fast-global-8x8-w4-rowmajor": dict(pt_kernel=7, pt_waves=4, pt_tile=1, pt_order=0),
    "refill-fast-32x32-w4": dict(pt_kernel=4, pt_waves=4, pt_tile=2),
    "refill-global-32x16-w5": dict(pt_kernel=7, pt_waves=5, pt_tile=3),
    "refill-global-32x32-rowmajor": dict(pt_kernel=7, pt_tile=2, pt_order=0),
    "refill-wave-16x16-w4": dict(pt_kernel=7, pt_waves=4, pt_tile=4),
    "refill-wave-16x8-rowmajor": dict(pt_kernel=7, pt_tile=5, pt_order=0),
    "refill-global-32x16-w4": dict(pt_kernel=7, pt_waves=4, pt_tile=3),
    "refill-wave-16x8-w4": dict(pt_kernel=7, pt_waves=4, pt_tile=5),
    "wave-w4-m16-s32": dict(pt_kernel=8, pt_waves=4, pt_min_refill=16, pt_shade_min=32, pt_blocks=0, pt_tile=0),
    "wave-w4-m32-s32": dict(pt_kernel=8, pt_waves=4, pt_min_refill=32, pt_shade_min=32, pt_blocks=0),
    "wave-w4-m1-s1-97blocks-rowmajor": dict(pt_kernel=8, pt_waves=4, pt_min_refill=1, pt_shade_min=1, pt_blocks=97,
                                            pt_order=0),
    "wave-m64-s64-300blocks": dict(pt_kernel=8, pt_waves=0, pt_min_refill=64, pt_shade_min=64, pt_blocks=300),
    "wave-w5-m16-s48-rowmajor": dict(pt_kernel=8, pt_waves=5, pt_min_refill=16, pt_shade_min=48, pt_blocks=0,
                                     pt_order=0),
    "wave-global-w4-m12-s24": dict(pt_kernel=8, pt_waves=4, pt_min_refill=12, pt_shade_min=24, pt_blocks=0,
                                   pt_tile=1),
    "wave-global-w6-m0-s0-rev": dict(pt_kernel=8, pt_waves=6, pt_min_refill=0, pt_shade_min=0, pt_blocks=1024,
                                     pt_tile=1, pt_order=3),
}


@pytest.fixture(params=list(PT_VARIANTS))
def pt_variant(request, ctx):
    old = ctx.tuning().as_dict()
    ctx.set_tuning(**PT_VARIANTS[request.param])
    yield PT_VARIANTS[request.param]
    ctx.set_tuning(**old)


def _render_one(ctx, sc, view_id, u, fclamp, spp, w, h, accum_format=0):
    view = v.View(sc.handle, u, fclamp, view_id, spp, accum_format)
    outs = {"radiance": np.zeros((h, w, 4), np.uint16), "albedo": np.zeros((h, w, 4), np.uint8),
            "gbuf": np.zeros((h, w, 4), np.float32)}
    fr = ctx.render([view], [outs])
    ctx.release_view(view_id)
    return outs, fr[0]


def test_pathtrace_frame128_bitexact(ctx, frame128, pt_variant):
    vox, sc, og = frame128
    pos, yaw, pitch = cf.frame_128_pose()
    p = cf.FRAME_128
    u, _ = v.make_uniforms(pos, yaw, pitch, p["width"], p["height"], frame=0, bounces=p["bounces"],
                           sun_azimuth=p["sun_azimuth"], sun_elevation=p["sun_elevation"],
                           sun_cos_radius=p["sun_cos_radius"])
    outs, fr = _render_one(ctx, sc, 900, u, p["fclamp"], 1, p["width"], p["height"])
    rad, alb, gb, cnt = oracle.pathtrace(og, to_oracle_uniforms(u), p["fclamp"], 1)
    assert_bits_equal("gbuf", outs["gbuf"], gb)
    assert_bits_equal("albedo", outs["albedo"], alb)
    assert_bits_equal("radiance", outs["radiance"], rad)
    assert fr.rays == cnt.rays
    if pt_variant.get("pt_kernel") == 8:
        # the step-level kernel ends rays at the box of occupied bricks (a
        # miss takes fewer brick steps than the grid walk: the scene's bricks
        # stop well below the grid's top) and fine walks at the entered
        # brick's box of solid voxels (fewer voxel steps in bricks passed
        # without a hit); the images above are the grid walk's
        assert 0 < fr.brick_steps < cnt.brick_steps
        assert 0 < fr.voxel_steps < cnt.voxel_steps
    else:
        assert (fr.brick_steps, fr.voxel_steps) == (cnt.brick_steps, cnt.voxel_steps)


@pytest.mark.parametrize("bounces,spp,fclamp,frame", [(4, 3, 0.0, 7), (0, 1, 48.0, 123456789), (6, 2, 10.0, 3)])
def test_pathtrace_variants_bitexact(ctx, frame128, bounces, spp, fclamp, frame, pt_variant):
    """Russian roulette (bounce >= 2), several samples, clamp off/on, seeds."""
    vox, sc, og = frame128
    pos, yaw, pitch = cf.frame_128_pose()
    w, h = 160, 96
    u, _ = v.make_uniforms(pos, yaw + 0.3, pitch, w, h, frame=frame, bounces=bounces)
    outs, fr = _render_one(ctx, sc, 901, u, fclamp, spp, w, h)
    rad, alb, gb, cnt = oracle.pathtrace(og, to_oracle_uniforms(u), fclamp, spp)
    assert_bits_equal("gbuf", outs["gbuf"], gb)
    assert_bits_equal("albedo", outs["albedo"], alb)
    assert_bits_equal("radiance", outs["radiance"], rad)
    assert fr.rays == cnt.rays


def test_pathtrace_ragged_size_bitexact(ctx, frame128, pt_variant):
    """An image no tile form divides (77 x 45: partial 8x8, 16x8, 16x16,
    32x16 and 32x32 tiles at the right and bottom edges)."""
    vox, sc, og = frame128
    pos, yaw, pitch = cf.frame_128_pose()
    w, h = 77, 45
    u, _ = v.make_uniforms(pos, yaw - 0.2, pitch, w, h, frame=11, bounces=3)
    outs, fr = _render_one(ctx, sc, 903, u, 48.0, 2, w, h)
    rad, alb, gb, cnt = oracle.pathtrace(og, to_oracle_uniforms(u), 48.0, 2)
    assert_bits_equal("gbuf", outs["gbuf"], gb)
    assert_bits_equal("albedo", outs["albedo"], alb)
    assert_bits_equal("radiance", outs["radiance"], rad)
    assert fr.rays == cnt.rays


def test_pathtrace_converged_relmse(ctx, frame128):
    """4096 spp over a crop.  With the same PCG per-pixel stream the GPU's
    converged image reaches relMSE < 1e-3 against the oracle's (it is 0: the
    two are equal bit for bit).  Against the oracle's estimate from an
    independent stream (frame seed + 1) the two differ by noise only: their
    per-channel image means agree within 4 standard errors of the
    difference, and the relMSE between them is the noise level (< 1e-2)."""
    vox, sc, og = frame128
    pos, yaw, pitch = cf.frame_128_pose()
    w, h = 64, 36
    kw = dict(bounces=2, sun_elevation=cf.SUN_40DEG["sun_elevation"], sun_cos_radius=cf.SUN_40DEG["sun_cos_radius"])
    u, _ = v.make_uniforms(pos, yaw, pitch, w, h, frame=0, **kw)
    outs, _ = _render_one(ctx, sc, 902, u, 10.0, 4096, w, h)
    a = f16(outs["radiance"])[..., :3].astype(np.float64)

    def relmse(x, y):
        return float(np.mean((x - y) ** 2 / (y * y + 1e-2)))

    same, _, _, _ = oracle.pathtrace(og, to_oracle_uniforms(u), 10.0, 4096)
    assert relmse(a, f16(same)[..., :3].astype(np.float64)) < 1e-3
    u2, _ = v.make_uniforms(pos, yaw, pitch, w, h, frame=1, **kw)
    rad, _, _, _ = oracle.pathtrace(og, to_oracle_uniforms(u2), 10.0, 4096)
    b = f16(rad)[..., :3].astype(np.float64)
    d = (a - b).reshape(-1, 3)
    se = d.std(0) / np.sqrt(len(d))
    rel = relmse(a, b)
    print(f"independent 4096-spp estimates: mean diff {d.mean(0)} (se {se}), relMSE {rel:.2e}", flush=True)
    assert np.all(np.abs(d.mean(0)) < 4 * se), (d.mean(0), se)
    assert 0.0 < rel < 1e-2, rel


# ---------------------------------------------------------------------------
# temporal + a-trous (temporal.wgsl, atrous.wgsl)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("fmt", [0, 1])
def test_filter_passes_bitexact(ctx, frame128, fmt):
    vox, sc, og = frame128
    pos, yaw, pitch = cf.frame_128_pose()
    w, h = 200, 120
    u0, vp0 = v.make_uniforms(pos, yaw, pitch, w, h, frame=0, flags=3)
    pos1 = (pos[0] + 0.05, pos[1], pos[2] + 0.02)
    u1, _ = v.make_uniforms(pos1, yaw + 0.01, pitch, w, h, frame=1, flags=3, prev_view_proj=vp0,
                            prev_cam_pos=pos)
    o0, o1 = to_oracle_uniforms(u0), to_oracle_uniforms(u1)
    rad0, alb0, gb0, _ = oracle.pathtrace(og, o0, 48.0, 1)
    acc0 = oracle.temporal(o0, rad0, gb0, np.zeros((h, w, 4), np.uint16 if fmt == 0 else np.float32),
                           np.zeros((h, w, 4), np.float32), fmt)
    rad1, alb1, gb1, _ = oracle.pathtrace(og, o1, 48.0, 1)
    want = oracle.temporal(o1, rad1, gb1, acc0, gb0, fmt)
    got = ctx.temporal_pass(u1, rad1, gb1, acc0, gb0, fmt)
    assert_bits_equal("temporal", got, want)
    assert (f16(want)[..., 3] if fmt == 0 else want[..., 3]).max() == 2.0  # history grew
    src = want
    for step in (1, 2, 4):
        want_d = oracle.atrous(o1, src, gb1, step, fmt if step == 1 else 0)
        got_d = ctx.atrous_pass(u1, src, gb1, step, fmt if step == 1 else 0)
        assert_bits_equal(f"atrous step {step}", got_d, want_d)
        src = want_d


def test_pathtrace_indirect_only_bitexact(ctx, frame128, pt_variant):
    """VRT_FLAG_INDIRECT (used by the VoxelBench adapter)."""
    vox, sc, og = frame128
    pos, yaw, pitch = cf.frame_128_pose()
    w, h = 128, 72
    u, _ = v.make_uniforms(pos, yaw, pitch, w, h, frame=11, bounces=2, flags=4)
    outs, fr = _render_one(ctx, sc, 903, u, 48.0, 2, w, h)
    rad, alb, gb, cnt = oracle.pathtrace(og, to_oracle_uniforms(u), 48.0, 2)
    assert_bits_equal("radiance", outs["radiance"], rad)
    assert_bits_equal("gbuf", outs["gbuf"], gb)
    assert fr.rays == cnt.rays
    # with one sample, a primary miss contributes nothing (no sky)
    rad, alb, gb, cnt = oracle.pathtrace(og, to_oracle_uniforms(u), 48.0, 1)
    sky = gb[..., 3] < 0
    assert sky.any() and np.all(f16(rad)[sky][..., :3] == 0)


def test_autotune_pathtrace_keeps_results(ctx, frame128):
    vox, sc, og = frame128
    pos, yaw, pitch = cf.frame_128_pose()
    w, h = 160, 96
    u, _ = v.make_uniforms(pos, yaw, pitch, w, h, frame=2, bounces=3)
    old = ctx.tuning().as_dict()
    t = ctx.autotune(sc, 0, u)
    outs, fr = _render_one(ctx, sc, 904, u, 48.0, 1, w, h)
    rad, alb, gb, cnt = oracle.pathtrace(og, to_oracle_uniforms(u), 48.0, 1)
    assert_bits_equal(f"radiance (tuned {t.as_dict()})", outs["radiance"], rad)
    ctx.set_tuning(**old)


def test_frame_sequence_bitexact(ctx, pt_variant):
    """Five frames of an orbit through the whole chain with temporal and
    denoise on: every target of every frame matches the oracle chain."""
    desc = dict(cf.FRAME_1080_SCENE)
    vox = v.generate_scene(**desc)
    sc = ctx.upload(vox)
    og = oracle.OracleGrid(vox)
    w, h = 240, 136
    r = v.Renderer(ctx, sc, w, h, bounces=2, temporal=True, denoise=True, fclamp=48.0, view_id=950)
    chain = OracleChain(og, w, h, 0)
    for i in range(5):
        r.set_pose(*cf.world_orbit(desc, i, 64))
        got = r.frame(("radiance", "albedo", "gbuf", "accum", "denoised", "hdr", "ldr"))
        want = chain.frame(to_oracle_uniforms(r.last_view.u), 48.0)
        for k in ("radiance", "albedo", "gbuf", "accum", "denoised", "hdr", "ldr"):
            assert_bits_equal(f"frame {i} {k}", got[k], want[k])
    ctx.release_view(950)
    sc.release()


@pytest.mark.parametrize("fmt", [0, 1], ids=["rgba16f", "rgba32f"])
def test_filter_chain_ragged(ctx, frame128, fmt):
    """Temporal + a-trous on a ragged size (partial 16x16 tiles at the right
    and bottom edges, halos leaving the image) over four moving frames, in
    both accumulation formats: every target equals the oracle chain."""
    vox, sc, og = frame128
    pos, yaw, pitch = cf.frame_128_pose()
    w, h = 157, 83
    try:
        r = v.Renderer(ctx, sc, w, h, bounces=2, temporal=True, denoise=True, fclamp=48.0, accum_format=fmt,
                       view_id=960)
        chain = OracleChain(og, w, h, fmt)
        for i in range(4):
            r.set_pose((pos[0] + 0.04 * i, pos[1], pos[2] - 0.03 * i), yaw + 0.01 * i, pitch)
            got = r.frame(("accum", "denoised", "hdr", "ldr"))
            want = chain.frame(to_oracle_uniforms(r.last_view.u), 48.0)
            for k in ("accum", "denoised", "hdr", "ldr"):
                assert_bits_equal(f"frame {i} {k}", got[k], want[k])
    finally:
        ctx.release_view(960)


# ---------------------------------------------------------------------------
# batched form: every view of a mixed batch equals the same view alone
# ---------------------------------------------------------------------------
BATCH_VARIANTS = {
    "lane": dict(pt_kernel=0, pt_waves=0),
    "fast-w5": dict(pt_kernel=4, pt_waves=5),
    "queue-fast": dict(pt_kernel=4, pt_waves=5, batch_mode=1),
    "queue-7blocks": dict(pt_kernel=0, pt_waves=0, batch_mode=1, batch_blocks=7),
    "graph-fast": dict(pt_kernel=4, pt_waves=5, batch_mode=2),
    "grid-fast-rowmajor": dict(pt_kernel=4, pt_waves=5, batch_mode=0, pt_order=0),
    "queue-fast-rowmajor": dict(pt_kernel=4, pt_waves=5, batch_mode=1, pt_order=0),
    "queue-fast-global-w4": dict(pt_kernel=7, pt_waves=4, batch_mode=1),
    "graph-lane-w4": dict(pt_kernel=0, pt_waves=4, batch_mode=2),
    "grid-refill16x8": dict(pt_kernel=7, pt_waves=4, pt_tile=5, batch_mode=0),
    "grid-refill16x16-rowmajor": dict(pt_kernel=7, pt_waves=5, pt_tile=4, batch_mode=0, pt_order=0),
    "queue-refill16x8": dict(pt_kernel=7, pt_waves=4, pt_tile=5, batch_mode=1),
    "queue-refill16x16-5blocks": dict(pt_kernel=7, pt_waves=0, pt_tile=4, batch_mode=1, batch_blocks=5),
    "graph-refill16x8": dict(pt_kernel=7, pt_waves=4, pt_tile=5, batch_mode=2, pt_order=1),
    "grid-refill32x16-global": dict(pt_kernel=7, pt_waves=4, pt_tile=3, batch_mode=0),
    "queue-refill32x32-lds": dict(pt_kernel=4, pt_waves=5, pt_tile=2, batch_mode=1),
    "graph-refill32x16-lds-rowmajor": dict(pt_kernel=4, pt_waves=4, pt_tile=3, batch_mode=2, pt_order=0),
    "queue-refill32x32-global-3blocks": dict(pt_kernel=7, pt_waves=5, pt_tile=2, batch_mode=1, batch_blocks=3),
    # the step-level refill kernel's batched form (batch_tile 0), or pt 7 on
    # batch_tile.  Its work items are chunks of patches // (8 x waves) 8x8
    # patches, 1..8: the five views below have 495 patches, so the default
    # grid and 97 waves take chunks of 1, 30 waves of 2, 15 of 4 and 7 of 8
    "wave-grid": dict(pt_kernel=8, pt_waves=4, batch_mode=0, batch_tile=0),
    "wave-grid-7blocks-chunk8": dict(pt_kernel=8, pt_waves=4, batch_mode=0, batch_tile=0, pt_blocks=7),
    "wave-graph-15blocks-chunk4": dict(pt_kernel=8, pt_waves=4, batch_mode=2, batch_tile=0, pt_blocks=15),
    "wave-queue-30blocks-chunk2": dict(pt_kernel=8, pt_waves=4, batch_mode=1, batch_tile=0, pt_blocks=30),
    "wave-queue-97blocks-m1-s1": dict(pt_kernel=8, pt_waves=4, batch_mode=1, batch_tile=0, pt_blocks=97,
                                      pt_min_refill=1, pt_shade_min=1),
    "wave-graph-w6-rowmajor": dict(pt_kernel=8, pt_waves=6, batch_mode=2, batch_tile=0, pt_order=0),
    "wave-graph-refill32x16": dict(pt_kernel=8, pt_waves=6, pt_tile=1, batch_mode=2, batch_tile=3),
}


@pytest.mark.parametrize("tuning", list(BATCH_VARIANTS.values()), ids=list(BATCH_VARIANTS))
def test_batched_views_match_single_views(ctx, tuning):
    descs = [cf.world_desc(64, 1), cf.world_desc(128, 2), dict(cf.world_desc(64, 3), ny=32)]
    scenes = [ctx.upload(v.generate_scene(**d)) for d in descs]
    specs = [  # scene, w, h, bounces, flags, fmt, spp
        (0, 96, 64, 1, 3, 0, 1), (1, 160, 90, 4, 3, 1, 1), (2, 64, 64, 2, 1, 0, 2),
        (1, 33, 17, 6, 0, 0, 1), (0, 128, 40, 3, 3, 1, 1)]
    names = ("radiance", "albedo", "gbuf", "accum", "denoised")

    def make_views(frame, base_id):
        views = []
        for k, (si, w, h, nb, flags, fmt, spp) in enumerate(specs):
            d = descs[si]
            pos, yaw, pitch = cf.world_orbit(d, frame * 5 + k, 64)
            u, _ = v.make_uniforms(pos, yaw, pitch, w, h, frame=frame + 100 * k, bounces=nb, flags=flags)
            views.append(v.View(scenes[si].handle, u, 48.0, base_id + k, spp, fmt))
        return views

    def alloc(w, h, fmt):
        return {"radiance": np.zeros((h, w, 4), np.uint16), "albedo": np.zeros((h, w, 4), np.uint8),
                "gbuf": np.zeros((h, w, 4), np.float32),
                "accum": np.zeros((h, w, 4), np.uint16 if fmt == 0 else np.float32),
                "denoised": np.zeros((h, w, 4), np.uint16)}

    old = ctx.tuning().as_dict()
    ctx.set_tuning(**tuning)
    for frame in range(3):
        batch = make_views(frame, 2000)
        outs_b = [alloc(s[1], s[2], s[5]) for s in specs]
        ctx.render(batch, outs_b)
        for k, view in enumerate(make_views(frame, 3000)):
            outs_s = [alloc(specs[k][1], specs[k][2], specs[k][5])]
            ctx.render([view], outs_s)
            for n in names:
                if n == "denoised" and not (specs[k][4] & 2):
                    continue
                assert_bits_equal(f"frame {frame} view {k} {n}", outs_b[k][n], outs_s[0][n])
    ctx.set_tuning(**old)
    for k in range(len(specs)):
        ctx.release_view(2000 + k)
        ctx.release_view(3000 + k)
    for sc in scenes:
        sc.release()


def test_autotune_batch_keeps_histories(ctx):
    """vrt_autotune_batch picks the batched form's launch shape on the caller's
    views without touching their histories or the ray statistics: the frame
    after the sweep equals the same views' frame rendered without it."""
    descs = [cf.world_desc(64, 4), cf.world_desc(128, 5)]
    scenes = [ctx.upload(v.generate_scene(**d)) for d in descs]
    specs = [(0, 80, 48, 2, 3), (1, 120, 64, 3, 3), (0, 40, 40, 1, 1)]  # scene, w, h, bounces, flags

    def views(frame, base):
        out = []
        for k, (si, w, h, nb, flags) in enumerate(specs):
            pos, yaw, pitch = cf.world_orbit(descs[si], frame * 3 + k, 64)
            u, _ = v.make_uniforms(pos, yaw, pitch, w, h, frame=frame + 10 * k, bounces=nb, flags=flags)
            out.append(v.View(scenes[si].handle, u, 48.0, base + k, 1, 0))
        return out

    def outs():
        return [{"radiance": np.zeros((h, w, 4), np.uint16), "accum": np.zeros((h, w, 4), np.uint16),
                 "denoised": np.zeros((h, w, 4), np.uint16)} for _, w, h, _, _ in specs]

    old = ctx.tuning().as_dict()
    ctx.set_tuning(pt_kernel=7, pt_waves=4)
    ctx.render(views(0, 4000))
    ctx.render(views(0, 5000))
    before = ctx.stats()["rays"]
    t = ctx.autotune_batch(views(1, 4000)).as_dict()
    assert t["batch_mode"] in (0, 1, 2) and t["pt_tile"] in (0, 2, 3, 4, 5)
    t = ctx.autotune_batch(views(1, 4000), modes=(1, 2)).as_dict()  # the persistent modes only
    assert t["batch_mode"] in (1, 2) and t["pt_tile"] in (0, 2, 3, 4, 5)
    assert ctx.stats()["rays"] == before
    assert ctx.tuning().as_dict() == t
    a, b = outs(), outs()
    ctx.render(views(1, 4000), a)
    ctx.render(views(1, 5000), b)
    for k in range(len(specs)):
        for n in ("radiance", "accum", "denoised"):
            if n == "denoised" and not (specs[k][4] & 2):
                continue
            assert_bits_equal(f"view {k} {n}", a[k][n], b[k][n])
    ctx.set_tuning(**old)
    for k in range(len(specs)):
        ctx.release_view(4000 + k)
        ctx.release_view(5000 + k)
    for sc in scenes:
        sc.release()


# ---------------------------------------------------------------------------
# editable scenes: device rebuild after a box edit
# ---------------------------------------------------------------------------
def test_scene_edits_rebuild_on_device(ctx):
    """Recolour solid voxels, clear single voxels of occupied bricks, carve a
    hole, fill a block and rewrite everything in an editable scene: after each
    edit the device bricks/masks, traced hits and a rendered frame equal the
    oracle's on the edited dense grid; the edits that keep every touched
    brick's occupancy are applied in place, the others rebuild the scene (and
    the pool grows when the occupied-brick count does)."""
    desc = cf.world_desc(64, 7)
    vox = v.generate_scene(**desc)
    sc = ctx.upload(vox, editable=True)
    rng = np.random.default_rng(11)
    recolour = np.where(vox[0:16, 0:16, 0:64] != 0, np.uint32(0x10A0B0C0), np.uint32(0))
    # one voxel cleared in each of a few occupied bricks that keep others solid
    nz, ny, nx = vox.shape
    keep = np.argwhere(vox.reshape(nz // 8, 8, ny // 8, 8, nx // 8, 8).sum(axis=(1, 3, 5), dtype=np.uint64) > 0)
    solid = None
    for bz, by, bx in keep:
        blk = vox[bz * 8:bz * 8 + 8, by * 8:by * 8 + 8, bx * 8:bx * 8 + 8]
        if np.count_nonzero(blk) >= 2:
            z, y, x = np.argwhere(blk != 0)[0]
            solid = (int(bx * 8 + x), int(by * 8 + y), int(bz * 8 + z))
            break
    assert solid is not None
    edits = [((0, 0, 0), recolour, True),                                            # in place
             (solid, np.zeros((1, 1, 1), np.uint32), True),                          # in place
             ((8, 0, 8), np.zeros((20, 24, 30), np.uint32), False),                  # carve
             ((30, 30, 30), np.full((10, 17, 9), 0x10406080, np.uint32), False),      # fill into empty air
             ((0, 0, 0), rng.integers(0, 2, (64, 64, 64), dtype=np.uint32) * 0x01808080, False)]  # rewrite all
    rays = cf.traverse_rays(1 << 15, n=64, seed=12)
    for k, (origin, box, in_place) in enumerate(edits):
        x0, y0, z0 = origin
        sz, sy, sx = box.shape
        vox[z0:z0 + sz, y0:y0 + sy, x0:x0 + sx] = box
        before = sc.edit_counts()
        sc.edit(box, origin)
        after = sc.edit_counts()
        assert after == ((before[0] + 1, before[1]) if in_place else (before[0], before[1] + 1)), (k, before, after)
        words, masks = sc.bricks()
        ob, om = oracle.build_bricks(vox)
        assert np.array_equal(words != 0, ob != 0), k
        assert np.array_equal(masks, om), k
        og = oracle.OracleGrid(vox)
        for tk, skip in ((0, 0), (4, 0), (4, 1), (6, 0)):  # 6: global-memory bitmap/prefix pairs
            old = ctx.tuning().as_dict()
            ctx.set_tuning(trace_kernel=tk, empty_skip=skip)
            _hits_equal(ctx.trace_rays(sc, rays), oracle.trace_rays(og, rays), f"edit {k} trace_kernel {tk}")
            ctx.set_tuning(**old)
        pos, yaw, pitch = cf.world_orbit(desc, k, 64)
        u, _ = v.make_uniforms(pos, yaw, pitch, 96, 64, frame=k, bounces=2, flags=0)
        got, _ = _render_one(ctx, sc, 960 + k, u, 48.0, 1, 96, 64)
        rad, alb, gb, _ = oracle.pathtrace(og, to_oracle_uniforms(u), 48.0, 1)
        assert_bits_equal(f"edit {k} radiance", got["radiance"], rad)
        assert_bits_equal(f"edit {k} gbuf", got["gbuf"], gb)
    with pytest.raises(v.VrtError):
        sc.edit(np.zeros((4, 4, 4), np.uint32), (62, 0, 0))  # outside the grid
    sc.release()
    plain = ctx.upload(vox)
    with pytest.raises(v.VrtError):
        plain.edit(np.zeros((1, 1, 1), np.uint32), (0, 0, 0))  # not editable
    plain.release()


@pytest.mark.parametrize("kind", ["empty", "corner-noncubic"])
def test_pathtrace_occupied_box_edges(ctx, kind):
    """Scenes at the edges of the path tracer's occupied boxes: an empty grid
    (the box is empty: every ray ends at setup and sees the sky) and a
    non-cubic grid whose occupied bricks fill only one corner region; frames
    equal the oracle's for every step-level variant form."""
    if kind == "empty":
        vox = np.zeros((64, 64, 64), np.uint32)
        pos, yaw, pitch = (2.0, 2.0, 2.0), 0.7, -0.3
    else:
        vox = v.generate_scene(v.SCENE_RANDOM_BRICKS, n=96, ny=32, nz=64, seed=17, p_brick=0.5, p_voxel=0.4)
        vox[:, :, 40:] = 0  # x >= 40
        vox[:, 16:, :] = 0  # y >= 16
        vox[:24, :, :] = 0  # z < 24
        pos, yaw, pitch = (0.5, 1.5, 0.5), 0.6, -0.2  # above the box, looking down into it
    sc = ctx.upload(vox)
    og = oracle.OracleGrid(vox)
    u, _ = v.make_uniforms(pos, yaw, pitch, 80, 48, frame=3, bounces=3, flags=0)
    rad, alb, gb, cnt = oracle.pathtrace(og, to_oracle_uniforms(u), 48.0, 1)
    old = ctx.tuning().as_dict()
    try:
        for tile in (0, 1):  # LDS and global-memory DDA
            ctx.set_tuning(pt_kernel=8, pt_waves=4, pt_tile=tile, pt_blocks=0)
            got, fr = _render_one(ctx, sc, 975 + tile, u, 48.0, 1, 80, 48)
            assert_bits_equal(f"{kind} tile {tile} gbuf", got["gbuf"], gb)
            assert_bits_equal(f"{kind} tile {tile} radiance", got["radiance"], rad)
            assert fr.rays == cnt.rays
            if kind == "empty":
                assert fr.brick_steps == 0 and fr.voxel_steps == 0 and (gb[..., 3] < 0).all()
    finally:
        ctx.set_tuning(**old)
    if kind != "empty":
        assert (gb[..., 3] >= 0).mean() > 0.1  # the corner region is in view
    sc.release()


def test_inplace_edit_grows_voxel_boxes(ctx):
    """An in-place edit that fills voxels of already occupied bricks outside
    their boxes of solid voxels.  The path tracer ends fine walks at each
    brick's box, so the edit must recompute the boxes (a stale one would let
    rays pass the new voxels): frames before and after equal the oracle's, and
    the camera sees the new voxels."""
    vox = np.zeros((64, 64, 64), np.uint32)
    vox[0:64, 0:4, 0:64] = 0x10508050  # floor
    for bz in range(2, 6):
        for bx in range(2, 6):
            vox[bz * 8, 24, bx * 8] = 0x10B04040  # one voxel in each brick's low corner
    sc = ctx.upload(vox, editable=True)
    # looking along +z and down at the bricks (32, 28, 32) from (32, 48, -8), voxel units
    u, _ = v.make_uniforms((2.0, 3.0, -0.5), 0.0, math.atan2(-1.25, 2.5), 96, 64, frame=1, bounces=2, flags=0)
    frames = []
    for k in range(2):
        if k == 1:
            slab = np.full((32, 4, 32), 0x10B0B040, np.uint32)  # y 28..31 of bricks bx, bz 2..5, by 3
            vox[16:48, 28:32, 16:48] = slab
            before = sc.edit_counts()
            sc.edit(slab, (16, 28, 16))
            assert sc.edit_counts() == (before[0] + 1, before[1])  # in place: occupancy unchanged
        got, fr = _render_one(ctx, sc, 970 + k, u, 48.0, 1, 96, 64)
        og = oracle.OracleGrid(vox)
        rad, alb, gb, cnt = oracle.pathtrace(og, to_oracle_uniforms(u), 48.0, 1)
        assert_bits_equal(f"box edit {k} gbuf", got["gbuf"], gb)
        assert_bits_equal(f"box edit {k} radiance", got["radiance"], rad)
        assert fr.rays == cnt.rays
        frames.append(got["gbuf"])
    changed = np.any(frames[0] != frames[1], axis=-1).mean()
    assert changed > 0.05, changed  # the slabs cover a good part of the view
    sc.release()


def test_large_scene_runs_the_global_dda(ctx):
    """A 512^3 grid: what the LDS form of the single-loop DDA stages (padded
    bitmap, prefix counts, row distances: 72 KiB) exceeds the LDS budget, so
    trace_kernel 4 and pt_kernel 4 run its global-memory form; hits and a
    small frame stay bit-exact, as for the explicit global variants."""
    rng = np.random.default_rng(21)
    vox = np.zeros((512, 512, 512), np.uint32)
    for _ in range(40):
        x, y, z = (int(c) for c in rng.integers(0, 480, 3))
        s = int(rng.integers(4, 32))
        vox[z:z + s, y:y + s, x:x + s] = int(rng.integers(1, 1 << 24)) | 1
    vox[:, :16, :] = 0x00406080 | 1
    print("large scene: 512^3, 40 blocks + floor", flush=True)
    sc = ctx.upload(vox)
    og = oracle.OracleGrid(vox)
    rays = cf.traverse_rays(1 << 16, n=512, seed=4)
    want = oracle.trace_rays(og, rays)
    old = ctx.tuning().as_dict()
    try:
        for tuning in (dict(trace_kernel=4), dict(trace_kernel=6, dda_batch=16), dict(trace_kernel=0)):
            ctx.set_tuning(**tuning)
            _hits_equal(ctx.trace_rays(sc, rays), want, f"512^3 {tuning}")
        u, _ = v.make_uniforms((16.0, 6.0, 3.0), 0.3, -0.1, 96, 64, frame=3, bounces=2, flags=0)
        rad, alb, gb, _ = oracle.pathtrace(og, to_oracle_uniforms(u), 48.0, 1)
        for tuning in (dict(pt_kernel=4, pt_waves=5), dict(pt_kernel=7)):
            ctx.set_tuning(**tuning)
            got, _ = _render_one(ctx, sc, 990, u, 48.0, 1, 96, 64)
            assert_bits_equal(f"512^3 radiance {tuning}", got["radiance"], rad)
            assert_bits_equal(f"512^3 gbuf {tuning}", got["gbuf"], gb)
    finally:
        ctx.set_tuning(**old)
        sc.release()
