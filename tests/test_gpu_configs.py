"""The benchmark configurations at their own shapes, against the oracle.

  frame_1080  1920x1080 on the 256^3 frame scene (partial tile rows: 1080 is
              67.5 tiles of 16), three orbit frames through the whole chain
              (pathtrace -> temporal -> a-trous x3 -> HDR), with the path
              tracer the benchmark times (step-level refill, 4 waves per SIMD,
              the DDA in LDS, min_wait 12 / shade_min 24), its global-memory
              form, and the refill-tile forms (32x16 and 16x8 tiles);
  batch       the eight batch scenes (64^3 .. 512^3, two of each) with one
              view each at reduced size, bounces 1..6, mixed accumulation
              formats, rendered as ONE batched call (grid launches, the
              persistent tile queue, the captured graph) over two frames;
              every view's targets equal its own oracle chain;
  serve       a served mix over a 1024^3 scene (4 GiB grid) and two small
              ones: requests submitted while others run, each result's HDR
              equal to the oracle chain of its view.
Every comparison is bit for bit.
"""
import numpy as np
import pytest

import oracle
import voxelrt_amd as v
from voxelrt_amd import configs as cf
from tests.chain import OracleChain, assert_bits_equal, to_oracle_uniforms

pytestmark = pytest.mark.gpu

TARGETS = ("radiance", "albedo", "gbuf", "accum", "denoised", "hdr", "ldr")


@pytest.fixture(scope="module")
def frame1080(ctx):
    vox = cf.frame_1080_scene()
    sc = ctx.upload(vox)
    yield vox, sc, oracle.OracleGrid(vox)
    sc.release()


@pytest.mark.parametrize("tuning", [dict(pt_kernel=8, pt_waves=4, pt_tile=0, pt_order=2, pt_blocks=0,
                                         pt_min_refill=16, pt_shade_min=40),
                                    dict(pt_kernel=8, pt_waves=4, pt_tile=1, pt_order=2, pt_blocks=0,
                                         pt_min_refill=16, pt_shade_min=32),
                                    dict(pt_kernel=7, pt_waves=4, pt_tile=3, pt_order=2),
                                    dict(pt_kernel=7, pt_waves=4, pt_tile=5, pt_order=2)],
                         ids=["wave-lds", "wave-global", "refill32x16", "refill16x8"])
def test_frame_1080_chain_bitexact(ctx, frame1080, tuning):
    vox, sc, og = frame1080
    p = cf.FRAME_1080
    w, h = p["width"], p["height"]
    old = ctx.tuning().as_dict()
    ctx.set_tuning(**tuning)
    try:
        r = v.Renderer(ctx, sc, w, h, bounces=p["bounces"], temporal=True, denoise=True, fclamp=p["fclamp"],
                       max_history=p["max_history"], view_id=8100)
        chain = OracleChain(og, w, h, 0)
        for i in range(3):
            r.set_pose(*cf.world_orbit(cf.FRAME_1080_SCENE, i))
            print(f"[frame_1080] {tuning} frame {i}", flush=True)
            got = r.frame(TARGETS)
            want = chain.frame(to_oracle_uniforms(r.last_view.u), p["fclamp"])
            for k in TARGETS:
                assert_bits_equal(f"frame {i} {k}", got[k], want[k])
    finally:
        ctx.release_view(8100)
        ctx.set_tuning(**old)


@pytest.fixture(scope="module")
def batch_scenes(ctx):
    descs, _ = cf.batch_views(8)
    vox = [v.generate_scene(**d) for d in descs]
    print("[batch] scenes " + ", ".join(f"{d['n']}^3" for d in descs), flush=True)
    scenes = [ctx.upload(x) for x in vox]
    grids = [oracle.OracleGrid(x) for x in vox]
    yield descs, scenes, grids
    for sc in scenes:
        sc.release()


@pytest.mark.parametrize("mode,kernel", [(0, 7), (1, 7), (2, 7), (0, 8), (2, 8)],
                         ids=["grid", "queue", "graph", "wave-grid", "wave-graph"])
def test_batch_views_match_oracle(ctx, batch_scenes, mode, kernel):
    descs, scenes, grids = batch_scenes
    rng = np.random.default_rng(40 + mode)
    specs = []  # one view per scene: size, bounces 1..6, accumulation format
    for si in range(8):
        h = int(rng.integers(64, 271))
        w = min(480, int(round(h * [1.0, 4 / 3, 16 / 9, 21 / 9][si % 4])))
        specs.append(dict(si=si, w=w, h=h, nb=1 + si % 6, fmt=si % 2, seed=7717 * si))
    old = ctx.tuning().as_dict()
    if kernel == 8:  # the step-level refill kernel's batched form
        ctx.set_tuning(pt_kernel=8, pt_waves=4, pt_tile=0, batch_tile=0, batch_mode=mode, pt_blocks=0,
                       pt_min_refill=16, pt_shade_min=32)
    else:
        ctx.set_tuning(pt_kernel=7, pt_waves=4, pt_tile=5, batch_mode=mode)
    chains = [OracleChain(grids[s["si"]], s["w"], s["h"], s["fmt"]) for s in specs]
    prev = [(None, None)] * len(specs)
    try:
        for frame in range(2):
            views, outs = [], []
            for k, s in enumerate(specs):
                pos, yaw, pitch = cf.world_orbit(descs[s["si"]], 5 * k + frame, 64)
                u, vp = v.make_uniforms(pos, yaw, pitch, s["w"], s["h"], frame=frame + s["seed"], bounces=s["nb"],
                                        flags=3, prev_view_proj=prev[k][0], prev_cam_pos=prev[k][1])
                prev[k] = (vp, pos)
                views.append(v.View(scenes[s["si"]].handle, u, 48.0, 8200 + k, 1, s["fmt"]))
                h, w = s["h"], s["w"]
                outs.append({"radiance": np.zeros((h, w, 4), np.uint16), "albedo": np.zeros((h, w, 4), np.uint8),
                             "gbuf": np.zeros((h, w, 4), np.float32),
                             "accum": np.zeros((h, w, 4), np.uint16 if s["fmt"] == 0 else np.float32),
                             "denoised": np.zeros((h, w, 4), np.uint16), "hdr": np.zeros((h, w, 3), np.float32)})
            print(f"[batch] mode {mode} frame {frame}: {len(views)} views", flush=True)
            ctx.render(views, outs)
            for k, s in enumerate(specs):
                want = chains[k].frame(to_oracle_uniforms(views[k].u), 48.0)
                for name in outs[k]:
                    assert_bits_equal(f"mode {mode} frame {frame} view {k} ({descs[s['si']]['n']}^3) {name}",
                                      outs[k][name], want[name])
    finally:
        ctx.set_tuning(**old)
        for k in range(len(specs)):
            ctx.release_view(8200 + k)


def test_served_mix_with_1024_scene_matches_oracle():
    descs = [cf.world_desc(1024, 315), cf.world_desc(64, 1), cf.world_desc(128, 2)]
    print("[serve] generating 1024^3 + 64^3 + 128^3", flush=True)
    vox = [v.generate_scene(**d) for d in descs]
    grids = [oracle.OracleGrid(x) for x in vox]
    # (scene, w, h, spp, bounces, flags, fmt, view); the 1024^3 view runs 3
    # frames in order, the others interleave with it
    specs = [(0, 160, 90, 1, 2, 3, 0, 0), (1, 96, 54, 2, 3, 3, 1, 1), (2, 128, 72, 1, 1, 1, 0, 2),
             (0, 160, 90, 1, 2, 3, 0, 0), (1, 64, 36, 4, 4, 0, 0, 3), (0, 160, 90, 1, 2, 3, 0, 0),
             (0, 200, 112, 2, 4, 3, 1, 4), (2, 48, 48, 3, 2, 3, 0, 5)]
    with v.Context(0) as ctx:
        ctx.set_tuning(pt_kernel=8, pt_waves=4, batch_tile=0, batch_mode=2)
        ctx.serve_configure(streams=2, max_batch_views=4)
        scenes = [ctx.upload(x) for x in vox]
        chains, prev, frames, tickets, wants = {}, {}, {}, [], []
        for k, (si, w, h, spp, nb, flags, fmt, vid) in enumerate(specs):
            f = frames.get(vid, 0)
            frames[vid] = f + 1
            pos, yaw, pitch = cf.world_orbit(descs[si], 3 * vid + f, 64)
            pvp, ppos = prev.get(vid, (None, None))
            u, vp = v.make_uniforms(pos, yaw, pitch, w, h, frame=100 * vid + f, bounces=nb, flags=flags,
                                    prev_view_proj=pvp, prev_cam_pos=ppos)
            prev[vid] = (vp, pos)
            view = v.View(scenes[si].handle, u, 48.0, 8300 + vid, spp, fmt)
            tickets.append(ctx.submit(view))
            if vid not in chains:
                chains[vid] = OracleChain(grids[si], w, h, fmt)
            wants.append(chains[vid].frame(to_oracle_uniforms(u), 48.0, spp))
        for k, (t, s) in enumerate(zip(tickets, specs)):
            hdr = ctx.host_array((s[2], s[1], 3), np.float32)
            res = ctx.wait(t, hdr, timeout_ms=120000)
            assert res is not None, f"request {k} not finished"
            assert_bits_equal(f"request {k} (scene {descs[s[0]]['n']}^3) hdr", hdr, wants[k]["hdr"])
            assert res["rays"] == wants[k]["counters"].rays, (k, res["rays"], wants[k]["counters"].rays)
        st = ctx.serve_stats()
        assert st["completed"] == len(specs) and st["failed"] == 0
