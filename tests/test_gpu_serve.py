"""Serving (continuous batching, vrt_submit / vrt_poll / vrt_wait) through
the C ABI.

The claim under test: a request's result is the same whether it is rendered
alone (vrt_render of that one view) or served in a mix with other requests
arriving while it runs -- bit for bit, since the batched kernels run the
single-view per-pixel code.  Frames of one view keep their order (temporal
history), admission respects the HBM budget (eviction of idle view state,
failure of a request that can never fit), and cancel / release rules hold.
"""
import time

import numpy as np
import pytest

import voxelrt_amd as v
from voxelrt_amd import configs as cf
from tests.chain import assert_bits_equal

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def world():
    descs = [cf.world_desc(64, 1), cf.world_desc(128, 2), dict(cf.world_desc(64, 3), ny=32)]
    return descs, [v.generate_scene(**d) for d in descs]


def _request_specs(n, seed):
    rng = np.random.default_rng(seed)
    specs = []
    for k in range(n):
        specs.append(dict(si=k % 3, w=int(rng.integers(16, 240)), h=int(rng.integers(8, 160)),
                          nb=int(rng.integers(0, 5)), flags=int(rng.choice([0, 1, 3])), fmt=k % 2,
                          spp=1 + (k % 3 == 2), pose=int(rng.integers(0, 64)), frame=int(rng.integers(0, 1000))))
    return specs


def _view(scenes, descs, s, vid, prev=None):
    pos, yaw, pitch = cf.world_orbit(descs[s["si"]], s["pose"], 64)
    pvp, ppos = prev if prev else (None, None)
    u, vp = v.make_uniforms(pos, yaw, pitch, s["w"], s["h"], frame=s["frame"], bounces=s["nb"],
                            flags=s["flags"], prev_view_proj=pvp, prev_cam_pos=ppos)
    return v.View(scenes[s["si"]].handle, u, 48.0, vid, s["spp"], s["fmt"]), (vp, pos)


def _hdr(s, ctx=None):
    """HDR destination: pageable numpy, or page-locked (vrt_host_alloc) when
    a context is given."""
    if ctx is not None:
        return ctx.host_array((s["h"], s["w"], 3), np.float32)
    return np.zeros((s["h"], s["w"], 3), np.float32)


def _alone(ctx, scenes, descs, specs, base):
    out = []
    for k, s in enumerate(specs):
        view, _ = _view(scenes, descs, s, base + k)
        hdr = _hdr(s)
        fr = ctx.render([view], [{"hdr": hdr}])
        out.append((hdr, fr[0].rays))
        ctx.release_view(base + k)
    return out


@pytest.mark.parametrize("mode", [0, 2], ids=["grid", "graph"])
def test_served_equals_alone(world, mode):
    descs, vox = world
    with v.Context(0) as ctx:
        ctx.set_tuning(pt_kernel=4, pt_waves=5, batch_mode=mode)
        ctx.serve_configure(streams=2, max_batch_views=8)
        scenes = [ctx.upload(x) for x in vox]
        specs = _request_specs(40, seed=5)
        # a long first request keeps both streams busy while the rest arrive
        specs.insert(0, dict(si=1, w=1280, h=720, nb=4, flags=3, fmt=0, spp=4, pose=3, frame=7))
        alone = _alone(ctx, scenes, descs, specs, 5000)
        tickets = []
        for k, s in enumerate(specs):
            view, _ = _view(scenes, descs, s, 6000 + k)
            tickets.append(ctx.submit(view))
            if k % 7 == 6:
                time.sleep(0.001)  # arrivals while earlier batches run
        batch_sizes = []
        for k, (t, s) in enumerate(zip(tickets, specs)):
            hdr = _hdr(s, ctx if mode == 2 else None)  # graph form: page-locked destinations
            res = ctx.wait(t, hdr, timeout_ms=120000)
            assert res is not None, f"request {k} not finished"
            assert_bits_equal(f"request {k} hdr", hdr, alone[k][0])
            assert res["rays"] == alone[k][1], (k, res["rays"], alone[k][1])
            tm = res["timing"]
            assert tm["submit_ms"] <= tm["launch_ms"] <= tm["done_ms"]
            assert tm["latency_ms"] > 0 and tm["gpu_ms"] > 0
            batch_sizes.append(tm["batch_views"])
        st = ctx.serve_stats()
        assert st["completed"] == len(specs) and st["failed"] == 0 and st["queued"] == 0
        assert st["batched_requests"] == len(specs)
        assert max(batch_sizes) > 1, batch_sizes  # requests were batched
        with pytest.raises(v.VrtError):
            ctx.poll(tickets[0])  # retired


def test_served_view_keeps_frame_order(world):
    """Five frames of one view (temporal + denoise), all submitted at once,
    equal the same five frames rendered one by one."""
    descs, vox = world
    with v.Context(0) as ctx:
        ctx.serve_configure(streams=3, max_batch_views=16)
        scenes = [ctx.upload(x) for x in vox]
        base = dict(si=0, w=120, h=72, nb=2, flags=3, fmt=1, spp=1)
        frames = [dict(base, pose=p, frame=f) for f, p in enumerate((0, 1, 2, 4, 7))]
        want, prev = [], None
        for s in frames:
            view, prev = _view(scenes, descs, s, 7100, prev)
            hdr = _hdr(s)
            ctx.render([view], [{"hdr": hdr}])
            want.append(hdr)
        tickets, prev = [], None
        other = _request_specs(12, seed=9)
        for k, s in enumerate(frames):
            view, prev = _view(scenes, descs, s, 7000, prev)
            tickets.append(ctx.submit(view))
            for j in range(2):  # other traffic between the frames
                o = other[2 * k + j]
                ov, _ = _view(scenes, descs, o, 7200 + 2 * k + j)
                tickets.append(ctx.submit(ov))
        got = {}
        for k, t in enumerate(tickets):
            if k % 3 == 0:
                hdr = _hdr(frames[k // 3])
                assert ctx.wait(t, hdr, timeout_ms=120000) is not None
                got[k // 3] = hdr
            else:
                assert ctx.wait(t, None, timeout_ms=120000) is not None
        for f in range(len(frames)):
            assert_bits_equal(f"frame {f}", got[f], want[f])


def test_cancel_and_release_rules(world):
    descs, vox = world
    with v.Context(0) as ctx:
        ctx.serve_configure(streams=1, max_batch_views=1)
        scenes = [ctx.upload(x) for x in vox]
        big = dict(si=1, w=1920, h=1080, nb=6, flags=3, fmt=0, spp=4, pose=5, frame=1)
        small = dict(si=0, w=32, h=32, nb=1, flags=0, fmt=0, spp=1, pose=1, frame=2)
        t_big = ctx.submit(_view(scenes, descs, big, 8000)[0])
        t_small = ctx.submit(_view(scenes, descs, small, 8001)[0])
        # one stream, one request per batch: the small request waits behind
        # the big one, so it can still be cancelled
        ctx.cancel(t_small)
        with pytest.raises(v.VrtError):
            ctx.poll(t_small)
        with pytest.raises(v.VrtError):
            scenes[1].release()  # in use by the big request
        with pytest.raises(v.VrtError):
            ctx.release_view(8000)  # queued or running
        assert ctx.wait(t_big, None, timeout_ms=120000) is not None
        with pytest.raises(v.VrtError):
            ctx.cancel(t_big)  # retired
        st = ctx.serve_stats()
        assert st["cancelled"] == 1 and st["completed"] == 1
        ctx.release_view(8000)
        scenes[1].release()


def test_admission_by_hbm_footprint(world):
    """A budget that holds the scene plus about two small views: serving
    evicts idle view state to admit new views, results stay exact, and a
    request larger than the whole budget fails with VRT_ERR_BUDGET."""
    descs, vox = world
    s0 = dict(si=0, w=128, h=128, nb=2, flags=3, fmt=1, spp=1, pose=2, frame=3)
    with v.Context(0) as ref:
        scenes = [ref.upload(vox[0])]
        want = _alone(ref, scenes, descs[:1], [s0], 100)[0][0]
        scene_bytes = scenes[0].nbytes()
    view_bytes = 128 * 128 * (8 + 4 + 32 + 32 + 16)
    result_bytes = 256 * 1024
    budget = scene_bytes + 2 * view_bytes + 3 * result_bytes
    with v.Context(0, hbm_budget=budget) as ctx:
        ctx.serve_configure(streams=1, max_batch_views=4)
        scenes = [ctx.upload(vox[0])]
        for k in range(5):
            hdr = _hdr(s0)
            t = ctx.submit(_view(scenes, descs[:1], s0, 200 + k)[0])
            assert ctx.wait(t, hdr, timeout_ms=120000) is not None
            assert_bits_equal(f"view {200 + k}", hdr, want)
        st = ctx.serve_stats()
        assert st["evictions"] >= 3, st
        assert ctx.stats()["hbm_resident"] <= budget
        huge = dict(s0, w=2048, h=2048)
        t = ctx.submit(_view(scenes, descs[:1], huge, 300)[0])
        with pytest.raises(v.VrtError) as e:
            ctx.wait(t, None, timeout_ms=120000)
        assert e.value.code == -4  # VRT_ERR_BUDGET
        assert ctx.serve_stats()["failed"] == 1


def test_tight_budget_keeps_the_pooled_result_buffer(world):
    """Admission under a budget where freeing the whole result pool would
    cover the new view's state but leave no room for a fresh result buffer:
    the request takes the pooled buffer that fits (it is kept out of the
    pool before room is made), evicts an idle view instead, and completes.

    Sizes (fmt rgba16f, 76 B/px of view state, results 12 B/px rounded to
    64 KiB): A and B 256x256 (state 4980736 B, result 786432 B), a big view
    512x288 (state 11206656 B, result 1769472 B, too large to serve B).
    Budget = scene + states of A, big and B + 393216 B of slack."""
    descs, vox = world
    with v.Context(0) as probe:
        sc = probe.upload(vox[0])
        scene_bytes = probe.stats()["hbm_resident"]
        sc.release()
    vs_small, vs_big, slack = 256 * 256 * 76, 512 * 288 * 76, 393216
    budget = scene_bytes + 2 * vs_small + vs_big + slack
    with v.Context(0, hbm_budget=budget) as ctx:
        ctx.serve_configure(streams=1, max_batch_views=1)
        scenes = [ctx.upload(vox[0])]
        assert ctx.stats()["hbm_resident"] == scene_bytes
        for k, (vid, w, h) in enumerate(((1, 256, 256), (2, 512, 288), (3, 256, 256))):
            s = dict(si=0, w=w, h=h, nb=1, flags=0, fmt=0, spp=1, pose=k, frame=k)
            view, _ = _view(scenes, descs, s, vid)
            t = ctx.submit(view)
            res = ctx.wait(t, np.zeros((h, w, 3), np.float32), timeout_ms=60000)
            assert res is not None, f"request {k} (view {vid}) failed"
        st = ctx.serve_stats()
        assert st["completed"] == 3 and st["failed"] == 0, st
        assert st["evictions"] == 1, st  # view 1's idle state made the room
        assert ctx.stats()["hbm_resident"] <= budget
