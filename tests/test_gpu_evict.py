"""Evictable scenes (VRT_SCENE_EVICTABLE) under an HBM budget that holds two
of three scenes.

The claim under test: when the budget is needed, the least recently used
idle evictable scene leaves the device (its buffers to host memory) and any
later use brings it back unchanged -- hits, frames and served results stay
bit-exact against the oracle and against a roomy context; a scene in use
(queued served requests, a call in progress) is never evicted; scenes
without the flag are never evicted; the allocation-free entry point rejects
an evicted scene until it is restored."""
import numpy as np
import pytest

import oracle
import voxelrt_amd as v
from voxelrt_amd import configs as cf
from tests.chain import assert_bits_equal, to_oracle_uniforms

pytestmark = pytest.mark.gpu

W, H = 96, 64


@pytest.fixture(scope="module")
def world():
    descs = [cf.world_desc(256, 31), cf.world_desc(256, 32), cf.world_desc(256, 33)]
    return descs, [v.generate_scene(**d) for d in descs]


def _scene_bytes(vox):
    with v.Context(0) as ctx:
        sc = ctx.upload(vox, evictable=True)
        b = sc.nbytes()
        sc.release()
    return b


def _view(sc, desc, vid, pose):
    u, _ = v.make_uniforms(*cf.world_orbit(desc, pose, 64), W, H, frame=pose, bounces=2, flags=0)
    return v.View(sc.handle, u, 48.0, vid, 1, 0)


def _hits_equal(got, want, name):
    for f in ("voxel", "face", "brick_steps", "voxel_steps"):
        assert np.array_equal(got[f], want[f]), f"{name} {f}"
    assert np.array_equal(got["t"].view(np.uint32), want["t"].view(np.uint32)), f"{name} t"


def test_lru_eviction_and_restore(world):
    descs, vox = world
    sb = [_scene_bytes(x) for x in vox]
    # two scenes and a few small view states fit, three scenes do not
    budget = max(sb) + sorted(sb)[1] + (2 << 20)
    assert budget < sum(sb)
    with v.Context(0, hbm_budget=budget) as ctx:
        a, b = ctx.upload(vox[0], evictable=True), ctx.upload(vox[1], evictable=True)
        assert a.resident() and b.resident()
        c = ctx.upload(vox[2], evictable=True)  # evicts a, the least recently used
        assert not a.resident() and b.resident() and c.resident()
        st = ctx.stats()
        assert st["scene_evictions"] == 1 and st["scene_restores"] == 0
        rays = cf.traverse_rays(1 << 14, n=256, seed=40)
        # using a restores it and evicts b (now the least recently used)
        _hits_equal(ctx.trace_rays(a, rays), oracle.trace_rays(oracle.OracleGrid(vox[0]), rays), "a restored")
        assert a.resident() and not b.resident() and c.resident()
        # frames of every scene in turn: each brings its scene back
        for k, (sc, d, x) in enumerate([(b, descs[1], vox[1]), (c, descs[2], vox[2]), (a, descs[0], vox[0]),
                                        (b, descs[1], vox[1])]):
            view = _view(sc, d, 70 + k, 3 * k)
            fr = {"radiance": np.zeros((H, W, 4), np.uint16), "gbuf": np.zeros((H, W, 4), np.float32)}
            ctx.render([view], [fr])
            rad, _, gb, _ = oracle.pathtrace(oracle.OracleGrid(x), to_oracle_uniforms(view.u), 48.0, 1)
            assert_bits_equal(f"frame {k} radiance", fr["radiance"], rad)
            assert_bits_equal(f"frame {k} gbuf", fr["gbuf"], gb)
            assert sc.resident()
            ctx.release_view(70 + k)
        st = ctx.stats()
        assert st["scene_evictions"] >= 4 and st["scene_restores"] >= 3, st
        # the allocation-free entry point must not restore (that allocates)
        victim = next(s for s in (a, b, c) if not s.resident())
        d_rays = ctx.device_array(rays.nbytes).upload(rays)
        d_hits = ctx.device_array(len(rays) * 16)
        with pytest.raises(v.VrtError):
            ctx.trace_rays_device_noalloc(victim, d_rays.ptr, len(rays), d_hits.ptr)
        victim.restore()
        assert victim.resident()
        for s in (a, b, c):
            s.release()


def test_resident_scenes_are_not_evicted(world):
    descs, vox = world
    sb = [_scene_bytes(x) for x in vox]
    with v.Context(0, hbm_budget=max(sb) + sorted(sb)[1] + (2 << 20)) as ctx:
        a = ctx.upload(vox[0])
        b = ctx.upload(vox[1])
        with pytest.raises(v.VrtError):
            ctx.upload(vox[2], evictable=True)  # no evictable scene to make room with
        assert a.resident() and b.resident()
        assert ctx.stats()["scene_evictions"] == 0
        a.release()
        b.release()


def test_served_mix_with_evictions_equals_alone(world):
    """Requests for views of three scenes in rotation, served under a budget
    that holds two: scenes of queued requests stay, idle ones are evicted and
    restored as requests arrive; every result equals the same view rendered
    alone on a roomy context."""
    descs, vox = world
    sb = [_scene_bytes(x) for x in vox]
    specs = [(k % 3, 3 * k + 1) for k in range(18)]
    with v.Context(0) as roomy:
        scenes = [roomy.upload(x) for x in vox]
        alone = []
        for k, (si, pose) in enumerate(specs):
            hdr = np.zeros((H, W, 3), np.float32)
            roomy.render([_view(scenes[si], descs[si], 500 + k, pose)], [{"hdr": hdr}])
            alone.append(hdr)
            roomy.release_view(500 + k)
    with v.Context(0, hbm_budget=max(sb) + sorted(sb)[1] + (4 << 20)) as ctx:
        ctx.serve_configure(streams=2, max_batch_views=2)
        scenes = [ctx.upload(x, evictable=True) for x in vox]
        for k, (si, pose) in enumerate(specs):
            # one request at a time per scene group: submit, wait, so the other
            # scenes are idle and evictable meanwhile
            t = ctx.submit(_view(scenes[si], descs[si], 600 + k, pose))
            hdr = np.zeros((H, W, 3), np.float32)
            ctx.wait(t, hdr)
            assert_bits_equal(f"served request {k} (scene {si})", hdr, alone[k])
        st = ctx.stats()
        assert st["scene_evictions"] >= 6 and st["scene_restores"] >= 6, st
        for s in scenes:
            s.release()
