"""vrt_trace_rays_device_noalloc (the allocation-free traversal entry point)
against vrt_trace_rays_device through the C ABI: two calls in a row, the
second with fewer rays, for the lane and the persistent (ray-queue) kernels;
hits equal bit for bit, the tail of the hit buffer beyond n untouched, and
an n beyond the maximum rejected before any device work."""
import numpy as np
import pytest

import voxelrt_amd as v
from voxelrt_amd import configs as cf

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scene(ctx):
    sc = ctx.upload(cf.traverse_scene(0.25, n=128, seed=3))
    yield sc
    sc.release()


@pytest.mark.parametrize("tuning", [dict(trace_kernel=5, trace_blocks=512, trace_steps=1, trace_min_refill=1),
                                    dict(trace_kernel=5, trace_blocks=1024, trace_steps=8, trace_min_refill=16,
                                         dda_batch=16),
                                    dict(trace_kernel=4), dict(trace_kernel=6),
                                    dict(trace_kernel=7, trace_blocks=512, trace_min_refill=16, trace_steps=32)],
                         ids=["persist-fast-s1", "persist-fast", "lane-fast", "lane-global", "wave"])
def test_noalloc_equals_trace_rays_device(ctx, scene, tuning):
    old = ctx.tuning().as_dict()
    ctx.set_tuning(**tuning)
    n0 = 1 << 18
    d_rays = ctx.device_array(n0 * 24)
    d_ref = ctx.device_array(n0 * 16)
    d_out = ctx.device_array(n0 * 16)
    try:
        for call, n in enumerate((n0, 1 << 15)):
            rays = cf.traverse_rays(n, n=128, seed=50 + call)
            d_rays.upload(rays)
            sentinel = np.full(n0 * 4, 0x5A5A5A5A, np.uint32)
            d_out.upload(sentinel)
            ctx.trace_rays_device(scene, d_rays.ptr, n, d_ref.ptr)
            ctx.trace_rays_device_noalloc(scene, d_rays.ptr, n, d_out.ptr)
            want = d_ref.download(np.zeros(n, v.HIT_DTYPE))
            got = d_out.download(np.zeros(n0 * 4, np.uint32))
            assert np.array_equal(got[:n * 4], want.view(np.uint32)), (call, tuning)
            assert np.all(got[n * 4:] == 0x5A5A5A5A), (call, tuning)
            assert (want["t"] >= 0).mean() > 0.3
    finally:
        ctx.set_tuning(**old)
        for d in (d_rays, d_ref, d_out):
            d.free()


def test_noalloc_rejects_too_many_rays(ctx, scene):
    d = ctx.device_array(1024)
    try:
        with pytest.raises(v.VrtError):
            ctx.trace_rays_device_noalloc(scene, d.ptr, v.NOALLOC_MAX_RAYS + 1, d.ptr)
    finally:
        d.free()
