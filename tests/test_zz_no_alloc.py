"""bvht_trace_camera_bounce_noalloc (the camera pass that allocates nothing,
its scratch a fixed array of the library) against bvht_trace_camera_bounce:
hits, bounce rays, source indices and counts bit for bit, on two calls in a
row with the second frame smaller than the first, and its size limits."""
import numpy as np
import pytest

import bvhtracer_amd as B
from bvhtracer_amd import _lib as L
from bvhtracer_amd import scenes
from helpers import comb_tree, upload

pytestmark = pytest.mark.gpu


def frame_buffers(ctx, n):
    bufs = [ctx.alloc(n * B.HIT_DTYPE.itemsize), ctx.alloc(n * B.RAY_DTYPE.itemsize), ctx.alloc(n * 4),
            ctx.alloc(4)]
    for b, size in zip(bufs, (n * 24, n * 32, n * 4, 4)):
        b.upload(np.full(size, 0xA5, np.uint8))  # anything left unwritten shows
    return bufs


def camera_pass(ctx, fn, bvh, cam, W, H, order, seed, bufs):
    cam = np.ascontiguousarray(cam, B.CAMERA_DTYPE).reshape(1)
    L.check(fn(ctx.handle, bvh.handle, L.ptr(cam), W, H, order, seed, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr,
               bufs[3].ptr))


def read(bufs, n):
    nb = int(bufs[3].download(np.uint32, 1)[0])
    assert 0 < nb <= n
    return (nb, bufs[0].download(np.uint8, n * 24).tobytes(), bufs[1].download(np.uint8, nb * 32).tobytes(),
            bufs[2].download(np.uint8, nb * 4).tobytes())


def test_noalloc_camera_pass_equals_the_allocating_pass():
    ctx = B.Context(0)
    try:
        lib = L.lib()
        mesh = scenes.icosphere(64, radius=32.0)
        bvh = B.BvhBuilder(B.BUILD_SAH, 2, ctx=ctx).build_for(mesh.copy())
        frames = [(1920, 1080, B.ORDER_TILES, 11), (640, 360, B.ORDER_ROWS, 12)]
        refs = []
        for W, H, order, seed in frames:
            bufs = frame_buffers(ctx, W * H)
            camera_pass(ctx, lib.bvht_trace_camera_bounce, bvh, scenes.sphere_camera(32.0, W, H), W, H, order, seed,
                        bufs)
            refs.append(read(bufs, W * H))
        # two allocation-free calls in a row, the second on a smaller frame
        outs = [frame_buffers(ctx, W * H) for W, H, _, _ in frames]
        for (W, H, order, seed), bufs in zip(frames, outs):
            camera_pass(ctx, lib.bvht_trace_camera_bounce_noalloc, bvh, scenes.sphere_camera(32.0, W, H), W, H,
                        order, seed, bufs)
        ctx.sync()
        for (W, H, _, _), bufs, ref in zip(frames, outs, refs):
            got = read(bufs, W * H)
            assert got[0] == ref[0], "bounce ray count"
            assert got[1] == ref[1], "hits"
            assert got[2] == ref[2], "bounce rays"
            assert got[3] == ref[3], "source indices"
    finally:
        ctx.close()


def test_noalloc_camera_pass_limits():
    ctx = B.Context(0)
    try:
        lib = L.lib()
        mesh = scenes.icosphere(8, radius=32.0)
        bvh = B.BvhBuilder(B.BUILD_SAH, 2, ctx=ctx).build_for(mesh.copy())
        cam = np.ascontiguousarray(scenes.sphere_camera(32.0, 2048, 2056), B.CAMERA_DTYPE).reshape(1)
        small = frame_buffers(ctx, 64)
        # one row of pixels beyond the fixed scratch: refused before any launch
        rc = lib.bvht_trace_camera_bounce_noalloc(ctx.handle, bvh.handle, L.ptr(cam), 2048, 2056, B.ORDER_TILES, 1,
                                                  small[0].ptr, small[1].ptr, small[2].ptr, small[3].ptr)
        assert rc == L.EINVAL and b"pixels" in lib.bvht_last_error()
        # a tree whose stack does not fit 32 LDS entries: refused
        nodes, used, tris = comb_tree(40)
        deep = upload(ctx, nodes, used, tris)
        rc = lib.bvht_trace_camera_bounce_noalloc(ctx.handle, deep.handle, L.ptr(cam), 8, 8, B.ORDER_TILES, 1,
                                                  small[0].ptr, small[1].ptr, small[2].ptr, small[3].ptr)
        assert rc == L.EINVAL and b"stack" in lib.bvht_last_error()
        # an empty frame writes a zero count
        camera_pass(ctx, lib.bvht_trace_camera_bounce_noalloc, bvh, cam, 0, 0, B.ORDER_ROWS, 1, small)
        assert int(small[3].download(np.uint32, 1)[0]) == 0
    finally:
        ctx.close()
