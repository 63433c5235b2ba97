"""Failure reporting and buffer lifetime on the GPU: a traversal stack limit
that a tree exceeds is reported as BVHT_ESTACK (every kernel stays in
bounds), a graph captured on a context pins the context's buffers, and the
batched TLAS rebuild leaves the per-scene traces a sufficient stack bound."""
import numpy as np
import pytest

import bvhtracer_amd as B
from bvhtracer_amd import _lib as L
from bvhtracer_amd import scenes
from helpers import assert_hits_equal, comb_tree, upload

pytestmark = pytest.mark.gpu


def comb_rays(n=1024):
    k = np.arange(n)
    rays = np.zeros(n, B.RAY_DTYPE)
    rays["o"] = np.stack([np.full(n, -5.0), (k % 32) / 32.0 - 0.5, (k // 32) / 32.0 - 0.5], 1)
    rays["d"] = (1.0, 1e-3, -1e-3)
    rays["tmax"] = np.finfo(np.float32).max
    return rays


def test_stack_limit_reports_estack_and_recovers(orc):
    """A ray along a 150-level comb pushes 149 far children.  With the stack
    capped at 40 entries the pushes beyond 40 are dropped inside the spill
    buffer sized for 40, and the read-back returns BVHT_ESTACK; the flag is
    cleared by that report, and with the limit lifted the same context gives
    the oracle's hits."""
    ctx = B.Context(0)
    try:
        nodes, used, tris = comb_tree(150)
        bvh = upload(ctx, nodes, used, tris)
        rays = comb_rays()
        ref = orc.bvh_trace(nodes, tris, rays.view(orc.RAY_DTYPE))
        ctx.set_stack_limit(40)
        with pytest.raises(B.BvhtError) as e:
            bvh.intersect(rays)
        assert e.value.code == L.ESTACK
        ctx.sync()  # the report cleared the status
        ctx.set_stack_limit(0)
        assert_hits_equal(bvh.intersect(rays), ref, "comb after the limit is lifted")
        ctx.sync()
        # the same through a one-object TLAS (the stack bound adds the TLAS depth)
        tlas = B.TlasBuilder(ctx=ctx).build_for([B.SceneObject(0)], [bvh])
        ctx.set_stack_limit(40)
        with pytest.raises(B.BvhtError) as e:
            tlas.intersect(rays)
        assert e.value.code == L.ESTACK
        ctx.set_stack_limit(0)
        # identity transform: the object-space ray is the world ray, object 0
        assert_hits_equal(tlas.intersect(rays), ref, "comb TLAS")
        # a limit the tree fits in changes nothing
        ctx.set_stack_limit(149)
        assert_hits_equal(bvh.intersect(rays), ref, "comb at a sufficient limit")
        ctx.sync()
    finally:
        ctx.close()


def test_graph_pins_context_buffers():
    """While a graph captured on a context lives, a call that would have to
    grow the context scratch is refused (BVHT_EINVAL) instead of freeing
    memory the graph still launches on; the graph keeps replaying, and after
    it is destroyed the larger call runs."""
    ctx = B.Context(0)
    try:
        mesh = scenes.icosphere(16, radius=3.0, seed=5)
        bvh = B.BvhBuilder(B.BUILD_SAH, 2, ctx=ctx).build_for(mesh.copy())
        cam = scenes.camera(eye=(0, 0, -12), forward=(0, 0, 1), right=(1, 0, 0), up=(0, 1, 0), fovy_deg=40)
        small = B.trace_camera_bounce(ctx, bvh, cam, 128, 128, B.ORDER_TILES, 3)
        eager = small[0].download(B.HIT_DTYPE, 128 * 128)

        def enqueue():
            B.trace_camera_bounce(ctx, bvh, cam, 128, 128, B.ORDER_TILES, 3, *small)

        g = B.Graph.capture(ctx, enqueue)
        big = [ctx.alloc(1024 * 1024 * B.HIT_DTYPE.itemsize), ctx.alloc(1024 * 1024 * B.RAY_DTYPE.itemsize),
               ctx.alloc(1024 * 1024 * 4), ctx.alloc(4)]
        with pytest.raises(B.BvhtError) as e:
            B.trace_camera_bounce(ctx, bvh, cam, 1024, 1024, B.ORDER_TILES, 3, *big)
        assert e.value.code == L.EINVAL and "graph" in str(e.value)
        for _ in range(3):
            small[0].upload(np.zeros(128 * 128, B.HIT_DTYPE))
            g.launch()
            assert small[0].download(B.HIT_DTYPE, 128 * 128).tobytes() == eager.tobytes()
        g.close()
        B.trace_camera_bounce(ctx, bvh, cam, 1024, 1024, B.ORDER_TILES, 3, *big)
        ctx.sync()
    finally:
        ctx.close()


def test_batch_rebuild_leaves_a_sufficient_stack_bound(ctx, orc):
    """Objects first in a compact cluster (a shallow TLAS, its depth read
    back), then moved onto a chain with growing gaps (a TLAS of one level per
    object) and rebuilt only by the batch: the per-scene trace that follows,
    with no synchronous rebuild, gives the oracle's hits on the chain."""
    n = 18
    mesh = scenes.icosphere(6, radius=1.0, seed=3)
    nodes, used, perm = orc.bvh_build(mesh, False, 2)
    model = B.Bvh.from_arrays(nodes, used, perm, ctx=ctx)
    M, Minv = scenes.random_affines(n, 91, trans_range=3.0)
    objs = [B.SceneObject(0, M[k].reshape(4, 4), Minv[k]) for k in range(n)]
    tlas = B.TlasBuilder(ctx=ctx).build_for(objs, [model])
    gaps = 3.0 * 2.0 ** np.arange(n)  # each gap exceeds all before it: one level per object
    xs = np.cumsum(gaps) - gaps.sum() / 2
    chain = []
    for k in range(n):
        m = np.eye(4, dtype=np.float32)
        m[3, 0] = xs[k]
        mi = np.eye(4, dtype=np.float32)
        mi[3, 0] = -xs[k]
        chain.append(B.SceneObject(0, m, mi))
    inst = B.api._instances(chain)
    B.api.check(L.lib().bvht_tlas_set_objects(ctx.handle, tlas.handle, L.ptr(inst), 0, n))
    batch = B.Batch(ctx, [tlas])
    batch.rebuild()
    y, z = np.meshgrid(np.linspace(-0.9, 0.9, 32), np.linspace(-0.9, 0.9, 32))
    rays = np.zeros(2 * y.size, B.RAY_DTYPE)
    rays["o"][: y.size] = np.stack([np.full(y.size, xs[0] - 100.0), y.ravel(), z.ravel()], 1)
    rays["d"][: y.size] = (1.0, 0.0, 0.0)
    rays["o"][y.size:] = np.stack([np.full(y.size, xs[-1] + 100.0), y.ravel(), z.ravel()], 1)
    rays["d"][y.size:] = (-1.0, 0.0, 0.0)
    rays["tmax"] = np.finfo(np.float32).max
    got = tlas.intersect(rays)
    bounds = np.stack([np.concatenate(orc.instance_bounds(o.m, nodes[0]["bmin"], nodes[0]["bmax"])) for o in chain])
    tnodes, tused = orc.tlas_build(bounds)
    g_nodes, g_used = tlas.download()
    assert g_nodes.tobytes() == tnodes.tobytes()
    oinst = np.zeros(n, orc.INSTANCE_DTYPE)
    oinst["minv"] = np.stack([o.minv for o in chain])
    ref = orc.tlas_trace(tnodes, oinst, [(nodes, perm)], rays.view(orc.RAY_DTYPE), nthreads=8)
    assert_hits_equal(got, ref, "per-scene trace after the batched rebuild")
    assert (got["prim"] != B.MISS).sum() > 500
    batch.close()


def test_camera_schedule_of_another_size_between_graph_replays():
    """A graph captures a scheduled 128x128 camera pass; between its replays
    an eager 64x48 pass rebuilds the context's wave order for its own size.
    Each replay then runs its groups in slot order and equals the eager
    128x128 pass bit for bit (hits, bounce rays and their sources)."""
    ctx = B.Context(0)
    try:
        mesh = scenes.icosphere(16, radius=3.0, seed=5)
        bvh = B.BvhBuilder(B.BUILD_SAH, 2, ctx=ctx).build_for(mesh.copy())
        cam = scenes.camera(eye=(0, 0, -12), forward=(0, 0, 1), right=(1, 0, 0), up=(0, 1, 0), fovy_deg=40)
        bufs = B.trace_camera_bounce(ctx, bvh, cam, 128, 128, B.ORDER_TILES, 3)
        B.trace_camera_bounce(ctx, bvh, cam, 128, 128, B.ORDER_TILES, 3, *bufs)  # scheduled from now on

        def snap():
            nb = int(bufs[3].download(np.uint32, 1)[0])
            return (bufs[0].download(B.HIT_DTYPE, 128 * 128).tobytes(), bufs[1].download(B.RAY_DTYPE, nb).tobytes(),
                    bufs[2].download(np.uint32, nb).tobytes())

        want = snap()

        def enqueue():
            B.trace_camera_bounce(ctx, bvh, cam, 128, 128, B.ORDER_TILES, 3, *bufs)

        g = B.Graph.capture(ctx, enqueue)
        small = [ctx.alloc(64 * 48 * B.HIT_DTYPE.itemsize), ctx.alloc(64 * 48 * B.RAY_DTYPE.itemsize),
                 ctx.alloc(64 * 48 * 4), ctx.alloc(4)]
        for _ in range(3):
            for _ in range(2):
                B.trace_camera_bounce(ctx, bvh, cam, 64, 48, B.ORDER_TILES, 3, *small)
            bufs[0].upload(np.zeros(128 * 128, B.HIT_DTYPE))
            g.launch()
            ctx.sync()
            assert snap() == want
        g.close()
    finally:
        ctx.close()
