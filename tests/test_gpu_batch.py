"""Batched multi-scene path and hipGraph frames: every result equals the
per-scene trace (bit for bit) and the oracle."""
import numpy as np
import pytest

import bvhtracer_amd as B
from bvhtracer_amd import autotune, scenes
from helpers import assert_hits_equal

pytestmark = pytest.mark.gpu


def build_batch(ctx, n_scenes=6, seed=0x51, tri_range=(500, 20000), ray_range=(1000, 20000), mode=B.BUILD_SAH):
    work = scenes.batch_workload(n_scenes, seed=seed, tri_range=tri_range, ray_range=ray_range)
    models, tlases, origs = [], [], []
    for w in work:
        m = w["mesh"].copy()
        bvh = B.BvhBuilder(mode, 2 if mode == B.BUILD_SAH else 0, ctx=ctx).build_for(m)
        objs = [B.SceneObject(0, w["M"][k].reshape(4, 4), w["Minv"][k]) for k in range(w["M"].shape[0])]
        models.append(bvh)
        origs.append(m)
        tlases.append(B.TlasBuilder(ctx=ctx).build_for(objs, [bvh]))
    rays = np.concatenate([w["rays"] for w in work])
    offsets = np.concatenate([[0], np.cumsum([w["rays"].shape[0] for w in work])]).astype(np.uint32)
    return work, models, tlases, origs, rays, offsets


def plain_hits(ctx, tlases, rays, offsets):
    out = []
    for s, t in enumerate(tlases):
        out.append(t.intersect(rays[offsets[s]:offsets[s + 1]]))
    return np.concatenate(out)


def oracle_scene_hits(orc, bvh, tlas, objs_minv, rays):
    nodes, used, tris = bvh.download()
    tnodes, _ = tlas.download()
    inst = np.zeros(objs_minv.shape[0], orc.INSTANCE_DTYPE)
    inst["minv"] = objs_minv
    return orc.tlas_trace(tnodes, inst, [(nodes, tris)], rays.view(orc.RAY_DTYPE), nthreads=8)


def test_batched_trace_equals_plain_and_oracle(ctx, orc):
    work, models, tlases, origs, rays, offsets = build_batch(ctx)
    batch = B.Batch(ctx, tlases)
    batch.set_rays(offsets, 128)
    d_rays = ctx.to_device(rays)
    d_hits = ctx.alloc(rays.shape[0] * B.HIT_DTYPE.itemsize)
    batch.trace(d_rays, d_hits)
    got = d_hits.download(B.HIT_DTYPE, rays.shape[0])
    assert_hits_equal(got, plain_hits(ctx, tlases, rays, offsets), "batched vs per-scene")
    for s in range(len(work)):
        sl = slice(offsets[s], offsets[s + 1])
        ref = oracle_scene_hits(orc, models[s], tlases[s], work[s]["Minv"], rays[sl])
        assert_hits_equal(got[sl], ref, f"batched scene {s} vs oracle")
    assert (got["prim"] != B.MISS).sum() > 1000
    # the queue order differs from trace to trace: every ray's hit does not
    for _ in range(2):
        d_again = ctx.alloc(rays.shape[0] * B.HIT_DTYPE.itemsize)
        batch.trace(d_rays, d_again)
        assert d_again.download(B.HIT_DTYPE, rays.shape[0]).tobytes() == got.tobytes()


def test_every_batch_config_is_bit_identical(ctx):
    work, models, tlases, origs, rays, offsets = build_batch(ctx, n_scenes=4, seed=0x77)
    batch = B.Batch(ctx, tlases)
    d_rays = ctx.to_device(rays)
    d_hits = ctx.alloc(rays.shape[0] * B.HIT_DTYPE.itemsize)
    ref = plain_hits(ctx, tlases, rays, offsets)
    for wg, d, r in autotune.BATCH_CONFIGS:
        batch.set_rays(offsets, r)
        batch.config(wg, d, r)
        batch.trace(d_rays, d_hits)
        assert_hits_equal(d_hits.download(B.HIT_DTYPE, rays.shape[0]), ref, f"config {wg}x{d}x{r}")


@pytest.mark.parametrize("mode", [B.BUILD_SAH, B.BUILD_REFERENCE])
def test_batched_refit_and_rebuild_equal_per_scene(ctx, mode):
    """Both builders: the reference builder's leaves of hundreds of
    triangles go through the wave-per-leaf refit."""
    work, models, tlases, origs, rays, offsets = build_batch(ctx, n_scenes=5, seed=0x99, mode=mode)
    batch = B.Batch(ctx, tlases)
    for bvh, orig in zip(models, origs):
        bvh.refit(scenes.deform_wave(orig, 3))
    before = [t.download()[0].copy() for t in tlases]
    batch.refit()
    batch.rebuild()
    ctx.sync()
    got_nodes = [m.download()[0] for m in models]
    got_tlas = [t.download()[0] for t in tlases]
    for bvh in models:
        bvh.refit()
    for t in tlases:
        t.refresh_bounds()
        t.rebuild()
    for k, m in enumerate(models):
        assert m.download()[0].tobytes() == got_nodes[k].tobytes()
    for k, t in enumerate(tlases):
        assert t.download()[0].tobytes() == got_tlas[k].tobytes()
    assert any(b.tobytes() != g.tobytes() for b, g in zip(before, got_tlas))  # the deformation moved boxes


def test_graph_frame_replays_match_eager(ctx, orc):
    """Frame = device deformation (frame counter on device) + batched refit +
    batched rebuild + batched trace + counter increment, captured once and
    replayed three times with eager work and synchronisations in between;
    each replay equals the same frame computed eagerly through the
    per-scene API, and the last one equals the oracle."""
    work, models, tlases, origs, rays, offsets = build_batch(ctx, n_scenes=4, seed=0x33)
    batch = B.Batch(ctx, tlases)
    batch.set_rays(offsets, 128)
    d_origs = [ctx.to_device(o) for o in origs]
    d_frame = ctx.to_device(np.array([1], np.uint32))
    d_rays = ctx.to_device(rays)
    d_hits = ctx.alloc(rays.shape[0] * B.HIT_DTYPE.itemsize)
    amp, kx, w = 0.2, 4.0, 0.1

    def frame():
        for bvh, d_o in zip(models, d_origs):
            B.deform_wave(ctx, bvh, d_o, amp, kx, w, d_frame)
        batch.refit()
        batch.rebuild()
        batch.trace(d_rays, d_hits)
        B.counter_add(ctx, d_frame, 1)

    graph = B.Graph.capture(ctx, frame)
    eager_frame = ctx.to_device(np.array([0], np.uint32))
    for f in (1, 2, 3):
        graph.launch()
        ctx.sync()
        got = d_hits.download(B.HIT_DTYPE, rays.shape[0])
        assert int(d_frame.download(np.uint32, 1)[0]) == f + 1
        # eager recomputation of frame f through the per-scene API
        eager_frame.upload(np.array([f], np.uint32))
        for bvh, d_o in zip(models, d_origs):
            B.deform_wave(ctx, bvh, d_o, amp, kx, w, eager_frame)
            bvh.refit()
        for t in tlases:
            t.refresh_bounds()
            t.rebuild()
        ref = plain_hits(ctx, tlases, rays, offsets)
        assert_hits_equal(got, ref, f"graph replay {f}")
    sl = slice(offsets[0], offsets[1])
    assert_hits_equal(got[sl], oracle_scene_hits(orc, models[0], tlases[0], work[0]["Minv"], rays[sl]),
                      "graph frame vs oracle")
    graph.close()
