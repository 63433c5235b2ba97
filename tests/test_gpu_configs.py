"""GPU parity at the benchmark workloads' full sizes, each against the oracle
on identical buffers (GPU-built trees and GPU-made rays and vertices are
copied back; the oracle refits, rebuilds and traces them itself):

- the 1,310,720-triangle displaced icosphere: the fused camera pass at
  1920x1080 (camera rays, primary trace, grouped bounce rays) and the bounce
  trace, every hit, including primitives >= 2^20 whose packed
  InstancePrimitiveIndex word aliases (query/intersection.rs:41-47);
- 16 instances of a 200,000-triangle icosphere at 1280x720: deform, refit,
  bounds, TLAS rebuild, primary and bounce traces for frames 0 and 1
  eagerly, then frames 2..63 replayed as one captured frame graph and frame
  63 checked (examples/sixteen_armadillos.rs:132-163, scene/tlas.rs:123-250);
- the 64 uneven scenes in one persistent batched trace, after a deform +
  refit + rebuild frame, every scene against its own oracle trace."""
import os

import numpy as np
import pytest

import bvhtracer_amd as B
from bvhtracer_amd import _lib as L
from bvhtracer_amd import scenes
from helpers import assert_hits_equal, assert_nodes_equal

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


def deformed_as_formula(got, orig, frame, k=2):
    """The device's deformation y + 0.2 sin(4x + 0.1 frame) against the
    float32 formula, within k ulps of the larger operand: sinf on the device
    and numpy's float32 sin may round differently, and where the sum cancels
    the result's own ulp is far below the operands' rounding.  (The
    deformation is input synthesis: the oracle refits and traces the
    vertices the device made.)"""
    want = scenes.deform_wave(orig, frame)
    scale = np.maximum(np.maximum(np.abs(orig), np.abs(want)), np.float32(0.2))
    return bool((np.abs(got.astype(np.float64) - want) <= k * np.spacing(scale).astype(np.float64)).all())


def tile_order(width, height):
    """Row-order index of ray j of the 8x8-tile order (renderer.rs:353-361)."""
    j = np.arange(width * height)
    tile, w = j >> 6, j & 63
    tiles_x = width // 8
    x = (tile % tiles_x) * 8 + (w & 7)
    y = (tile // tiles_x) * 8 + (w >> 3)
    return y * width + x


def test_sphere_1m_camera_pass_and_bounce_every_hit(ctx, orc):
    W, H = 1920, 1080
    n = W * H
    mesh = scenes.icosphere(256, radius=32.0)
    assert mesh.shape[0] == 1310720
    bvh = B.BvhBuilder(B.BUILD_SAH, 2, ctx=ctx).build_for(mesh.copy())
    nodes, used, tris = bvh.download()
    cam = scenes.sphere_camera(32.0, W, H)
    seed = 0xB0B + 7
    d_hits, d_out, d_src, d_count = B.trace_camera_bounce(ctx, bvh, cam, W, H, B.ORDER_TILES, seed)
    d_bhits = ctx.alloc(n * B.HIT_DTYPE.itemsize)
    L.check(L.lib().bvht_trace_blas_count(ctx.handle, bvh.handle, d_out.ptr, d_count.ptr, n, d_bhits.ptr, None))
    nb = int(d_count.download(np.uint32, 1)[0])
    hits = d_hits.download(B.HIT_DTYPE, n)
    brays = d_out.download(B.RAY_DTYPE, nb)
    src = d_src.download(np.uint32, nb)
    bhits = d_bhits.download(B.HIT_DTYPE, nb)
    # the camera rays are the oracle's (Camera::get_ray_world), in tile order
    rays = orc.raygen(cam, W, H)[tile_order(W, H)]
    ref = orc.bvh_trace(nodes, tris, rays, nthreads=THREADS)
    assert_hits_equal(hits, ref, "1.3M sphere, 1920x1080 camera rays")
    found = hits["prim"] != B.MISS
    assert nb == found.sum() > 1_000_000
    # every bounce ray comes from a distinct hit of the frame, grouped by bin
    assert np.unique(src).size == nb and found[src].all()
    ref_b = orc.bvh_trace(nodes, tris, brays.view(orc.RAY_DTYPE), nthreads=THREADS)
    assert_hits_equal(bhits, ref_b, "1.3M sphere, bounce rays")
    # primitive ids beyond 20 bits alias in the reference's packed word
    for h in (hits, bhits):
        big = (h["prim"] != B.MISS) & (h["prim"] >= (1 << 20))
        assert big.sum() > 100
        assert np.array_equal(h["packed"][big], h["prim"][big] & 0xFFFFF)
        assert (h["packed"][big] < h["prim"][big]).all()


def test_instanced_16x200k_frames_0_1_63(ctx, orc):
    W, H = 1280, 720
    n = W * H
    mesh = scenes.icosphere(100, radius=4.0)
    assert mesh.shape[0] == 200000
    M, Minv = scenes.random_affines(16, seed=0x16, trans_range=20.0)
    cam = scenes.camera(eye=(0, 0, -60), forward=(0, 0, 1), right=(1, 0, 0), up=(0, 1, 0), fovy_deg=45,
                        aspect=W / H)
    m = mesh.copy()
    bvh = B.BvhBuilder(B.BUILD_SAH, 2, ctx=ctx).build_for(m)
    nodes0, used, tris0 = bvh.download()
    objs = [B.SceneObject(0, M[i].reshape(4, 4), Minv[i]) for i in range(16)]
    tlas = B.TlasBuilder(ctx=ctx).build_for(objs, [bvh])
    batch = B.Batch(ctx, [tlas])
    inst = np.zeros(16, orc.INSTANCE_DTYPE)
    inst["minv"] = Minv
    d_orig = ctx.to_device(m)
    d_frame = ctx.to_device(np.array([0], np.uint32))
    d_rays, d_hits = ctx.alloc(n * 32), ctx.alloc(n * 24)
    d_brays, d_bhits, d_cnt = ctx.alloc(n * 32), ctx.alloc(n * 24), ctx.alloc(4)
    lib = L.lib()
    cam_rec = np.ascontiguousarray(cam)
    seed = 0xB0B

    def frame():
        B.deform_wave(ctx, bvh, d_orig, 0.2, 4.0, 0.1, d_frame)
        batch.refit()
        batch.rebuild()
        h, t = ctx.handle, tlas.handle
        L.check(lib.bvht_raygen(h, L.ptr(cam_rec), W, H, B.ORDER_TILES, d_rays.ptr))
        L.check(lib.bvht_trace_tlas(h, t, d_rays.ptr, n, d_hits.ptr, None))
        L.check(lib.bvht_bounce_tlas(h, t, d_rays.ptr, d_hits.ptr, n, seed, d_brays.ptr, d_cnt.ptr))
        L.check(lib.bvht_trace_tlas_count(h, t, d_brays.ptr, d_cnt.ptr, n, d_bhits.ptr, None))

    def check_frame(f):
        g_nodes, g_used, g_tris = bvh.download()
        # the deformation is input synthesis: close to the float32 formula
        assert deformed_as_formula(g_tris, tris0, f), f"frame {f} deformation"
        nodes = nodes0.copy()
        orc.bvh_refit(nodes, used, g_tris)
        assert_nodes_equal(g_nodes, nodes, used)
        bounds = np.stack([np.concatenate(orc.instance_bounds(M[i], nodes[0]["bmin"], nodes[0]["bmax"]))
                           for i in range(16)])
        tnodes, tused = orc.tlas_build(bounds)
        t_nodes, t_used = tlas.download()
        assert t_used == tused and t_nodes.tobytes() == tnodes.tobytes(), f"frame {f} TLAS"
        rays = d_rays.download(B.RAY_DTYPE, n)
        ref = orc.tlas_trace(tnodes, inst, [(nodes, g_tris)], rays.view(orc.RAY_DTYPE), nthreads=THREADS)
        hits = d_hits.download(B.HIT_DTYPE, n)
        assert_hits_equal(hits, ref, f"frame {f} primary")
        nb = int(d_cnt.download(np.uint32, 1)[0])
        assert nb == (hits["prim"] != B.MISS).sum() > 50000
        brays = d_brays.download(B.RAY_DTYPE, nb)
        ref_b = orc.tlas_trace(tnodes, inst, [(nodes, g_tris)], brays.view(orc.RAY_DTYPE), nthreads=THREADS)
        assert_hits_equal(d_bhits.download(B.HIT_DTYPE, nb), ref_b, f"frame {f} bounce")
        assert len(np.unique(hits["inst"][hits["prim"] != B.MISS])) >= 12

    for f in (0, 1):
        d_frame.upload(np.array([f], np.uint32))
        frame()
        check_frame(f)
    # frames 2..63 as one captured graph per frame, the frame counter
    # advanced inside the graph
    d_frame.upload(np.array([2], np.uint32))

    def graph_frame():
        frame()
        B.counter_add(ctx, d_frame, 1)

    g = B.Graph.capture(ctx, graph_frame)
    for _ in range(62):
        g.launch()
    ctx.sync()
    g.close()
    assert int(d_frame.download(np.uint32, 1)[0]) == 64
    check_frame(63)
    batch.close()


def test_batched_64_scenes_every_scene(ctx, orc):
    work = scenes.batch_workload(64)
    models, tlases, d_origs = [], [], []
    for w in work:
        m = w["mesh"].copy()
        bvh = B.BvhBuilder(B.BUILD_SAH, 2, ctx=ctx).build_for(m)
        objs = [B.SceneObject(0, w["M"][k].reshape(4, 4), w["Minv"][k]) for k in range(w["M"].shape[0])]
        tlases.append(B.TlasBuilder(ctx=ctx).build_for(objs, [bvh]))
        models.append(bvh)
        d_origs.append(ctx.to_device(m))
    rays = np.concatenate([w["rays"] for w in work])
    offsets = np.concatenate([[0], np.cumsum([w["rays"].shape[0] for w in work])]).astype(np.uint32)
    n = int(offsets[-1])
    d_rays = ctx.to_device(rays)
    d_hits = ctx.alloc(n * 24)
    batch = B.Batch(ctx, tlases)
    batch.set_rays(offsets, 128)
    d_frame = ctx.to_device(np.array([5], np.uint32))
    for bvh, d_o in zip(models, d_origs):
        B.deform_wave(ctx, bvh, d_o, 0.2, 4.0, 0.1, d_frame)
    batch.refit()
    batch.rebuild()
    batch.trace(d_rays, d_hits)
    hits = d_hits.download(B.HIT_DTYPE, n)
    checked = 0
    for s, (w, bvh, t) in enumerate(zip(work, models, tlases)):
        g_nodes, used, g_tris = bvh.download()
        assert deformed_as_formula(g_tris, w["mesh"][bvh.permutation()], 5), f"scene {s} deformation"
        nodes = g_nodes.copy()
        orc.bvh_refit(nodes, used, g_tris)  # the batched refit left exactly Bvh::refit's boxes
        assert_nodes_equal(g_nodes, nodes, used)
        bounds = np.stack([np.concatenate(orc.instance_bounds(w["M"][k], g_nodes[0]["bmin"], g_nodes[0]["bmax"]))
                           for k in range(w["M"].shape[0])])
        tnodes, tused = orc.tlas_build(bounds)
        t_nodes, t_used = t.download()
        assert t_nodes.tobytes() == tnodes.tobytes(), f"scene {s} TLAS"
        inst = np.zeros(w["M"].shape[0], orc.INSTANCE_DTYPE)
        inst["minv"] = w["Minv"]
        a, b = int(offsets[s]), int(offsets[s + 1])
        ref = orc.tlas_trace(tnodes, inst, [(g_nodes, g_tris)], rays[a:b].view(orc.RAY_DTYPE), nthreads=THREADS)
        assert_hits_equal(hits[a:b], ref, f"scene {s} ({w['mesh'].shape[0]} tris, {b - a} rays)")
        checked += b - a
    assert checked == n > 10_000_000
    batch.close()
