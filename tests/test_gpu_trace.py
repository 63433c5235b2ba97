"""GPU traversal parity: libbvht_hip.so against the oracle on identical BVHs
and rays, bit for bit (t, u, v compared as raw binary32 words)."""
import numpy as np
import pytest

import bvhtracer_amd as B
from bvhtracer_amd import scenes
from helpers import assert_hits_equal, assert_nodes_equal, comb_tree, oracle_rays_to_product, upload

pytestmark = pytest.mark.gpu
f32 = np.float32


@pytest.mark.parametrize("faithful,max_leaf", [(True, 0), (False, 2)])
def test_soup_primary_parity(ctx, orc, faithful, max_leaf):
    """12,582-triangle soup, 640x640 camera rays (the CPU-runnable workload)."""
    mesh = scenes.triangle_soup()
    nodes, used, perm = orc.bvh_build(mesh, faithful, max_leaf)
    bvh = upload(ctx, nodes, used, perm)
    rays = orc.raygen(scenes.soup_camera(), 640, 640)
    ref, st_ref = orc.bvh_trace(nodes, perm, rays, nthreads=8, stats=True)
    got, st = bvh.intersect(oracle_rays_to_product(rays), stats=True)
    assert_hits_equal(got, ref, "soup primary")
    assert np.array_equal(st.view(np.uint32), st_ref.view(np.uint32)), "work counters differ"
    assert (got["prim"] != B.MISS).sum() > 10000


@pytest.mark.parametrize("order", [B.ORDER_ROWS, B.ORDER_TILES])
def test_raygen_parity(ctx, orc, order):
    cam = scenes.soup_camera()
    ref = orc.raygen(cam, 640, 640)
    d = B.raygen(ctx, cam, 640, 640, order)
    got = d.download(B.RAY_DTYPE, 640 * 640)
    if order == B.ORDER_TILES:
        j = np.arange(640 * 640)
        tile, w = j // 64, j % 64
        x = (tile % 80) * 8 + w % 8
        y = (tile // 80) * 8 + w // 8
        ref = ref[y * 640 + x]
    assert got.tobytes() == oracle_rays_to_product(ref).tobytes()


def random_rays(n, seed, lo=-6.0, hi=6.0):
    rng = scenes.Pcg32(seed, np.arange(n, dtype=np.uint64))
    o = np.stack([rng.next_f32() for _ in range(3)], 1) * f32(hi - lo) + f32(lo)
    d = np.stack([rng.next_f32() for _ in range(3)], 1) * f32(2) - f32(1)
    r = np.zeros(n, B.RAY_DTYPE)
    r["o"], r["d"] = o, d
    r["tmax"] = np.where(np.arange(n) % 4 == 0, f32(3.0), np.finfo(np.float32).max)
    return r


def test_incoherent_rays_parity(ctx, orc):
    mesh = scenes.triangle_soup(8000, seed=17)
    nodes, used, perm = orc.bvh_build(mesh, False, 2)
    bvh = upload(ctx, nodes, used, perm)
    rays = random_rays(200000, 5)
    ref = orc.bvh_trace(nodes, perm, rays.view(orc.RAY_DTYPE), nthreads=8)
    assert_hits_equal(bvh.intersect(rays), ref, "incoherent")


def test_axis_aligned_and_degenerate_rays(ctx, orc):
    """Zero direction components (1/d = +-inf, NaN slabs), -0 components,
    tiny tmax, rays starting inside boxes, and triangles under the 1e-4
    determinant threshold."""
    mesh = scenes.triangle_soup(500, seed=1)
    mesh[:20, 6:9] = mesh[:20, 3:6]  # degenerate (zero area) triangles
    nodes, used, perm = orc.bvh_build(mesh, False, 2)
    bvh = upload(ctx, nodes, used, perm)
    rays = []
    for o in [(0, 0, -20), (0.5, -0.25, -20), (-3, 2, 0), (0, 0, 0), (1e-3, 1e-3, -1e-3)]:
        for d in [(0, 0, 1), (0, 0, -1), (1, 0, 0), (0, -1, 0), (-0.0, 0.0, 1), (0, 1, 1), (1, 1, 1), (1e-20, 0, 1)]:
            for t in [np.finfo(np.float32).max, 1e-5, 7.5, np.inf]:
                r = np.zeros(1, B.RAY_DTYPE)
                r["o"][0], r["d"][0], r["tmax"][0] = o, d, t
                rays.append(r)
    rays = np.concatenate(rays)
    ref = orc.bvh_trace(nodes, perm, rays.view(orc.RAY_DTYPE))
    assert_hits_equal(bvh.intersect(rays), ref, "edge rays")


def test_deep_tree_uses_spill(ctx, orc):
    nodes, used, tris = comb_tree(150)
    depth = orc.bvh_depth(nodes)
    assert depth == 150  # deeper than the largest LDS stack (64 entries)
    bvh = upload(ctx, nodes, used, tris)
    assert bvh.depth() == depth
    n = 4096
    k = np.arange(n)
    rays = np.zeros(n, B.RAY_DTYPE)
    rays["o"] = np.stack([np.full(n, -5.0), (k % 64) / 64.0 - 0.5, (k // 64) / 64.0 - 0.5], 1)
    rays["d"] = np.stack([np.ones(n), np.full(n, 1e-3), np.full(n, -1e-3)], 1)
    rays["tmax"] = np.where(k % 2 == 0, np.finfo(np.float32).max, f32(40.5))
    ref, st = orc.bvh_trace(nodes, tris, rays.view(orc.RAY_DTYPE), stats=True)
    assert st["pairs"].max() >= 149
    assert_hits_equal(bvh.intersect(rays), ref, "deep tree")


def test_empty_and_single(ctx, orc):
    tri = np.array([[0, 0.5, 0, -0.5773503, -0.5, 0, 0.5773503, -0.5, 0]], np.float32)
    nodes, used, perm = orc.bvh_build(tri)
    bvh = upload(ctx, nodes, used, perm)
    assert bvh.intersect(np.zeros(0, B.RAY_DTYPE)).shape == (0,)
    r = B.Ray.from_origin_dir((0, 0, 5), (0, 0, -1))
    h = bvh.intersect(r)[0]
    assert h["prim"] == 0 and h["t"] == f32(5) and h["packed"] == 0
    with pytest.raises(B.BvhtError):
        B.Bvh.from_arrays(np.zeros(0, B.NODE_DTYPE), 0, np.zeros((0, 9), np.float32), ctx=ctx)
    bad = nodes.copy()
    bad[0]["count"] = 5  # leaf range beyond the triangle array: rejected on the host
    with pytest.raises(B.BvhtError):
        B.Bvh.from_arrays(bad, used, perm, ctx=ctx)


def test_bounce_rays_parity(ctx, orc):
    """Primary pass on the GPU, bounce rays generated on the GPU, then both
    passes traced by the GPU and by the oracle on the same buffers."""
    mesh = scenes.icosphere(24, radius=32.0)
    nodes, used, perm = orc.bvh_build(mesh, False, 2)
    bvh = upload(ctx, nodes, used, perm)
    W, H = 320, 184
    cam = scenes.sphere_camera(32.0, W, H)
    d_rays = B.raygen(ctx, cam, W, H, B.ORDER_TILES)
    d_hits = ctx.alloc(W * H * B.HIT_DTYPE.itemsize)
    bvh.intersect_device(d_rays, W * H, d_hits)
    rays = d_rays.download(B.RAY_DTYPE, W * H)
    hits = d_hits.download(B.HIT_DTYPE, W * H)
    assert_hits_equal(hits, orc.bvh_trace(nodes, perm, rays.view(orc.RAY_DTYPE), nthreads=8), "primary")
    d_b, d_cnt = B.bounce(ctx, bvh, d_rays, d_hits, W * H, seed=1234)
    cnt = int(d_cnt.download(np.uint32, 1)[0])
    assert cnt == (hits["prim"] != B.MISS).sum() > 1000
    brays = d_b.download(B.RAY_DTYPE, cnt)
    # bounce origins sit 1e-4 off the surface on the incoming side
    assert np.isfinite(brays["o"]).all() and np.allclose(np.linalg.norm(brays["d"], axis=1), 1, atol=1e-5)
    got = bvh.intersect(brays)
    assert_hits_equal(got, orc.bvh_trace(nodes, perm, brays.view(orc.RAY_DTYPE), nthreads=8), "bounce")


def test_every_trace_config_is_bit_identical(ctx, orc):
    """Each workgroup size x LDS depth variant (the autotuner's search space)
    gives the oracle's hits, on a normal tree and on the 150-level comb tree
    that spills past every LDS depth."""
    from bvhtracer_amd import autotune

    mesh = scenes.triangle_soup(3000, seed=23)
    nodes, used, perm = orc.bvh_build(mesh, False, 2)
    bvh = upload(ctx, nodes, used, perm)
    rays = random_rays(30000, 8)
    ref = orc.bvh_trace(nodes, perm, rays.view(orc.RAY_DTYPE), nthreads=8)
    cnodes, cused, ctris = comb_tree(150)
    comb = upload(ctx, cnodes, cused, ctris)
    k = np.arange(2048)
    crays = np.zeros(2048, B.RAY_DTYPE)
    crays["o"] = np.stack([np.full(2048, -5.0), (k % 64) / 64.0 - 0.5, (k // 64) / 64.0 - 0.5], 1)
    crays["d"] = np.stack([np.ones(2048), np.full(2048, 1e-3), np.full(2048, -1e-3)], 1)
    crays["tmax"] = np.finfo(np.float32).max
    cref = orc.bvh_trace(cnodes, ctris, crays.view(orc.RAY_DTYPE))
    try:
        for wg, d in autotune.configurations():
            autotune.set_config(ctx, wg, d)
            assert autotune.current(ctx) == (wg, d)
            assert_hits_equal(bvh.intersect(rays), ref, f"config {wg}x{d}")
            assert_hits_equal(comb.intersect(crays), cref, f"comb config {wg}x{d}")
    finally:
        autotune.set_config(ctx, 0, 0)
    with pytest.raises(B.BvhtError):
        autotune.set_config(ctx, 96, 16)


def dir_bins(d):
    """24 direction bins: dominant axis (first of equal magnitudes) and its
    sign, then the signs of the next two components in axis order."""
    a = np.abs(d)
    axis = np.where((a[:, 0] >= a[:, 1]) & (a[:, 0] >= a[:, 2]), 0, np.where(a[:, 1] >= a[:, 2], 1, 2))
    r = np.arange(d.shape[0])
    c0, c1, c2 = d[r, axis], d[r, (axis + 1) % 3], d[r, (axis + 2) % 3]
    return (axis * 2 + (c0 < 0)) * 4 + 2 * (c1 >= 0) + (c2 >= 0)


@pytest.mark.parametrize("tlas", [False, True])
def test_sorted_bounce_is_a_grouping_of_the_ray_order_bounce(ctx, orc, tlas):
    """bvht_bounce_sorted emits exactly bvht_bounce's rays, grouped by
    direction bin with ray order kept inside each bin; d_src names each ray's
    source; the sorted rays trace to the oracle's hits."""
    mesh = scenes.icosphere(24, radius=32.0)
    nodes, used, perm = orc.bvh_build(mesh, False, 2)
    bvh = upload(ctx, nodes, used, perm)
    W, H = 320, 184
    cam = scenes.sphere_camera(32.0, W, H)
    d_rays = B.raygen(ctx, cam, W, H, B.ORDER_TILES)
    d_hits = ctx.alloc(W * H * B.HIT_DTYPE.itemsize)
    if tlas:
        M, Minv = scenes.random_affines(1, 3, scale_range=(1, 1), trans_range=0.5)
        t = B.TlasBuilder(ctx=ctx).build_for([B.SceneObject(0, M[0].reshape(4, 4), Minv[0])], [bvh])
        t.intersect_device(d_rays, W * H, d_hits)
        d_b, d_cnt = B.bounce_tlas(ctx, t, d_rays, d_hits, W * H, seed=99)
        d_s, d_src, d_cnt2 = B.bounce_tlas_sorted(ctx, t, d_rays, d_hits, W * H, seed=99)
    else:
        bvh.intersect_device(d_rays, W * H, d_hits)
        d_b, d_cnt = B.bounce(ctx, bvh, d_rays, d_hits, W * H, seed=99)
        d_s, d_src, d_cnt2 = B.bounce_sorted(ctx, bvh, d_rays, d_hits, W * H, seed=99)
    cnt = int(d_cnt.download(np.uint32, 1)[0])
    assert cnt == int(d_cnt2.download(np.uint32, 1)[0]) > 1000
    plain = d_b.download(B.RAY_DTYPE, cnt)
    srt = d_s.download(B.RAY_DTYPE, cnt)
    src = d_src.download(np.uint32, cnt)
    hits = d_hits.download(B.HIT_DTYPE, W * H)
    hit_idx = np.nonzero(hits["prim"] != B.MISS)[0]
    assert np.array_equal(np.sort(src), hit_idx)
    pos = np.searchsorted(hit_idx, src)          # where each sorted ray sits in ray order
    assert srt.tobytes() == plain[pos].tobytes()
    bins = dir_bins(srt["d"])
    assert (np.diff(bins) >= 0).all()
    for b in np.unique(bins):
        assert (np.diff(src[bins == b]) > 0).all()
    assert len(np.unique(bins)) >= 8
    if not tlas:
        assert_hits_equal(bvh.intersect(srt), orc.bvh_trace(nodes, perm, srt.view(orc.RAY_DTYPE), nthreads=8),
                          "sorted bounce")


@pytest.mark.parametrize("left_count", [63, 64, 65])
def test_stack_entry_packing_boundary(ctx, orc, left_count):
    """Two-leaf trees around the packed-stack limit (a pushed leaf of 63
    triangles packs into its stack entry, 64 and more do not): the same hits
    as the oracle, with finite and infinite ray lengths."""
    rng = np.random.default_rng(left_count)
    n = left_count + 40
    tris = np.zeros((n, 9), f32)
    # two overlapping clusters along z so that rays hit both leaf boxes
    for lo, hi, z in ((0, left_count, 0.0), (left_count, n, 0.5)):
        c = rng.uniform(-1, 1, size=(hi - lo, 1, 3)).astype(f32)
        c[..., 2] = z + rng.uniform(-0.1, 0.1, size=(hi - lo, 1)).astype(f32)
        tris[lo:hi] = (c + rng.uniform(-0.3, 0.3, size=(hi - lo, 3, 3)).astype(f32)).reshape(-1, 9)
    nodes = np.zeros(4, orc.NODE_DTYPE)
    nodes["left_first"][0], nodes["count"][0] = 2, 0
    for k, (lo, hi) in ((2, (0, left_count)), (3, (left_count, n))):
        v = tris[lo:hi].reshape(-1, 3)
        nodes["bmin"][k], nodes["bmax"][k] = v.min(0), v.max(0)
        nodes["left_first"][k], nodes["count"][k] = lo, hi - lo
    v = tris.reshape(-1, 3)
    nodes["bmin"][0], nodes["bmax"][0] = v.min(0), v.max(0)
    bvh = upload(ctx, nodes, 4, tris)
    m = 4096
    r = np.zeros(m, B.RAY_DTYPE)
    r["o"] = np.stack([rng.uniform(-1.2, 1.2, m), rng.uniform(-1.2, 1.2, m), np.full(m, -5.0)], 1).astype(f32)
    r["d"] = np.stack([rng.uniform(-0.05, 0.05, m), rng.uniform(-0.05, 0.05, m), np.ones(m)], 1).astype(f32)
    r["tmax"] = np.where(np.arange(m) % 3 == 0, f32(np.inf), np.finfo(np.float32).max)
    ref = orc.bvh_trace(nodes, tris, r.view(orc.RAY_DTYPE))
    got = bvh.intersect(r)
    assert_hits_equal(got, ref, f"two-leaf tree, left leaf {left_count}")
    assert (got["prim"] != B.MISS).sum() > m // 4


@pytest.mark.parametrize("faithful,width,height,order", [(False, 640, 640, B.ORDER_TILES), (True, 200, 150, B.ORDER_ROWS),
                                                          (False, 328, 120, B.ORDER_TILES)])
def test_camera_bounce_pass_equals_separate_calls(ctx, orc, faithful, width, height, order):
    """bvht_trace_camera_bounce = bvht_raygen + bvht_trace_blas +
    bvht_bounce_sorted: hits, bounce rays, source indices and count bit for
    bit, for packed and unpacked trees and a frame that is not a multiple of
    64 rays."""
    mesh = scenes.triangle_soup()
    nodes, used, perm = orc.bvh_build(mesh, faithful, 0 if faithful else 2)
    bvh = upload(ctx, nodes, used, perm)
    cam = scenes.soup_camera()
    n = width * height
    d_rays = B.raygen(ctx, cam, width, height, order)
    d_hits = ctx.alloc(n * B.HIT_DTYPE.itemsize)
    bvh.intersect_device(d_rays, n, d_hits)
    d_out, d_src, d_cnt = B.bounce_sorted(ctx, bvh, d_rays, d_hits, n, 0x5EED)
    f_hits, f_out, f_src, f_cnt = B.trace_camera_bounce(ctx, bvh, cam, width, height, order, 0x5EED)
    ctx.sync()
    nb = int(d_cnt.download(np.uint32, 1)[0])
    assert int(f_cnt.download(np.uint32, 1)[0]) == nb and nb > 1000
    assert f_hits.download(B.HIT_DTYPE, n).tobytes() == d_hits.download(B.HIT_DTYPE, n).tobytes()
    assert f_out.download(B.RAY_DTYPE, nb).tobytes() == d_out.download(B.RAY_DTYPE, nb).tobytes()
    assert f_src.download(np.uint32, nb).tobytes() == d_src.download(np.uint32, nb).tobytes()
    # and the hits are the oracle's
    ref = orc.bvh_trace(nodes, perm, orc.raygen(cam, width, height) if order == B.ORDER_ROWS else
                        d_rays.download(B.RAY_DTYPE, n).view(orc.RAY_DTYPE), nthreads=8)
    assert_hits_equal(f_hits.download(B.HIT_DTYPE, n), ref, "camera pass")


def test_camera_bounce_pass_full_frame_configs_and_no_src(ctx):
    """1920x1080 on an 81,920-triangle sphere: the wave counts span 190 scan
    tiles.  With 64- and 256-lane workgroups, an 8-entry LDS stack (the tree
    is deeper, so the stack spills) and without source indices, the fused
    pass equals bvht_raygen + bvht_trace_blas + bvht_bounce_sorted bit for
    bit."""
    from bvhtracer_amd import _lib as L
    from bvhtracer_amd import autotune

    W, H, seed = 1920, 1080, 0xB0B
    n = W * H
    bvh = B.BvhBuilder(B.BUILD_SAH, 2, ctx=ctx).build_for(scenes.icosphere(64, radius=32.0))
    assert bvh.depth() - 1 > 8
    cam = scenes.sphere_camera(32.0, W, H)
    d_rays = B.raygen(ctx, cam, W, H, B.ORDER_TILES)
    d_hits = ctx.alloc(n * B.HIT_DTYPE.itemsize)
    bvh.intersect_device(d_rays, n, d_hits)
    d_out, d_src, d_cnt = B.bounce_sorted(ctx, bvh, d_rays, d_hits, n, seed)
    ctx.sync()
    nb = int(d_cnt.download(np.uint32, 1)[0])
    assert nb > n // 2
    hits, out, src = (d_hits.download(B.HIT_DTYPE, n).tobytes(), d_out.download(B.RAY_DTYPE, nb).tobytes(),
                      d_src.download(np.uint32, nb).tobytes())
    try:
        for wg, depth in ((64, 16), (256, 8)):
            autotune.set_config(ctx, wg, depth)
            f_hits, f_out, f_src, f_cnt = B.trace_camera_bounce(ctx, bvh, cam, W, H, B.ORDER_TILES, seed)
            ctx.sync()
            assert int(f_cnt.download(np.uint32, 1)[0]) == nb, (wg, depth)
            assert f_hits.download(B.HIT_DTYPE, n).tobytes() == hits, (wg, depth)
            assert f_out.download(B.RAY_DTYPE, nb).tobytes() == out, (wg, depth)
            assert f_src.download(np.uint32, nb).tobytes() == src, (wg, depth)
        cam_rec = np.ascontiguousarray(cam, B.CAMERA_DTYPE).reshape(1)
        g_hits, g_out, g_cnt = ctx.alloc(n * B.HIT_DTYPE.itemsize), ctx.alloc(n * B.RAY_DTYPE.itemsize), ctx.alloc(4)
        L.check(L.lib().bvht_trace_camera_bounce(ctx.handle, bvh.handle, L.ptr(cam_rec), W, H, B.ORDER_TILES, seed,
                                                 g_hits.ptr, g_out.ptr, None, g_cnt.ptr))
        ctx.sync()
        assert int(g_cnt.download(np.uint32, 1)[0]) == nb
        assert g_out.download(B.RAY_DTYPE, nb).tobytes() == out
    finally:
        autotune.set_config(ctx, 0, 0)


def test_camera_bounce_pass_empty_frame(ctx):
    bvh = B.BvhBuilder(B.BUILD_SAH, 2, ctx=ctx).build_for(scenes.triangle_soup(500, seed=3))
    d_cnt = ctx.alloc(4)
    d_cnt.upload(np.array([77], np.uint32))
    B.trace_camera_bounce(ctx, bvh, scenes.soup_camera(), 0, 0, B.ORDER_TILES, 1, d_count=d_cnt)
    ctx.sync()
    assert int(d_cnt.download(np.uint32, 1)[0]) == 0


def test_bounce_orders_agree_beyond_one_scan_tile(ctx):
    """1280x960 rays give 4800 blocks of 256 rays: the hit compaction of
    bvht_bounce scans more than one 4096-entry tile.  Its ray-order output
    equals bvht_bounce_sorted's put back in source order, and both count
    exactly the rays that hit."""
    W, H, seed = 1280, 960, 0xF00D
    n = W * H
    mesh = scenes.icosphere(64, radius=32.0)
    bvh = B.BvhBuilder(B.BUILD_SAH, 2, ctx=ctx).build_for(mesh.copy())
    d_rays = B.raygen(ctx, scenes.sphere_camera(32.0, W, H), W, H, B.ORDER_TILES)
    d_hits = ctx.alloc(n * B.HIT_DTYPE.itemsize)
    bvh.intersect_device(d_rays, n, d_hits)
    d_plain, d_pcnt = B.bounce(ctx, bvh, d_rays, d_hits, n, seed)
    d_sorted, d_src, d_scnt = B.bounce_sorted(ctx, bvh, d_rays, d_hits, n, seed)
    ctx.sync()
    hits = d_hits.download(B.HIT_DTYPE, n)
    expect = int((hits["prim"] < mesh.shape[0]).sum())
    assert expect > n // 4
    assert int(d_pcnt.download(np.uint32, 1)[0]) == expect
    assert int(d_scnt.download(np.uint32, 1)[0]) == expect
    src = d_src.download(np.uint32, expect)
    assert np.array_equal(np.sort(src), np.flatnonzero(hits["prim"] < mesh.shape[0]).astype(np.uint32))
    back = d_sorted.download(B.RAY_DTYPE, expect)[np.argsort(src)]
    assert back.tobytes() == d_plain.download(B.RAY_DTYPE, expect).tobytes()


def test_two_contexts_trace_one_tree_at_once(ctx):
    """Frames on two contexts (two streams) tracing the same BVH at the same
    time, with an 8-entry LDS stack so both spill: each context has its own
    spill buffer, and every frame equals the frame traced alone."""
    from bvhtracer_amd import _lib as L
    from bvhtracer_amd import autotune

    lib = L.lib()
    W, H, seed = 640, 480, 0x2C7
    n = W * H
    bvh = B.BvhBuilder(B.BUILD_SAH, 2, ctx=ctx).build_for(scenes.icosphere(64, radius=32.0))
    assert bvh.depth() - 1 > 8
    cam = np.ascontiguousarray(scenes.sphere_camera(32.0, W, H), B.CAMERA_DTYPE).reshape(1)
    other = B.Context(0)
    ctxs = [ctx, other]
    try:
        bufs = [dict(h=c.alloc(n * 24), b=c.alloc(n * 32), bh=c.alloc(n * 24), src=c.alloc(n * 4), cnt=c.alloc(4))
                for c in ctxs]

        def frame(k):
            c, d = ctxs[k], bufs[k]
            L.check(lib.bvht_trace_camera_bounce(c.handle, bvh.handle, L.ptr(cam), W, H, B.ORDER_TILES, seed,
                                                 d["h"].ptr, d["b"].ptr, d["src"].ptr, d["cnt"].ptr))
            L.check(lib.bvht_trace_blas_count(c.handle, bvh.handle, d["b"].ptr, d["cnt"].ptr, n, d["bh"].ptr, None))

        for c in ctxs:
            autotune.set_config(c, 64, 8)
        frame(0)
        ctx.sync()
        nb = int(bufs[0]["cnt"].download(np.uint32, 1)[0])
        ref = (bufs[0]["h"].download(B.HIT_DTYPE, n).tobytes(), bufs[0]["bh"].download(B.HIT_DTYPE, nb).tobytes())
        for _ in range(6):
            frame(0)
            frame(1)
        for c in ctxs:
            c.sync()
        for d in bufs:
            assert int(d["cnt"].download(np.uint32, 1)[0]) == nb
            assert d["h"].download(B.HIT_DTYPE, n).tobytes() == ref[0]
            assert d["bh"].download(B.HIT_DTYPE, nb).tobytes() == ref[1]
    finally:
        autotune.set_config(ctx, 0, 0)
        other.close()


def test_camera_pass_wave_schedule_keeps_results(ctx):
    """From the second pass over the same tree, camera and frame the camera
    pass runs its 64-ray groups longest first, ordered by the durations the
    previous pass recorded.  Every pass (other seeds, another camera and
    back) equals the same pass on a fresh context, which runs unscheduled,
    bit for bit."""
    mesh = scenes.icosphere(48, radius=32.0)
    bvh = B.BvhBuilder(B.BUILD_SAH, 2, ctx=ctx).build_for(mesh.copy())
    W, H = 640, 360
    cam = scenes.sphere_camera(32.0, W, H)
    cam2 = scenes.sphere_camera(32.0, W, H, distance=2.2)

    def run(c, camera, seed):
        out = B.trace_camera_bounce(c, bvh, camera, W, H, B.ORDER_TILES, seed)
        nb = int(out[3].download(np.uint32, 1)[0])
        return (nb, out[0].download(np.uint8).tobytes(), out[1].download(B.RAY_DTYPE, nb).tobytes(),
                out[2].download(np.uint32, nb).tobytes())

    for camera, seed in [(cam, 1), (cam, 1), (cam, 2), (cam, 3), (cam2, 3), (cam2, 4), (cam, 5)]:
        fresh = B.Context(0)
        try:
            want = run(fresh, camera, seed)
        finally:
            fresh.close()
        assert run(ctx, camera, seed) == want, (seed,)



def test_blas_wave_schedule_keeps_hits(orc):
    """bvht_set_wave_schedule on bvht_trace_blas: repeated traces of the same
    camera rays (ordered longest first from the second call), traces of other
    buffers and counts in between, and a captured trace replayed after eager
    calls of another count all give the unscheduled hits bit for bit, which
    equal the oracle's."""
    ctx = B.Context(0)
    try:
        mesh = scenes.icosphere(48, radius=32.0)
        m = mesh.copy()
        bvh = B.BvhBuilder(B.BUILD_SAH, 2, ctx=ctx).build_for(m)
        nodes, used, tris = bvh.download()
        W, H = 640, 360
        n = W * H
        cam = scenes.sphere_camera(32.0, W, H)
        d_rays = B.raygen(ctx, cam, W, H, B.ORDER_TILES)
        rays = d_rays.download(B.RAY_DTYPE, n)
        want = bvh.intersect(rays)
        assert_hits_equal(want, orc.bvh_trace(nodes, tris, rays.view(orc.RAY_DTYPE), nthreads=8), "unscheduled")
        d_hits = ctx.alloc(n * B.HIT_DTYPE.itemsize)
        d_part = ctx.alloc(n * B.HIT_DTYPE.itemsize)
        ctx.set_wave_schedule(True)
        for call in range(10):
            d_hits.upload(np.zeros(n, B.HIT_DTYPE))
            bvh.intersect_device(d_rays, n, d_hits)
            assert d_hits.download(B.HIT_DTYPE, n).tobytes() == want.tobytes(), call
            if call in (3, 6):
                part = n // 5 + 3
                bvh.intersect_device(d_rays, part, d_part)
                assert d_part.download(B.HIT_DTYPE, part).tobytes() == want[:part].tobytes()

        def enqueue():
            bvh.intersect_device(d_rays, n, d_hits)

        g = B.Graph.capture(ctx, enqueue, keep=(d_rays, d_hits))
        for _ in range(3):
            for _ in range(2):
                bvh.intersect_device(d_rays, n // 2 + 1, d_part)
            d_hits.upload(np.zeros(n, B.HIT_DTYPE))
            g.launch()
            ctx.sync()
            assert d_hits.download(B.HIT_DTYPE, n).tobytes() == want.tobytes()
        g.close()
    finally:
        ctx.close()


def test_trace_after_refit_sees_the_new_boxes(ctx, orc):
    """Trace, deform the vertices and refit on the device, trace again: the
    second trace (whose uniform node pairs go through the scalar data cache)
    must use the refitted boxes, giving the oracle's hits on the oracle's
    refit of the same tree; then the same again in the other direction."""
    mesh = scenes.icosphere(48, radius=8.0)
    nodes, used, perm = orc.bvh_build(mesh, False, 2)
    bvh = upload(ctx, nodes, used, perm)
    W, H = 256, 192
    cam = scenes.sphere_camera(8.0, W, H)
    rays = B.raygen(ctx, cam, W, H, B.ORDER_TILES).download(B.RAY_DTYPE, W * H)
    first = bvh.intersect(rays)
    assert_hits_equal(first, orc.bvh_trace(nodes, perm, rays.view(orc.RAY_DTYPE), nthreads=8), "before refit")
    rng = np.random.default_rng(5)
    shifted = perm.copy()
    for step in range(2):
        # push every triangle along its own random offset (boxes move by up to 0.4)
        shifted = (shifted + rng.uniform(-0.2, 0.2, size=(shifted.shape[0], 1)).astype(f32)).astype(f32)
        bvh.refit(shifted)
        ref_nodes = nodes.copy()
        orc.bvh_refit(ref_nodes, used, shifted)
        g_nodes, _, _ = bvh.download()
        assert_nodes_equal(g_nodes, ref_nodes, used)
        got = bvh.intersect(rays)
        ref = orc.bvh_trace(ref_nodes, shifted, rays.view(orc.RAY_DTYPE), nthreads=8)
        assert_hits_equal(got, ref, f"after refit {step}")
        assert got.tobytes() != first.tobytes()
