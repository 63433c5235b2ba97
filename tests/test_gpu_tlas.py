"""GPU TLAS parity: object bounds, the agglomerative rebuild and the
two-level trace against the oracle (scene/tlas.rs, scene/scene_object.rs)."""
import numpy as np
import pytest

import bvhtracer_amd as B
from bvhtracer_amd import scenes
from helpers import assert_hits_equal

pytestmark = pytest.mark.gpu


def oracle_instances(orc, objs):
    inst = np.zeros(len(objs), orc.INSTANCE_DTYPE)
    for k, o in enumerate(objs):
        inst["minv"][k] = o.minv
        inst["blas"][k] = o.model
    return inst


def build_scene(ctx, orc, n_obj, seed, freq=12, radius=3.0, n_models=1):
    models, blases = [], []
    for m in range(n_models):
        mesh = scenes.icosphere(freq + 3 * m, radius=radius, seed=100 + m)
        nodes, used, perm = orc.bvh_build(mesh, False, 2)
        blases.append((nodes, perm))
        models.append(B.Bvh.from_arrays(nodes, used, perm, ctx=ctx))
    M, Minv = scenes.random_affines(n_obj, seed, trans_range=10.0)
    objs = [B.SceneObject(k % n_models, M[k].reshape(4, 4), Minv[k]) for k in range(n_obj)]
    return models, blases, objs


def oracle_tlas(orc, blases, objs):
    bounds = np.stack([np.concatenate(orc.instance_bounds(o.m, blases[o.model][0][0]["bmin"],
                                                          blases[o.model][0][0]["bmax"])) for o in objs])
    return orc.tlas_build(bounds)


@pytest.mark.parametrize("n_obj,n_models", [(1, 1), (2, 1), (16, 1), (40, 3)])
def test_tlas_build_and_trace(ctx, orc, n_obj, n_models):
    models, blases, objs = build_scene(ctx, orc, n_obj, seed=n_obj, n_models=n_models)
    tlas = B.TlasBuilder(ctx=ctx).build_for(objs, models)
    tnodes, used = oracle_tlas(orc, blases, objs)
    g_nodes, g_used = tlas.download()
    assert g_used == used
    assert g_nodes.tobytes() == tnodes.tobytes()
    cam = scenes.camera(eye=(0, 0, -45), forward=(0, 0, 1), right=(1, 0, 0), up=(0, 1, 0), fovy_deg=45)
    rays = orc.raygen(cam, 200, 200)
    inst = oracle_instances(orc, objs)
    ref = orc.tlas_trace(tnodes, inst, blases, rays, nthreads=8)
    prays = np.zeros(rays.shape[0], B.RAY_DTYPE)
    prays.view(np.uint8)[:] = rays.view(np.uint8)
    got, st = tlas.intersect(prays, stats=True)
    assert_hits_equal(got, ref, "tlas primary")
    _, st_ref = orc.tlas_trace(tnodes, inst, blases, rays, nthreads=8, stats=True)
    assert np.array_equal(st.view(np.uint32), st_ref.view(np.uint32))
    assert (got["prim"] != B.MISS).sum() > 100


def translated(offsets):
    """Column-major (M, Minv) pairs of pure translations."""
    n = len(offsets)
    M = np.tile(np.eye(4, dtype=np.float32).reshape(16), (n, 1))
    Minv = M.copy()
    for k, (x, y, z) in enumerate(offsets):
        M[k, 12:15] = (x, y, z)
        Minv[k, 12:15] = (-x, -y, -z)
    return M, Minv


@pytest.mark.parametrize("case", ["random64", "random65", "chain48", "duplicates20", "duplicates70"])
def test_tlas_rebuild_at_the_wave_limit_chains_and_ties(ctx, orc, case):
    """Up to 64 objects take the one-wave merge loop, more the block-wide one;
    a chain with growing gaps makes a deep tree (the traversal's stack
    follows its depth), and identical boxes tie on every area (first-index
    rule) and chain all objects (height n).  Nodes bit for bit and hits
    equal the oracle's."""
    if case.startswith("random"):
        n = int(case[6:])
        M, Minv = scenes.random_affines(n, 0x5EED + n, trans_range=10.0)
    elif case == "chain48":
        gaps = 7.0 * 1.25 ** np.arange(48)
        M, Minv = translated([(x, 0.0, 0.0) for x in np.cumsum(gaps) - gaps.sum() / 2])
    else:
        M, Minv = translated([(0.0, 0.0, 0.0)] * int(case[10:]))
    mesh = scenes.icosphere(8, radius=3.0, seed=7)
    nodes, used, perm = orc.bvh_build(mesh, False, 2)
    blases = [(nodes, perm)]
    model = B.Bvh.from_arrays(nodes, used, perm, ctx=ctx)
    objs = [B.SceneObject(0, M[k].reshape(4, 4), Minv[k]) for k in range(M.shape[0])]
    tlas = B.TlasBuilder(ctx=ctx).build_for(objs, [model])
    tnodes, t_used = oracle_tlas(orc, blases, objs)
    g_nodes, g_used = tlas.download()
    assert g_used == t_used == 2 * len(objs)
    assert g_nodes.tobytes() == tnodes.tobytes()
    # rays along +x through the objects' rows, and back along -x
    y, z = np.meshgrid(np.linspace(-9.0, 9.0, 48), np.linspace(-9.0, 9.0, 48))
    rays = np.zeros(2 * y.size, B.RAY_DTYPE)
    rays["o"][: y.size] = np.stack([np.full(y.size, -2.0e6), y.ravel(), z.ravel()], 1)
    rays["d"][: y.size] = (1.0, 0.0, 0.0)
    rays["o"][y.size:] = np.stack([np.full(y.size, 2.0e6), y.ravel(), z.ravel()], 1)
    rays["d"][y.size:] = (-1.0, 0.0, 0.0)
    rays["tmax"] = np.float32(3.4028235e38)
    ref = orc.tlas_trace(tnodes, oracle_instances(orc, objs), blases, rays.view(orc.RAY_DTYPE), nthreads=8)
    got = tlas.intersect(rays)
    assert_hits_equal(got, ref, f"tlas {case}")
    assert (got["prim"] != B.MISS).sum() > 100


def test_tlas_frame_refit_refresh_rebuild_bounce(ctx, orc):
    """Per frame: deform the shared model, refit it, refresh the object
    bounds (set_transform with the same transform), rebuild the TLAS, trace
    primary rays, generate bounce rays, trace them -- every step against the
    oracle on the same buffers."""
    mesh = scenes.icosphere(20, radius=3.0)
    m = mesh.copy()
    model = B.BvhBuilder(B.BUILD_SAH, 2, ctx=ctx).build_for(m)
    nodes, used, perm = model.download()
    M, Minv = scenes.random_affines(16, seed=77, trans_range=10.0)
    objs = [B.SceneObject(0, M[k].reshape(4, 4), Minv[k]) for k in range(16)]
    tlas = B.TlasBuilder(ctx=ctx).build_for(objs, [model])
    inst = oracle_instances(orc, objs)
    cam = scenes.camera(eye=(0, 0, -45), forward=(0, 0, 1), right=(1, 0, 0), up=(0, 1, 0), fovy_deg=45,
                        aspect=16 / 9)
    W, H = 256, 144
    for frame in (1, 2):
        deformed = scenes.deform_wave(perm, frame)
        model.refit(deformed)
        tlas.refresh_bounds()
        tlas.rebuild()
        orc.bvh_refit(nodes, used, deformed)
        tnodes, tused = oracle_tlas(orc, [(nodes, deformed)], objs)
        g_nodes, g_used = tlas.download()
        assert g_nodes.tobytes() == tnodes.tobytes() and g_used == tused
        d_rays = B.raygen(ctx, cam, W, H, B.ORDER_TILES)
        d_hits = ctx.alloc(W * H * B.HIT_DTYPE.itemsize)
        tlas.intersect_device(d_rays, W * H, d_hits)
        rays = d_rays.download(B.RAY_DTYPE, W * H)
        hits = d_hits.download(B.HIT_DTYPE, W * H)
        ref = orc.tlas_trace(tnodes, inst, [(nodes, deformed)], rays.view(orc.RAY_DTYPE), nthreads=8)
        assert_hits_equal(hits, ref, f"frame {frame} primary")
        d_b, d_cnt = B.bounce_tlas(ctx, tlas, d_rays, d_hits, W * H, seed=frame)
        cnt = int(d_cnt.download(np.uint32, 1)[0])
        assert cnt == (hits["prim"] != B.MISS).sum() > 100
        brays = d_b.download(B.RAY_DTYPE, cnt)
        got = tlas.intersect(brays)
        ref = orc.tlas_trace(tnodes, inst, [(nodes, deformed)], brays.view(orc.RAY_DTYPE), nthreads=8)
        assert_hits_equal(got, ref, f"frame {frame} bounce")


def test_set_transform_then_rebuild(ctx, orc):
    models, blases, objs = build_scene(ctx, orc, 8, seed=3)
    tlas = B.TlasBuilder(ctx=ctx).build_for(objs, models)
    M, Minv = scenes.random_affines(8, seed=99, trans_range=10.0)
    for k, o in enumerate(objs):
        o.set_transform(M[k].reshape(4, 4), Minv[k])
    tlas.rebuild(objs)
    tnodes, used = oracle_tlas(orc, blases, objs)
    g_nodes, g_used = tlas.download()
    assert g_nodes.tobytes() == tnodes.tobytes() and g_used == used


def test_cube_and_quad_known_answers_on_gpu(ctx, orc):
    """bvhtracer/tests/test_scene_cube.rs ids 6 / 9 and the exact quad depths
    of bvhtracer/tests/test_scene_quad.rs through the GPU TLAS path."""
    import json
    import os

    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "golden", "cube_obj.json")) as f:
        g = json.load(f)
    v = np.array(g["vertices"], np.float32)
    cube = v[np.array(g["faces"]) - 1].reshape(-1, 9).astype(np.float32)
    model = B.BvhBuilder(B.BUILD_REFERENCE, ctx=ctx).build_for(cube)
    m = np.eye(4, dtype=np.float32)
    m[:3, :3] *= 2
    m[3, :3] = -1
    minv = np.eye(4, dtype=np.float32)
    minv[:3, :3] *= 0.5
    minv[3, :3] = 0.5
    tlas = B.TlasBuilder(ctx=ctx).build_for([B.SceneObject(0, m, minv)], [model])
    for target, prim in [((0.5, 1.0, -0.5), 6), ((-0.5, 1.0, 0.5), 9)]:
        o = np.array([0, 4, 0], np.float32)
        r = B.Ray.from_origin_dir(o, orc.normalize(np.array(target, np.float32) - o))
        h = tlas.intersect(r)[0]
        assert h["prim"] == prim and h["inst"] == 0 and h["packed"] == prim
        assert abs(float(h["t"]) - float(np.sqrt(np.float32(19) / np.float32(2)))) <= 3.7e-7


def test_tlas_with_wide_leaves_and_infinite_rays(ctx, orc):
    """A TLAS over a reference-built BLAS (leaves of hundreds of triangles:
    unpacked stack entries) and an SAH BLAS, traced with rays of t = FLT_MAX,
    t = +inf and short t in mixed waves: the same hits and counters as the
    oracle."""
    soup = scenes.triangle_soup(4000, seed=23)
    ref_nodes, ref_used, ref_perm = orc.bvh_build(soup, True, 0)
    assert ref_nodes["count"][:ref_used].max() >= 64
    sph = scenes.icosphere(10, radius=2.0, seed=7)
    s_nodes, s_used, s_perm = orc.bvh_build(sph, False, 2)
    models = [B.Bvh.from_arrays(ref_nodes, ref_used, ref_perm, ctx=ctx),
              B.Bvh.from_arrays(s_nodes, s_used, s_perm, ctx=ctx)]
    blases = [(ref_nodes, ref_perm), (s_nodes, s_perm)]
    M, Minv = scenes.random_affines(12, 77, trans_range=8.0)
    objs = [B.SceneObject(k % 2, M[k].reshape(4, 4), Minv[k]) for k in range(12)]
    tlas = B.TlasBuilder(ctx=ctx).build_for(objs, models)
    tnodes, used = oracle_tlas(orc, blases, objs)
    cam = scenes.camera(eye=(0, 0, -40), forward=(0, 0, 1), right=(1, 0, 0), up=(0, 1, 0), fovy_deg=50)
    rays = orc.raygen(cam, 160, 160)
    k = np.arange(rays.shape[0])
    rays["tmax"] = np.where(k % 3 == 0, np.float32(np.inf), np.where(k % 7 == 0, np.float32(30.0), rays["tmax"]))
    rays["tmax"][: 64 * 40] = np.finfo(np.float32).max  # whole waves of finite t
    inst = oracle_instances(orc, objs)
    ref, st_ref = orc.tlas_trace(tnodes, inst, blases, rays, nthreads=8, stats=True)
    prays = np.zeros(rays.shape[0], B.RAY_DTYPE)
    prays.view(np.uint8)[:] = rays.view(np.uint8)
    got, st = tlas.intersect(prays, stats=True)
    assert_hits_equal(got, ref, "wide-leaf TLAS")
    assert np.array_equal(st.view(np.uint32), st_ref.view(np.uint32))
    assert len(np.unique(got["inst"][got["prim"] != B.MISS])) > 4


def _hand_cases():
    import json
    import os

    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tlas_hand_trace.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("name", [c["name"] for c in _hand_cases()["cases"]])
def test_tlas_rebuild_paths_equal_hand_trace(ctx, orc, name):
    """Tlas::rebuild traced by hand (tests/golden/tlas_hand_trace.json, from
    scene/tlas.rs:179-250): the synchronous rebuild, the device-only rebuild
    and the batched rebuild each give the hand-derived node array bit for
    bit.  Before each path the objects are moved (reversed offsets) and
    rebuilt, so the path under test starts from another tree."""
    from bvhtracer_amd import _lib as L
    from bvhtracer_amd import api

    fx = _hand_cases()
    case = next(c for c in fx["cases"] if c["name"] == name)
    want = np.zeros(len(case["nodes"]), B.TNODE_DTYPE)
    for k, n in enumerate(case["nodes"]):
        want[k]["bmin"], want[k]["bmax"] = n["bmin"], n["bmax"]
        want[k]["left_right"], want[k]["blas"] = n["left_right"], n["blas"]
    tri = np.asarray(fx["model_triangle"], np.float32).reshape(1, 9)
    model = B.BvhBuilder(B.BUILD_REFERENCE, ctx=ctx).build_for(tri.copy())
    xs = case["offsets_x"]
    M, Minv = translated([(x, 0.0, 0.0) for x in xs])
    objs = [B.SceneObject(0, M[k].reshape(4, 4), Minv[k]) for k in range(len(xs))]
    Mr, Minvr = translated([(x, 0.0, 0.0) for x in xs[::-1]])
    moved = [B.SceneObject(0, Mr[k].reshape(4, 4), Minvr[k]) for k in range(len(xs))]
    tlas = B.TlasBuilder(ctx=ctx).build_for(objs, [model])
    g, used = tlas.download()
    assert used == case["nodes_used"] and g.tobytes() == want.tobytes(), "TlasBuilder::build_for"

    def set_objects(o):
        inst = api._instances(o)
        api.check(L.lib().bvht_tlas_set_objects(ctx.handle, tlas.handle, L.ptr(inst), 0, inst.shape[0]))

    tlas.rebuild(moved)
    tlas.rebuild(objs)
    g, used = tlas.download()
    assert used == case["nodes_used"] and g.tobytes() == want.tobytes(), "rebuild"
    set_objects(moved)
    tlas.rebuild_async()
    set_objects(objs)
    tlas.rebuild_async()
    g, used = tlas.download()
    assert used == case["nodes_used"] and g.tobytes() == want.tobytes(), "rebuild_async"
    batch = B.Batch(ctx, [tlas])
    set_objects(moved)
    batch.rebuild()
    set_objects(objs)
    batch.rebuild()
    g, used = tlas.download()
    assert used == case["nodes_used"] and g.tobytes() == want.tobytes(), "batch rebuild"
    batch.close()
    if "hit_inst" in case:
        # equal distances descend right() first; an equal t is not accepted
        ray = np.zeros(1, B.RAY_DTYPE)
        ray["o"], ray["d"] = fx["ray"]["o"], fx["ray"]["d"]
        ray["tmax"] = np.finfo(np.float32).max
        h, st = tlas.intersect(ray, stats=True)
        assert h[0]["t"] == np.float32(fx["hit"]["t"]) and h[0]["u"] == np.float32(fx["hit"]["u"])
        assert h[0]["v"] == np.float32(fx["hit"]["v"]) and h[0]["prim"] == fx["hit"]["prim"]
        assert h[0]["inst"] == case["hit_inst"]
        assert st[0]["tlas_pairs"] == case["stats"]["tlas_pairs"]
        assert st[0]["inst_visits"] == case["stats"]["inst_visits"]


def test_tlas_wave_schedule_keeps_hits(orc):
    """With bvht_set_wave_schedule on, repeated traces of the same rays run
    their 64-ray groups longest first from the second call (the order is
    rebuilt every 8 calls); every call equals the unscheduled trace and the
    oracle bit for bit, a different ray buffer in between included, and so
    does a captured trace replayed with eager work between the replays."""
    ctx = B.Context(0)
    try:
        models, blases, objs = build_scene(ctx, orc, 16, seed=7, freq=14)
        tlas = B.TlasBuilder(ctx=ctx).build_for(objs, models)
        cam = scenes.camera(eye=(0, 0, -45), forward=(0, 0, 1), right=(1, 0, 0), up=(0, 1, 0), fovy_deg=45,
                            aspect=16 / 9)
        W, H = 320, 184
        n = W * H
        d_rays = B.raygen(ctx, cam, W, H, B.ORDER_TILES)
        d_hits = ctx.alloc(n * B.HIT_DTYPE.itemsize)
        tlas.intersect_device(d_rays, n, d_hits)
        want = d_hits.download(B.HIT_DTYPE, n)
        rays = d_rays.download(B.RAY_DTYPE, n)
        tnodes, _ = oracle_tlas(orc, blases, objs)
        ref = orc.tlas_trace(tnodes, oracle_instances(orc, objs), blases, rays.view(orc.RAY_DTYPE), nthreads=8)
        assert_hits_equal(want, ref, "unscheduled")
        assert (want["prim"] != B.MISS).sum() > 1000
        other = rays.copy()
        other["o"][:, 0] += 0.25
        d_other = ctx.to_device(other)
        want_other = tlas.intersect(other)
        ctx.set_wave_schedule(True)
        for call in range(12):
            d_hits.upload(np.zeros(n, B.HIT_DTYPE))
            tlas.intersect_device(d_rays, n, d_hits)
            assert d_hits.download(B.HIT_DTYPE, n).tobytes() == want.tobytes(), call
            if call == 5:  # another buffer breaks the key; the next call starts over
                d_o = ctx.alloc(n * B.HIT_DTYPE.itemsize)
                tlas.intersect_device(d_other, n, d_o)
                assert d_o.download(B.HIT_DTYPE, n).tobytes() == want_other.tobytes()

        def enqueue():
            tlas.intersect_device(d_rays, n, d_hits)

        g = B.Graph.capture(ctx, enqueue, keep=(d_rays, d_hits))
        for _ in range(4):
            d_hits.upload(np.zeros(n, B.HIT_DTYPE))
            g.launch()
            ctx.sync()
            assert d_hits.download(B.HIT_DTYPE, n).tobytes() == want.tobytes()
            tlas.intersect_device(d_other, n, d_o)  # eager work between the replays
            assert d_o.download(B.HIT_DTYPE, n).tobytes() == want_other.tobytes()
            # an eager scheduled trace of another count rebuilds the order for
            # its own groups; the next replay must not take it
            for _ in range(2):
                tlas.intersect_device(d_rays, n // 3 + 7, d_o)
            assert d_o.download(B.HIT_DTYPE, n // 3 + 7).tobytes() == want[: n // 3 + 7].tobytes()
        g.close()
        ctx.set_wave_schedule(False)
        tlas.intersect_device(d_rays, n, d_hits)
        assert d_hits.download(B.HIT_DTYPE, n).tobytes() == want.tobytes()
    finally:
        ctx.close()
