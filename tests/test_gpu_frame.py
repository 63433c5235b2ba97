"""GPU frame path and host mirror: per-pair triangle tests, the tiled
PathTracer frame (raygen, trace, accumulator, pixel shader on the device)
and the Model / Scene mirror, all against the oracle."""
import numpy as np
import pytest

import bvhtracer_amd as B
from bvhtracer_amd import render, scenes
from bvhtracer_amd.scene import ModelBuilder, SceneBuilder, SceneObjectBuilder, Triangle, intersect_pairs
from helpers import assert_hits_equal
from oracle import frame as oframe

pytestmark = pytest.mark.gpu

F32_MAX = np.finfo(np.float32).max


def test_intersect_pairs_equals_oracle(ctx, orc):
    rng = np.random.default_rng(7)
    n = 20000
    v0 = rng.uniform(-5, 5, (n, 3))
    tris = np.concatenate([v0, v0 + rng.uniform(-1, 1, (n, 3)), v0 + rng.uniform(-1, 1, (n, 3))], 1)
    tris = tris.astype(np.float32)
    rays = np.zeros(n, B.RAY_DTYPE)
    rays["o"] = rng.uniform(-8, 8, (n, 3))
    target = tris.reshape(n, 3, 3).mean(1) + rng.normal(0, 0.3, (n, 3))
    d = target - rays["o"]
    rays["d"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays["tmax"] = np.where(np.arange(n) % 4 == 0, 3.0, F32_MAX)  # some hits beyond tmax: t = min(tmax, t)
    got = intersect_pairs(tris, rays, ctx)
    hit = 0
    for i in range(n):
        r = orc.triangle_intersect(tris[i], rays[i:i + 1].view(orc.RAY_DTYPE))
        if r is None:
            assert got[i]["prim"] == B.MISS
        else:
            hit += 1
            assert got[i]["prim"] == i
            assert [got[i]["t"], got[i]["u"], got[i]["v"]] == [np.float32(x) for x in r], i
    assert hit > 2000
    tri = Triangle(tris[0, :3], tris[0, 3:6], tris[0, 6:9])
    res = tri.intersect(rays[:50], ctx)
    for i in range(50):
        r = orc.triangle_intersect(tris[0], rays[i:i + 1].view(orc.RAY_DTYPE))
        assert (res[i] is None) == (r is None)


@pytest.fixture(scope="module")
def soup_scene(ctx, orc):
    mesh = scenes.triangle_soup()
    bvh = B.BvhBuilder(B.BUILD_REFERENCE, ctx=ctx).build_for(mesh)
    nodes, used, tris = bvh.download()
    cam = scenes.soup_camera()
    rays = orc.raygen(cam, 640, 640)  # row order
    hits = orc.bvh_trace(nodes, tris, rays, nthreads=8)
    return bvh, cam, hits


@pytest.mark.parametrize("acc,shader", [
    (render.DepthAccumulator(), render.DepthMappingShader(80.0, 3.0)),
    (render.DepthAccumulator(), render.DepthMappingShader(8.0, 10.0)),
    (render.IntersectionAccumulator((1.0, 0.5, 0.25), (0.0, 0.0, 0.0)), render.RadianceToRgbShader()),
    (render.IntersectionAccumulator((1.0, 1.0, 1.0), (0.0, 0.0, 0.0)),
     render.IntersectionShader((255, 128, 0, 255), (0, 0, 64, 255))),
    (render.UvMappingAccumulator(), render.RadianceToRgbShader()),
])
def test_path_tracer_frame_equals_oracle(ctx, soup_scene, acc, shader):
    bvh, cam, hits = soup_scene
    state = render.RendererState(acc, shader, 640, 640, ctx=ctx)
    n = render.Renderer(render.PathTracer()).render(state, bvh, cam)
    assert n == 640 * 640
    kinds = {type(render.DepthAccumulator()): "depth", type(render.UvMappingAccumulator()): "uv"}
    if isinstance(acc, render.IntersectionAccumulator):
        ref_acc = oframe.accumulate(hits, 640, 640, "intersection", acc.hit_value, acc.miss_value, order="rows")
    else:
        ref_acc = oframe.accumulate(hits, 640, 640, kinds[type(acc)], order="rows")
    got_acc = state.accumulation_buffer().reshape(-1, 3)
    assert got_acc.tobytes() == ref_acc.tobytes()
    if isinstance(shader, render.DepthMappingShader):
        ref = oframe.shade(ref_acc, "depth", scale=shader.scale, offset=shader.offset)
    elif isinstance(shader, render.IntersectionShader):
        ref = oframe.shade(ref_acc, "intersection", hit_rgba=shader.hit_value, miss_rgba=shader.miss_value)
    else:
        ref = oframe.shade(ref_acc, "radiance")
    got = state.frame_buffer().reshape(-1, 4)
    assert np.array_equal(got, ref)
    assert len(np.unique(got.view(np.uint32))) >= 2  # hits and misses both present


def test_scene_mirror_equals_oracle(ctx, orc):
    m1 = ModelBuilder(ctx).with_mesh(scenes.icosphere(10, radius=2.0)).with_build(B.BUILD_SAH, 2).build()
    m2 = ModelBuilder(ctx).with_mesh(scenes.icosphere(6, radius=1.0, seed=3)).build()
    M, Minv = scenes.random_affines(6, 0x5C, trans_range=6.0)
    objs = [SceneObjectBuilder(m1 if k % 2 else m2).with_transform(M[k]).build() for k in range(6)]
    cam = scenes.camera(eye=(0, 0, -25), forward=(0, 0, 1), right=(1, 0, 0), up=(0, 1, 0), fovy_deg=60)
    scene = SceneBuilder(cam, ctx=ctx).with_objects(objs).build()
    assert [o.model for o in objs] == [0, 1, 0, 1, 0, 1]
    rays = orc.raygen(cam, 128, 128)

    def oracle_hits():
        nodes1, _, t1 = m2.bvh.download()
        nodes2, _, t2 = m1.bvh.download()
        tnodes, _ = scene.tlas.download()
        inst = np.zeros(6, orc.INSTANCE_DTYPE)
        for k, o in enumerate(objs):
            inst["minv"][k] = o.minv
            inst["blas"][k] = o.model
        return orc.tlas_trace(tnodes, inst, [(nodes1, t1), (nodes2, t2)], rays, nthreads=8)

    got = scene.intersect(rays.view(B.RAY_DTYPE))
    assert_hits_equal(got, oracle_hits(), "scene")
    assert (got["prim"] != B.MISS).sum() > 500
    # move an object, shrink a model in place, refit, rebuild
    objs[0].set_transform(np.eye(4, dtype=np.float32))
    m1.primitives_mut()[:] *= np.float32(0.5)
    m1.refit()
    scene.rebuild()
    got2 = scene.intersect(rays.view(B.RAY_DTYPE))
    assert_hits_equal(got2, oracle_hits(), "scene after refit + rebuild")
    assert got2.tobytes() != got.tobytes()


def test_cube_scene_normal_mapping_known_answer(ctx):
    """test_scene_cube.rs:139-175,230-265: rays from (0, 4, 0) toward
    (0.5, 1, -0.5) and (-0.5, 1, 0.5) hit the top face of cube.obj scaled by 2
    and moved by (-1, -1, -1); the world normal of the hit triangle's corners
    is +y.  With the attributes permuted like the triangles, the
    normal-mapping accumulator gives (0.5, 1, 0.5) for both (the reference
    reads the unpermuted normal array there)."""
    import os

    from bvhtracer_amd import mesh as M

    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    model = ModelBuilder(ctx).with_mesh(M.open_mesh(os.path.join(golden, "cube.obj"))).build()
    T = np.diag([2.0, 2.0, 2.0, 1.0]).astype(np.float32)
    T[3, :3] = -1.0
    scene = SceneBuilder(None, ctx=ctx).with_object(SceneObjectBuilder(model).with_transform(T.reshape(16)).build()) \
        .build()
    o = np.array([0.0, 4.0, 0.0], np.float32)
    rays = np.zeros(2, B.RAY_DTYPE)
    for k, target in enumerate(([0.5, 1.0, -0.5], [-0.5, 1.0, 0.5])):
        d = np.asarray(target, np.float32) - o
        rays[k]["o"], rays[k]["d"], rays[k]["tmax"] = o, d / np.sqrt(np.float32((d * d).sum())), F32_MAX
    hits = scene.intersect(rays)
    assert hits["prim"].tolist() == [6, 9]  # test_scene_cube.rs:124-135,215-226
    d_hits = ctx.to_device(hits)
    d_acc = ctx.alloc(2 * 12)
    B._lib.check(B._lib.lib().bvht_accumulate_tlas(ctx.handle, scene.tlas.handle, d_hits.ptr, 2, 1, B.ORDER_ROWS,
                                                   B._lib.ACC_NORMAL, d_acc.ptr))
    assert d_acc.download(np.float32).reshape(2, 3).tolist() == [[0.5, 1.0, 0.5], [0.5, 1.0, 0.5]]


def sphere_attributes(tris):
    """Corner normals (radial) and spherical texture coordinates of a
    sphere-like mesh, binary32."""
    v = tris.reshape(-1, 3, 3).astype(np.float64)
    n = v / np.linalg.norm(v, axis=2, keepdims=True)
    uv = np.stack([0.5 + np.arctan2(n[..., 2], n[..., 0]) / (2 * np.pi), 0.5 - np.arcsin(n[..., 1]) / np.pi], -1)
    return n.astype(np.float32), (uv * 3.0).astype(np.float32)  # texture repeats 3x


@pytest.mark.parametrize("acc", ["normal", "texture"])
def test_attribute_accumulators_frame_equal_oracle(ctx, acc):
    """Two models (different spheres, own textures), five objects with
    rotations and non-uniform scales: the frame's accumulation buffer equals
    the oracle's, bit for bit, through the PathTracer and the TLAS."""
    from bvhtracer_amd.mesh import Mesh

    rng = np.random.default_rng(11)
    models, attrs = [], []
    for k, (freq, seed) in enumerate(((8, 1), (5, 2))):
        tris = scenes.icosphere(freq, radius=2.0, seed=seed)
        nrm, uv = sphere_attributes(tris)
        tex = rng.integers(0, 256, (16 + 8 * k, 32, 3), dtype=np.uint8)
        m = ModelBuilder(ctx).with_mesh(Mesh(tris, uv, nrm)).with_texture(tex).with_build(B.BUILD_SAH, 2).build()
        models.append(m)
        attrs.append({"normals": m.mesh().normals(), "tex_coords": m.mesh().tex_coords(), "texture": tex})
    M, _ = scenes.random_affines(5, 0xA77, scale_range=(0.6, 1.4), trans_range=5.0)
    M = M.reshape(5, 4, 4).copy()
    M[:, 0, :3] *= np.float32(1.3)  # non-uniform scale: normals still go through M, as in the reference
    objs = [SceneObjectBuilder(models[k % 2]).with_transform(M[k].reshape(16)).build() for k in range(5)]
    cam = scenes.camera(eye=(0, 0, -14), forward=(0, 0, 1), right=(1, 0, 0), up=(0, 1, 0), fovy_deg=50)
    scene = SceneBuilder(cam, ctx=ctx).with_objects(objs).build()
    accum = render.NormalMappingAccumulator() if acc == "normal" else render.TextureMaterialAccumulator()
    state = render.RendererState(accum, render.RadianceToRgbShader(), 64, 64, ctx=ctx)
    render.PathTracer().evaluate(state, scene)
    hits = state.d_hits.download(B.HIT_DTYPE, 64 * 64)
    assert (hits["prim"] != B.MISS).sum() > 500 and len(set(hits["inst"][hits["prim"] != B.MISS])) >= 3
    ref = oframe.accumulate_attributes(hits, 64, 64, acc, attrs, [(o.m, o.model) for o in objs], order="tiles")
    got = state.accumulation_buffer().reshape(-1, 3)
    assert got.tobytes() == ref.tobytes()
    assert np.array_equal(state.frame_buffer().reshape(-1, 4), oframe.shade(ref, "radiance"))
