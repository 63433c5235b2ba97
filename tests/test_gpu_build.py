"""GPU BLAS build and refit parity: node arrays, nodes_used and the triangle
permutation equal the oracle's (bvh.rs:333-541, 469-493)."""
import json
import os

import numpy as np
import pytest

import bvhtracer_amd as B
from bvhtracer_amd import scenes
from helpers import assert_hits_equal, assert_nodes_equal, upload

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def unit_triangle():
    s3 = np.float32(1) / np.sqrt(np.float32(3))
    return np.array([0, 0.5, 0, -s3, -0.5, 0, s3, -0.5, 0], np.float32)


def cube_mesh():
    with open(os.path.join(HERE, "golden", "cube_obj.json")) as f:
        g = json.load(f)
    v = np.array(g["vertices"], np.float32)
    return v[np.array(g["faces"]) - 1].reshape(-1, 9).astype(np.float32)


def meshes():
    base = unit_triangle().reshape(3, 3)
    diag = np.stack([(base + np.float32(i) * np.array([5, 5, 0], np.float32)).reshape(9) for i in range(-100, 100)])
    stacked = np.stack([(base - np.array([0, 0, i], np.float32)).reshape(9) for i in range(100)])
    same = np.tile(unit_triangle(), (40, 1))  # identical centroids: no usable axis
    flat = scenes.triangle_soup(300, seed=8)
    flat[:, 2::3] = 0.0  # every z is 0: the z axis is skipped
    # non-finite coordinates: NaN, +-inf and f32::MAX vertices among finite ones
    nonfinite = scenes.triangle_soup(300, seed=9)
    nonfinite[3, 0] = np.nan
    nonfinite[40, 4:6] = np.nan
    nonfinite[77, 8] = np.inf
    nonfinite[120, 1] = -np.inf
    nonfinite[121, 3:6] = np.float32(3.4028235e38)
    nonfinite[200] = np.nan
    # 60 tight clusters of 17..45 triangles far apart: subtrees of either
    # side of the 32-element boundary between the level passes and the wave
    # builder
    rng = np.random.default_rng(11)
    clusters = np.concatenate([
        (rng.random((int(k), 9), dtype=np.float32) * np.float32(0.5)
         + np.tile(rng.random(3, dtype=np.float32) * np.float32(400), 3))
        for k in rng.integers(17, 46, 60)]).astype(np.float32)
    # 5000 coplanar triangles on a coarse grid: many equal centroids
    g = np.arange(5000, dtype=np.float32)
    plane = np.zeros((5000, 9), np.float32)
    plane[:, 0] = (g % 50) // 5
    plane[:, 1] = g // 500
    plane[:, 3] = plane[:, 0] + 1
    plane[:, 4] = plane[:, 1]
    plane[:, 6] = plane[:, 0]
    plane[:, 7] = plane[:, 1] + 1
    return {
        "nonfinite": nonfinite,
        "clusters": clusters,
        "plane": plane,
        "one": unit_triangle().reshape(1, 9),
        "two": np.stack([(base - [5, 0, 0]).reshape(9), (base + [5, 0, 0]).reshape(9)]).astype(np.float32),
        "cube": cube_mesh(),
        "diagonal": diag,
        "stacked": stacked,
        "identical": same,
        "flat": flat,
        "soup33": scenes.triangle_soup(33, seed=4),
        "soup1000": scenes.triangle_soup(1000, seed=6),
        "soup": scenes.triangle_soup(),
        "sphere16": scenes.icosphere(16, radius=32.0),
    }


@pytest.mark.parametrize("mode,max_leaf", [(B.BUILD_REFERENCE, 0), (B.BUILD_SAH, 2), (B.BUILD_SAH, 0),
                                           (B.BUILD_SAH, 7)])
@pytest.mark.parametrize("name", list(meshes().keys()))
def test_build_matches_oracle(ctx, orc, name, mode, max_leaf):
    mesh = meshes()[name]
    faithful = mode == B.BUILD_REFERENCE
    nodes, used, perm = orc.bvh_build(mesh, faithful, max_leaf)
    m = np.ascontiguousarray(mesh.copy())
    bvh = B.BvhBuilder(mode, max_leaf, ctx=ctx).build_for(m)
    assert m.tobytes() == perm.tobytes(), "triangle permutation (written back in place) differs"
    g_nodes, g_used, g_tris = bvh.download()
    assert g_used == used == bvh.nodes_used()
    assert_nodes_equal(g_nodes, nodes, used)
    assert g_tris.tobytes() == perm.tobytes()
    assert bvh.depth() == orc.bvh_depth(nodes)


def test_cube_ids_through_gpu_build(ctx, orc):
    """IDs 6 and 9 of bvhtracer/tests/test_scene_cube.rs through the GPU build."""
    m = cube_mesh()
    bvh = B.BvhBuilder(B.BUILD_REFERENCE, ctx=ctx).build_for(m)
    for target, prim in [((0.5, 1.0, -0.5), 6), ((-0.5, 1.0, 0.5), 9)]:
        o = np.array([0, 4, 0], np.float32)
        t = np.array(target, np.float32)
        # model space of the x2 scale, -1 translation object: o' = (o + 1) / 2
        om = (o + np.float32(1)) * np.float32(0.5)
        d = orc.normalize(t - o) * np.float32(0.5)
        h = bvh.intersect(B.Ray.from_origin_dir(om, d))[0]
        assert h["prim"] == prim


@pytest.mark.parametrize("mode,max_leaf", [(B.BUILD_REFERENCE, 0), (B.BUILD_SAH, 2)])
def test_sphere_1m_build_matches_oracle(ctx, orc, mode, max_leaf):
    """Build parity at the 1,310,720-triangle size of the incoherent-ray workload."""
    mesh = scenes.icosphere(256, radius=32.0)
    faithful = mode == B.BUILD_REFERENCE
    nodes, used, perm = orc.bvh_build(mesh, faithful, max_leaf)
    m = mesh.copy()
    bvh = B.BvhBuilder(mode, max_leaf, ctx=ctx).build_for(m)
    g_nodes, g_used, g_tris = bvh.download()
    assert g_used == used
    assert_nodes_equal(g_nodes, nodes, used)
    assert g_tris.tobytes() == perm.tobytes()


@pytest.mark.parametrize("mode,max_leaf", [(B.BUILD_REFERENCE, 0), (B.BUILD_SAH, 2)])
def test_refit_matches_oracle(ctx, orc, mode, max_leaf):
    mesh = scenes.icosphere(40, radius=10.0)
    m = mesh.copy()
    bvh = B.BvhBuilder(mode, max_leaf, ctx=ctx).build_for(m)
    nodes, used, perm = bvh.download()
    for frame in (0, 5, 63):
        deformed = scenes.deform_wave(perm, frame)
        bvh.refit(deformed)
        orc.bvh_refit(nodes, used, deformed)
        g_nodes, g_used, _ = bvh.download()
        assert_nodes_equal(g_nodes, nodes, used)
    # refit on unchanged vertices is the identity (bvhtracer/tests/test_bvh_refit.rs:46-53)
    before = bvh.download()[0]
    bvh.refit()
    assert bvh.download()[0].tobytes() == before.tobytes()


def test_build_then_trace(ctx, orc):
    mesh = scenes.triangle_soup()
    m = mesh.copy()
    bvh = B.BvhBuilder(B.BUILD_SAH, 2, ctx=ctx).build_for(m)
    nodes, used, perm = orc.bvh_build(mesh, False, 2)
    rays = orc.raygen(scenes.soup_camera(), 256, 256)
    prays = np.zeros(rays.shape[0], B.RAY_DTYPE)
    prays.view(np.uint8)[:] = rays.view(np.uint8)
    assert_hits_equal(bvh.intersect(prays), orc.bvh_trace(nodes, perm, rays, nthreads=8), "built")


def test_build_rejects_empty(ctx):
    with pytest.raises(B.BvhtError):
        B.BvhBuilder(ctx=ctx).build_for(np.zeros((0, 9), np.float32))


def test_builds_on_reused_buffers_match_oracle(orc):
    """A destroyed BLAS's device buffers go back to its context and serve the
    next BLAS: a large build, a smaller one (reusing the large buffers'
    pool only where sizes are close), the large one again and an upload
    after a destroy all give the oracle's node arrays and triangles."""
    ctx = B.Context(0)
    try:
        big = scenes.icosphere(40, radius=10.0)
        small = scenes.icosphere(12, radius=3.0)
        ref = {k: orc.bvh_build(m, False, 2) for k, m in (("big", big), ("small", small))}
        for name, mesh in (("big", big), ("small", small), ("big", big), ("big", big)):
            m = np.ascontiguousarray(mesh.copy())
            bvh = B.BvhBuilder(B.BUILD_SAH, 2, ctx=ctx).build_for(m)
            nodes, used, perm = ref[name]
            g_nodes, g_used, g_tris = bvh.download()
            assert g_used == used
            assert_nodes_equal(g_nodes, nodes, used)
            assert g_tris.tobytes() == perm.tobytes()
            assert bvh.depth() == orc.bvh_depth(nodes)
            bvh.close()
        nodes, used, perm = ref["big"]
        up = upload(ctx, nodes, used, perm)
        g_nodes, g_used, g_tris = up.download()
        assert g_nodes.tobytes() == nodes.tobytes() and g_tris.tobytes() == perm.tobytes()
        up.close()
    finally:
        ctx.close()


def _large_nodes_per_level(nodes, used):
    """Nodes of more than 32 triangles at each depth of a node array (the
    levels the device build processes in its level loop)."""
    lf = nodes["left_first"][:used].astype(np.int64)
    cnt = nodes["count"][:used].astype(np.int64)
    levels, frontier = [], np.array([0])
    while frontier.size:
        levels.append(frontier)
        inner = frontier[cnt[frontier] == 0]
        frontier = np.concatenate([lf[inner], lf[inner] + 1])
    size = np.zeros(used, np.int64)
    for lv in reversed(levels):
        leaf = cnt[lv] > 0
        size[lv[leaf]] = cnt[lv[leaf]]
        inner = lv[~leaf]
        size[inner] = size[lf[inner]] + size[lf[inner] + 1]
    return [int((size[lv] > 32).sum()) for lv in levels]


def test_build_with_wide_levels_matches_oracle(ctx, orc):
    """4.8M triangles: the widest level holds more than 64 x 1024 large nodes,
    so each of the 64 workgroups that place a level takes several 1024-node
    rounds, and the level's task list takes over 64 rounds."""
    mesh = scenes.icosphere(490, radius=32.0)
    nodes, used, perm = orc.bvh_build(mesh, False, 2)
    assert max(_large_nodes_per_level(nodes, used)) > 64 * 1024
    m = mesh.copy()
    bvh = B.BvhBuilder(B.BUILD_SAH, 2, ctx=ctx).build_for(m)
    g_nodes, g_used, g_tris = bvh.download()
    assert g_used == used
    assert_nodes_equal(g_nodes, nodes, used)
    assert g_tris.tobytes() == perm.tobytes()
    assert bvh.depth() == orc.bvh_depth(nodes)
    bvh.close()


@pytest.mark.parametrize("freq,mode,max_leaf", [(128, B.BUILD_SAH, 2), (128, B.BUILD_REFERENCE, 0),
                                                 (181, B.BUILD_SAH, 2), (181, B.BUILD_SAH, 0)])
def test_build_mid_sizes_match_oracle(ctx, orc, freq, mode, max_leaf):
    """327,680 and 655,220 triangles: the level kernels cut node ranges into
    chunks of 1024 elements here (256 at the 12.6k soup, 512 at 82k, 2048 at
    1.3M and 4096 at 4.8M, all covered above), the copy back (11 and 22 MB)
    runs on the second stream beside the numbering kernels, and SAH mode
    without a leaf bound stops splitting on its own."""
    mesh = scenes.icosphere(freq, radius=16.0)
    faithful = mode == B.BUILD_REFERENCE
    nodes, used, perm = orc.bvh_build(mesh, faithful, max_leaf)
    m = mesh.copy()
    bvh = B.BvhBuilder(mode, max_leaf, ctx=ctx).build_for(m)
    g_nodes, g_used, g_tris = bvh.download()
    assert g_used == used
    assert_nodes_equal(g_nodes, nodes, used)
    assert g_tris.tobytes() == perm.tobytes()
    assert m.tobytes() == perm.tobytes()  # the caller's array holds the permuted triangles
    assert bvh.depth() == orc.bvh_depth(nodes)
    bvh.close()
