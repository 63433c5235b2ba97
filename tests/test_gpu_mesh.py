"""The reference's real mesh assets through the GPU path: unity.tri built
as BvhBuilder builds it (the construction benchmark) and traced; bigben.tri
deformed with the refit benchmark's animate(0.05), then refit and rebuilt.
Everything against the oracle, bit for bit."""
import os

import numpy as np
import pytest

import bvhtracer_amd as B
from bvhtracer_amd import mesh as M
from bvhtracer_amd import scenes
from bvhtracer_amd.scene import ModelBuilder
from helpers import assert_hits_equal, assert_nodes_equal

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_unity_build_and_trace(ctx, orc):
    tris = M.load_tri(os.path.join(GOLDEN, "unity.tri.gz"))
    mesh = M.Mesh(tris.copy())                             # the build permutes the mesh in place
    model = ModelBuilder(ctx).with_mesh(mesh).build()    # BvhBuilder::new(): the reference build
    ref_nodes, ref_used, ref_tris = orc.bvh_build(tris.copy(), True, 0)
    nodes, used, got_tris = model.bvh.download()
    assert used == ref_used
    assert_nodes_equal(nodes, ref_nodes, used)
    assert got_tris.tobytes() == ref_tris.tobytes() == model.primitives().tobytes()
    v = tris.reshape(-1, 3)
    v = v[(v < 900).all(1)]  # the asset ends with a sentinel triangle at (999, 999, 999)
    bmin, bmax = v.min(0), v.max(0)
    c = (bmin + bmax) / 2
    ext = float(np.max(bmax - bmin))
    cam = scenes.camera(eye=(c[0], c[1], c[2] - 2.0 * ext), forward=(0, 0, 1), right=(1, 0, 0), up=(0, 1, 0),
                        fovy_deg=45)
    rays = orc.raygen(cam, 320, 320)
    got = model.intersect(rays.view(B.RAY_DTYPE))
    assert_hits_equal(got, orc.bvh_trace(ref_nodes, ref_tris, rays, nthreads=8), "unity")
    assert (got["prim"] != B.MISS).sum() > 5000


@pytest.mark.parametrize("mode,max_leaf", [(B.BUILD_REFERENCE, 0), (B.BUILD_SAH, 2)])
def test_bigben_animate_refit_and_rebuild(ctx, orc, mode, max_leaf):
    tris = M.load_tri(os.path.join(GOLDEN, "bigben.tri.gz"))
    model = ModelBuilder(ctx).with_mesh(tris.copy()).with_build(mode, max_leaf).build()
    nodes0, used, built = model.bvh.download()
    animated = scenes.animate(built, 0.05)
    model.primitives_mut()[:] = animated
    model.refit()                                          # ModelInstance::refit
    nodes, used2, t = model.bvh.download()
    ref = orc.bvh_refit(nodes0.copy(), used, animated)
    assert used2 == used and t.tobytes() == animated.tobytes()
    assert_nodes_equal(nodes, ref, used)
    # the rebuild benchmark: a fresh build of the animated mesh
    rebuilt = ModelBuilder(ctx).with_mesh(scenes.animate(tris, 0.05)).with_build(mode, max_leaf).build()
    rn, ru, rt = orc.bvh_build(scenes.animate(tris, 0.05), mode == B.BUILD_REFERENCE, max_leaf)
    gn, gu, gt = rebuilt.bvh.download()
    assert gu == ru and gt.tobytes() == rt.tobytes()
    assert_nodes_equal(gn, rn, gu)


def test_decoded_attributes_follow_the_build_permutation(ctx):
    mesh = M.open_mesh(os.path.join(GOLDEN, "cube.obj"))
    normals_before = mesh.normals().copy()
    prims_before = mesh.primitives().copy()
    model = ModelBuilder(ctx).with_mesh(mesh).build()
    perm = model.bvh.permutation()
    assert model.primitives().tobytes() == prims_before[perm].tobytes()
    assert model.mesh().normals().tobytes() == normals_before[perm].tobytes()
