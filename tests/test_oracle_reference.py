"""The reference crate's own tests for the traversal path, restated against
the CPU oracle (paths relative to the reference repository root).

These pin the oracle: exact t values, hit/miss maps and primitive ids the
reference asserts, evaluated in binary32 exactly as the reference computes
its inputs (numpy float32 scalars round every operation).
"""
import json
import math
import os

import numpy as np
import pytest

f32 = np.float32
F32_MAX = np.finfo(np.float32).max
HERE = os.path.dirname(os.path.abspath(__file__))


def unit_triangle():
    # Triangle of bvhtracer/tests/test_triangle_intersection.rs:11-17
    s3 = f32(1) / np.sqrt(f32(3))
    return np.array([0, 0.5, 0, -s3, -0.5, 0, s3, -0.5, 0], np.float32)


def ray_towards(orc, origin, target, tmax=F32_MAX):
    o = np.asarray(origin, np.float32)
    d = orc.normalize(np.asarray(target, np.float32) - o)
    return orc.make_ray(o, d, tmax)


# ---- bvhtracer/tests/test_triangle_intersection.rs ----------------------

def test_triangle_hits_center_and_vertices(orc):
    tri = unit_triangle()
    o = (0, 0, 5)
    assert orc.triangle_intersect(tri, ray_towards(orc, o, (0, 0, 0))) is not None  # :19-29
    for v in (tri[0:3], tri[3:6], tri[6:9]):  # :31-65
        assert orc.triangle_intersect(tri, ray_towards(orc, o, v)) is not None


def test_triangle_exact_t(orc):
    tri = unit_triangle()
    o = (0, 0, 5)
    # :67-80 t == 5 (the reference compares whole Rays, so recip_direction too)
    r = ray_towards(orc, o, (0, 0, 0))
    assert orc.triangle_intersect(tri, r)[0] == f32(5)
    # :82-95 t == sqrt(101/4)
    assert orc.triangle_intersect(tri, ray_towards(orc, o, tri[0:3]))[0] == np.sqrt(f32(101) / f32(4))
    # :97-125 t == sqrt(307/12) for both lower vertices
    e = np.sqrt(f32(307) / f32(12))
    assert orc.triangle_intersect(tri, ray_towards(orc, o, tri[3:6]))[0] == e
    assert orc.triangle_intersect(tri, ray_towards(orc, o, tri[6:9]))[0] == e


def test_triangle_vertex_misses(orc):
    tri = unit_triangle()
    o = (0, 0, 5)
    # :127-167
    targets = [tri[0:3] + np.array([0, 0.5, 0], np.float32), tri[3:6] + np.array([0, -0.5, 0], np.float32),
               tri[6:9] + np.array([0, -0.5, 0], np.float32)]
    for t in targets:
        assert orc.triangle_intersect(tri, ray_towards(orc, o, t)) is None


@pytest.mark.parametrize("ray_t", [F32_MAX, 100.0, 0.01])
def test_triangle_t_clamped_to_ray_t(orc, ray_t):
    # :169-207: the returned t is min(ray.t, t)
    tri = unit_triangle()
    r = ray_towards(orc, (0, 0, 5), (0, 0, 0), f32(ray_t))
    res = orc.triangle_intersect(tri, r)
    assert res is not None and res[0] <= f32(ray_t)
    if ray_t == 0.01:
        assert res[0] == f32(0.01)


# ---- bvhtracer/tests/test_aabb_intersection.rs ---------------------------

BOX = (np.array([-1, -1, -1], np.float32), np.array([1, 1, 1], np.float32))


@pytest.mark.parametrize("plane", ["xy", "yz", "zx"])
def test_aabb_hit_miss_sweeps(orc, plane):
    # :29-147; FRAC_2_PI is 2/pi, so the sweep covers [0, 2/pi) radians
    frac_2_pi = f32(2.0 / math.pi)
    for i in range(64):
        a = (f32(i) / f32(64)) * frac_2_pi
        c, s = f32(10) * np.cos(a), f32(10) * np.sin(a)
        o = {"xy": (c, s, 0), "yz": (0, c, s), "zx": (s, 0, c)}[plane]
        o = np.array(o, np.float32)
        away = orc.make_ray(o, orc.normalize(o - np.zeros(3, np.float32)))
        towards = orc.make_ray(o, orc.normalize(np.zeros(3, np.float32) - o))
        assert orc.aabb_intersect(*BOX, away) is None
        assert orc.aabb_intersect(*BOX, towards) is not None


@pytest.mark.parametrize("origin", [(5, 0, 0), (0, 5, 0), (0, 0, 5), (-5, 0, 0), (0, -5, 0), (0, 0, -5)])
def test_aabb_exact_t(orc, origin):
    # :149-188: t == 4 exactly on the six axis rays
    o = np.array(origin, np.float32)
    r = orc.make_ray(o, orc.normalize(np.zeros(3, np.float32) - o))
    assert orc.aabb_intersect(*BOX, r) == 4.0


# ---- BVH on one, two, three triangles -------------------------------------

def test_bvh_one_triangle_structure(orc):
    # bvhtracer/src/model/bvh.rs:779-831: one root leaf + dummy, nodes_used 2
    tri = unit_triangle().reshape(1, 9)
    nodes, used, perm = orc.bvh_build(tri)
    assert used == 2 and nodes.shape[0] == 2
    assert nodes[0]["count"] == 1 and nodes[0]["left_first"] == 0
    v = tri.reshape(3, 3)
    assert np.array_equal(nodes[0]["bmin"], v.min(axis=0)) and np.array_equal(nodes[0]["bmax"], v.max(axis=0))
    assert nodes[1].tobytes() == bytes(32)


def test_bvh_one_triangle_exact_t(orc):
    # bvhtracer/tests/test_bvh_one_triangle.rs:81-139 (same answers through the BVH)
    tri = unit_triangle().reshape(1, 9)
    nodes, used, perm = orc.bvh_build(tri)
    o = (0, 0, 5)
    for target, t in [((0, 0, 0), f32(5)), (tri[0, 0:3], np.sqrt(f32(101) / f32(4))),
                      (tri[0, 3:6], np.sqrt(f32(307) / f32(12))), (tri[0, 6:9], np.sqrt(f32(307) / f32(12)))]:
        h = orc.bvh_trace(nodes, perm, ray_towards(orc, o, target))[0]
        assert h["prim"] == 0 and h["t"] == t


def test_bvh_two_triangles(orc):
    # bvhtracer/tests/test_bvh_two_triangles.rs:16-82
    base = unit_triangle().reshape(3, 3)
    disp = np.array([5, 0, 0], np.float32)
    mesh = np.stack([(base - disp).reshape(9), (base + disp).reshape(9)])
    nodes, used, perm = orc.bvh_build(mesh)
    o = (0, 0, 5)
    for k in range(2):
        c = orc.centroid(perm[k])
        assert orc.bvh_trace(nodes, perm, ray_towards(orc, o, c))[0]["prim"] == k
    assert orc.bvh_trace(nodes, perm, ray_towards(orc, o, (0, 0, 0)))[0]["prim"] == 0xFFFFFFFF


def test_bvh_three_triangles(orc):
    # bvhtracer/tests/test_bvh_three_triangles.rs:16-60
    base = unit_triangle().reshape(3, 3)
    disp = np.array([0, 10, 0], np.float32)
    mesh = np.stack([base.reshape(9), (base - disp).reshape(9), (base + disp).reshape(9)])
    nodes, used, perm = orc.bvh_build(mesh)
    for k in range(3):
        r = ray_towards(orc, (0, 0, 10), orc.centroid(perm[k]))
        assert orc.triangle_intersect(perm[k], r) is not None
        assert orc.bvh_trace(nodes, perm, r)[0]["prim"] == k


def stacked_mesh():
    # bvhtracer/tests/test_bvh_closest_intersection.rs:16-32: 100 copies at z = 0, -1, ..., -99
    base = unit_triangle().reshape(3, 3)
    return np.stack([(base - np.array([0, 0, i], np.float32)).reshape(9) for i in range(100)])


def test_bvh_closest_intersection(orc):
    # :49-77: the BVH returns the closest triangle's t
    mesh = stacked_mesh()
    nodes, used, perm = orc.bvh_build(mesh)
    top = unit_triangle()
    r = ray_towards(orc, (0, 0, 5), orc.centroid(top))
    h = orc.bvh_trace(nodes, perm, r)[0]
    assert h["t"] == orc.triangle_intersect(top, r)[0]
    # :83-118: every stacked triangle is hit by the axial ray; the fold picks the top one
    c = orc.centroid(top)
    r2 = ray_towards(orc, (c[0], c[1], 5), c)
    ts = [orc.triangle_intersect(perm[k], r2) for k in range(100)]
    assert all(t is not None for t in ts)
    assert min(t[0] for t in ts) == orc.triangle_intersect(top, r2)[0]


# ---- bvhtracer/src/model/bvh.rs:545-643 structural invariants -------------

def diagonal_mesh():
    base = unit_triangle().reshape(3, 3)
    dx = np.array([5, 0, 0], np.float32)
    dy = np.array([0, 5, 0], np.float32)
    return np.stack([(base + f32(i) * dx + f32(i) * dy).reshape(9) for i in range(-100, 100)])


def check_structure(nodes, used, n):
    assert nodes.shape[0] == 2 * n and used < nodes.shape[0]
    zero = bytes(32)
    for k in range(used):
        if k == 1:
            continue
        assert nodes[k].tobytes() != zero
        if nodes[k]["count"] == 0:
            lf = nodes[k]["left_first"]
            assert lf > k and lf + 1 > k and lf < lf + 1 and lf % 2 == 0
    assert nodes[1].tobytes() == zero
    assert all(nodes[k].tobytes() == zero for k in range(used, nodes.shape[0]))
    leaves = nodes[:used][nodes[:used]["count"] > 0]
    assert leaves["count"].sum() == n


def test_bvh_structure_invariants(orc):
    mesh = diagonal_mesh()
    nodes, used, perm = orc.bvh_build(mesh)
    check_structure(nodes, used, mesh.shape[0])
    # refit leaves node 1 untouched and is the identity on an unchanged mesh
    # (bvh.rs:635-643; bvhtracer/tests/test_bvh_refit.rs:46-53)
    before = nodes.copy()
    orc.bvh_refit(nodes, used, perm)
    assert nodes.tobytes() == before.tobytes()


def test_bvh_refit_idempotent_after_animate(orc):
    # bvhtracer/tests/test_bvh_refit.rs:55-65 with animate() of :34-44
    mesh = diagonal_mesh()
    nodes, used, perm = orc.bvh_build(mesh)
    anim = perm.reshape(-1, 3, 3).copy()
    anim[:, 0, :] += np.array([0.3, 0, 0], np.float32)
    anim[:, 1, :] += np.array([0, 0.3, 0], np.float32)
    anim[:, 2, :] += np.array([0, 0, 0.3], np.float32)
    anim = anim.reshape(-1, 9)
    orc.bvh_refit(nodes, used, anim)
    once = nodes.copy()
    orc.bvh_refit(nodes, used, anim)
    assert nodes.tobytes() == once.tobytes()
    # refit boxes contain their triangles exactly (closed form)
    for k in range(used):
        if k != 1 and nodes[k]["count"] > 0:
            lf, c = nodes[k]["left_first"], nodes[k]["count"]
            v = anim[lf:lf + c].reshape(-1, 3)
            assert np.array_equal(nodes[k]["bmin"], v.min(0)) and np.array_equal(nodes[k]["bmax"], v.max(0))


# ---- cube.obj: ids 6 and 9 (bvhtracer/tests/test_scene_cube.rs) ------------

def cube_fixture():
    with open(os.path.join(HERE, "golden", "cube_obj.json")) as f:
        g = json.load(f)
    v = np.array(g["vertices"], np.float32)
    faces = np.array(g["faces"], np.int64) - 1
    return v[faces].reshape(-1, 9).astype(np.float32), g


def test_cube_build_matches_hand_trace(orc):
    """The faithful build of cube.obj: final order [T3,T2,T9,T8,T1,T0,T4,T10,
    T11,T5,T7,T6] and node layout traced by hand from bvh.rs:333-541."""
    mesh, g = cube_fixture()
    nodes, used, perm = orc.bvh_build(mesh)
    order = [int(np.nonzero((mesh == row).all(axis=1))[0][0]) for row in perm]
    assert order == g["expected_order"]
    assert used == g["expected_nodes_used"]
    got = [[int(nodes[k]["left_first"]), int(nodes[k]["count"])] for k in range(used)]
    assert got == g["expected_left_count"]


def cube_scene(orc):
    mesh, g = cube_fixture()
    nodes, used, perm = orc.bvh_build(mesh)
    # Transform3::new(scale 2, translation -1, identity) (test_scene_cube.rs:53-65)
    m = np.eye(4, dtype=np.float32)
    m[:3, :3] *= 2
    m[3, :3] = -1  # m[col][row]: translation in column 3
    minv = np.eye(4, dtype=np.float32)
    minv[:3, :3] *= 0.5
    minv[3, :3] = 0.5
    bmin, bmax = orc.instance_bounds(m.reshape(16), nodes[0]["bmin"], nodes[0]["bmax"])
    tnodes, tused = orc.tlas_build(np.concatenate([bmin, bmax]).reshape(1, 6))
    inst = np.zeros(1, orc.INSTANCE_DTYPE)
    inst["minv"][0] = minv.reshape(16)
    inst["blas"][0] = 0
    return tnodes, inst, [(nodes, perm)]


@pytest.mark.parametrize("target,prim", [((0.5, 1.0, -0.5), 6), ((-0.5, 1.0, 0.5), 9)])
def test_cube_scene_ids_and_t(orc, target, prim):
    # :85-135 and :176-226: hit, t ~ sqrt(19/2) (epsilon 1e-7), instance 0, primitive id
    tnodes, inst, blases = cube_scene(orc)
    r = ray_towards(orc, (0, 4, 0), target)
    h = orc.tlas_trace(tnodes, inst, blases, r)[0]
    assert h["prim"] == prim
    assert (h["packed"] >> 20) == 0 and (h["packed"] & 0xFFFFF) == prim
    expected = np.sqrt(f32(19) / f32(2))
    diff = abs(float(h["t"]) - float(expected))
    assert diff <= 1e-7 or diff <= np.finfo(np.float32).eps * max(abs(h["t"]), abs(expected))


# ---- bvhtracer/tests/test_scene_quad.rs -----------------------------------

def quad_scene(orc):
    mesh = np.array([[-1, -1, 0, 1, 1, 0, -1, 1, 0], [-1, -1, 0, 1, -1, 0, 1, 1, 0]], np.float32)
    nodes, used, perm = orc.bvh_build(mesh)
    eye = np.eye(4, dtype=np.float32).reshape(16)
    bmin, bmax = orc.instance_bounds(eye, nodes[0]["bmin"], nodes[0]["bmax"])
    tnodes, _ = orc.tlas_build(np.concatenate([bmin, bmax]).reshape(1, 6))
    inst = np.zeros(1, orc.INSTANCE_DTYPE)
    inst["minv"][0] = eye
    # BoxSpec(-1,1,-1,1,1,100) -> corners (-1,1,-1),(1,1,-1),(-1,-1,-1); camera at (0,0,2)
    # looking down -z with right +x, up +y: view_inv = translation by (0,0,2)
    cam = np.zeros(1, orc.CAMERA_DTYPE)
    cam["tl"][0] = (-1, 1, -1)
    cam["tr"][0] = (1, 1, -1)
    cam["bl"][0] = (-1, -1, -1)
    vi = np.eye(4, dtype=np.float32)
    vi[3, :3] = (0, 0, 2)
    cam["view_inv"][0] = vi.reshape(16)
    return tnodes, inst, [(nodes, perm)], cam


def quad_test_case(orc):
    """test_case() of test_scene_quad.rs:133-198 in binary32."""
    o = np.array([0, 0, 2], np.float32)
    corners = {k: np.array(v, np.float32) for k, v in
               {"tl": (-1, 1, 0), "tr": (1, 1, 0), "bl": (-1, -1, 0), "br": (1, -1, 0)}.items()}
    tc = np.sqrt(f32(6)) / f32(2)
    cp = {k: o + orc.normalize(v - o) * tc for k, v in corners.items()}
    dim = np.array([(cp["tr"][0] - cp["tl"][0]) / f32(2), (cp["tl"][1] - cp["bl"][1]) / f32(2)], np.float32)
    tl = np.array([(cp["tl"][0] - corners["tl"][0]) / f32(2), (corners["tl"][1] - cp["tl"][1]) / f32(2)], np.float32)
    uv = {"tl": tl, "tr": np.array([tl[0] + dim[0], tl[1]], np.float32),
          "bl": np.array([tl[0], tl[1] + dim[1]], np.float32), "br": tl + dim}
    return uv, tc


def test_quad_depths_exact(orc):
    # :200-261: t == 2 at the centre and sqrt(6) at the four corners, exactly
    tnodes, inst, blases, cam = quad_scene(orc)
    uv, tc = quad_test_case(orc)
    r = orc.camera_ray(cam, f32(0.5), f32(0.5))
    assert orc.tlas_trace(tnodes, inst, blases, r)[0]["t"] == f32(2)
    for k in ("tl", "tr", "bl", "br"):
        r = orc.camera_ray(cam, uv[k][0], uv[k][1])
        assert orc.tlas_trace(tnodes, inst, blases, r)[0]["t"] == np.sqrt(f32(6)), k
        # :263-324: the camera-plane point of each corner ray is at z == 1
        z = r["o"][0][2] + r["d"][0][2] * tc
        assert z == f32(1)


def test_quad_viewport_hit_map(orc):
    """:354-378: hit iff (u, v) lies in the quad's solid angle, over 640x640.

    Exact for every pixel off the quad's four edge lines.  The 1280 pixels ON
    the edge lines (u or v exactly 0.25 / 0.75) land exactly on a triangle edge
    (u + v = 1 in exact arithmetic), so hit/miss is decided by the last bit of
    cglinalg's normalize/dot rounding, which is not pinned (unpinned git
    dependency, bvhtracer/Cargo.toml:20-22): this restatement misses 16 of them
    (top and right edges); every variant of the operation order tried misses
    12-16, so the reference's own assertion may fail as shipped.  Parity on
    those 16 rays is unpinned; the GPU path is checked against the oracle on
    them bit for bit like every other ray.
    """
    tnodes, inst, blases, cam = quad_scene(orc)
    uv, _ = quad_test_case(orc)
    rays = orc.raygen(cam, 640, 640)
    hits = orc.tlas_trace(tnodes, inst, blases, rays, nthreads=8)
    x = np.arange(640, dtype=np.float32)
    u = (x / f32(640))[None, :].repeat(640, 0)
    v = (x / f32(640))[:, None].repeat(640, 1)
    expected = (u >= uv["tl"][0]) & (u <= uv["tr"][0]) & (v >= uv["tl"][1]) & (v <= uv["bl"][1])
    got = (hits["prim"] != 0xFFFFFFFF).reshape(640, 640)
    on_edge = ((u == uv["tl"][0]) | (u == uv["tr"][0])) & (v >= uv["tl"][1]) & (v <= uv["bl"][1])
    on_edge |= ((v == uv["tl"][1]) | (v == uv["bl"][1])) & (u >= uv["tl"][0]) & (u <= uv["tr"][0])
    assert np.array_equal(got[~on_edge], expected[~on_edge])
    assert on_edge.sum() == 1280
    assert (got[on_edge] != expected[on_edge]).sum() == 16
