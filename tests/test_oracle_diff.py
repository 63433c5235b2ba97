"""Pins of the oracle's differentiable depth/shading checker (no reference
counterpart): the binary32 forward against the reference-path TLAS trace and
a float64 shading formula, the binary64 gradient against closed forms, an
invariant and Richardson-extrapolated central finite differences."""
import numpy as np
import pytest

from bvhtracer_amd import scenes

MISS = 0xFFFFFFFF


def face_order(orc, soup):
    """Oracle SAH build of a face soup; returns (nodes, build-order tris,
    face_of) with face_of recovered by matching the (distinct) triangles."""
    nodes, _, tris = orc.bvh_build(soup.copy(), faithful=False, max_leaf=2)
    key = {r.tobytes(): i for i, r in enumerate(soup)}
    assert len(key) == soup.shape[0]
    return nodes, tris, np.array([key[r.tobytes()] for r in tris], np.uint32)


@pytest.fixture(scope="module")
def small(orc):
    w = scenes.diff_workload(8, width=64, n_cameras=2)
    V, F = w["V"], w["F"]
    rays = np.concatenate([orc.raygen(w["cameras"][k:k + 1], 64, 64) for k in range(2)])
    nodes_t, tris_t, fo_t = face_order(orc, np.ascontiguousarray(w["V_target"][F].reshape(-1, 9)))
    Nt = orc.mesh_normals(w["V_target"], F)
    _, img_t = orc.diff_forward(nodes_t, tris_t, F[fo_t].reshape(-1), Nt, w["Minv"], w["light"], rays)
    nodes, tris, fo = face_order(orc, np.ascontiguousarray(V[F].reshape(-1, 9)))
    N = orc.mesh_normals(V, F)
    vid = F[fo].reshape(-1)
    hits, img = orc.diff_forward(nodes, tris, vid, N, w["Minv"], w["light"], rays)
    return dict(w=w, rays=rays, nodes=nodes, tris=tris, vid=vid, N=N, hits=hits, img=img, target=img_t)


def test_mesh_normals_are_area_weighted_unit_normals(orc):
    V, F = scenes.indexed(scenes.icosphere(6, radius=32.0, amp=0.15))
    N = orc.mesh_normals(V, F)
    ref = scenes.vertex_normals(V, F)
    assert np.abs(N - ref).max() < 2e-6
    assert np.abs(np.linalg.norm(N, axis=1) - 1.0).max() < 1e-6
    assert ((N * V).sum(1) > 0).all()  # outward on a star-shaped surface


def test_forward_hits_equal_one_object_tlas_trace(orc, small):
    """The forward's transform + BLAS traversal is SceneObject::intersect:
    its hits equal the oracle TLAS trace of a one-object scene bit for bit."""
    w = small["w"]
    nodes, used, tris = small["nodes"], None, small["tris"]
    bmin = np.array([nodes[0]["bmin"]], np.float32)[0]
    bmax = np.array([nodes[0]["bmax"]], np.float32)[0]
    wb = np.concatenate(orc.instance_bounds(w["M"], bmin, bmax)).astype(np.float32)
    tnodes, _ = orc.tlas_build(wb.reshape(1, 6))
    inst = np.zeros(1, orc.INSTANCE_DTYPE)
    inst["minv"][0] = w["Minv"]
    ref = orc.tlas_trace(tnodes, inst, [(nodes, tris)], small["rays"])
    got = small["hits"]
    assert got.tobytes() == ref.tobytes()
    assert (got["prim"] != MISS).sum() > 3000


def test_forward_image_matches_float64_shading(small):
    hits, img, N, vid = small["hits"], small["img"], small["N"], small["vid"]
    light = small["w"]["light"].astype(np.float64)
    hit = hits["prim"] != MISS
    assert (img[~hit, 0] == np.float32(np.finfo(np.float32).max)).all() and (img[~hit, 1] == 0).all()
    assert (img[hit, 0] == hits["t"][hit]).all()
    p = hits["prim"][hit]
    u, v = hits["u"][hit].astype(np.float64), hits["v"][hit].astype(np.float64)
    n = [N[vid.reshape(-1, 3)[p, k]].astype(np.float64) for k in range(3)]
    m = (1 - u - v)[:, None] * n[0] + u[:, None] * n[1] + v[:, None] * n[2]
    s = np.maximum(0.0, (m @ light) / np.linalg.norm(m, axis=1))
    assert np.abs(img[hit, 1] - s).max() < 2e-6
    assert (s > 0.1).sum() > 1000


def test_grad64_plane_closed_form(orc):
    """Triangle in the plane z = c, rays along +z, identity M^-1.  Depth loss:
    dL/dv_k = g_t b_k e_z (barycentric weights b), dL/dM^-1[2][c] = -g_t p_c
    (hit point p), dL/dM^-1[2][3] = -g_t, all else 0.  Shading with equal
    vertex normals contributes nothing."""
    c = 5.0
    V = np.array([[-4, -4, c], [6, -4, c], [-4, 6, c]], np.float64)
    vid = np.array([0, 1, 2], np.uint32)
    N = np.tile([0.0, 0.0, -1.0], (3, 1))
    rays = np.zeros(3, orc.RAY_DTYPE)
    rays["o"] = [[0.5, 0.25, -1.0], [-2.0, 1.0, 0.0], [1.0, -3.0, 2.0]]
    rays["d"] = [[0, 0, 1]] * 3
    rays["tmax"] = np.finfo(np.float32).max
    target = np.array([[4.0, 0.3], [6.5, 0.9], [2.5, 0.0]], np.float32)
    minv = np.eye(4).reshape(16)
    light = np.array([0.0, 0.0, -1.0])
    loss, gv, gm, img = orc.diff_grad64(V, vid, N, minv, light, rays, np.zeros(3, np.uint32), target, 1.0, 0.0)
    t = c - rays["o"][:, 2].astype(np.float64)
    gt = 2.0 * (t - target[:, 0])
    assert np.allclose(img[:, 0], t, rtol=0, atol=1e-12)
    assert loss == pytest.approx(((t - target[:, 0]) ** 2).sum(), rel=1e-14)
    p = rays["o"].astype(np.float64) + t[:, None] * rays["d"]
    # barycentrics of the hit points
    b1 = (p[:, 0] + 4) / 10.0
    b2 = (p[:, 1] + 4) / 10.0
    b0 = 1 - b1 - b2
    want = np.zeros((3, 3))
    for k, b in enumerate((b0, b1, b2)):
        want[k, 2] = (gt * b).sum()
    assert np.abs(gv - want).max() < 1e-12
    G = gm.reshape(4, 4).T  # [row, col]
    wantG = np.zeros((4, 4))
    wantG[2, :3] = -(gt[:, None] * p).sum(0)
    wantG[2, 3] = -gt.sum()
    assert np.abs(G - wantG).max() < 1e-11
    # shading only, equal normals: no gradient
    loss_s, gv_s, gm_s, _ = orc.diff_grad64(V, vid, N, minv, light, rays, np.zeros(3, np.uint32), target, 0.0, 1.0)
    assert loss_s > 0 and np.abs(gv_s).max() == 0 and np.abs(gm_s).max() == 0


def test_grad64_translation_invariant(small, orc):
    """Moving every vertex by delta equals moving o' by -delta, i.e. the
    translation column of M^-1 by -delta: sum_v dL/dV = -dL/dM^-1[0..2][3]."""
    w = small["w"]
    loss, gv, gm, _ = orc.diff_grad64(w["V"], small["vid"], small["N"], w["Minv"], w["light"], small["rays"],
                                      small["hits"]["prim"], small["target"])
    assert loss > 0
    col3 = gm.reshape(4, 4)[3, :3]
    assert np.abs(gv.sum(0) + col3).max() < 1e-9 * np.abs(gv).sum()


def richardson(f, x0, h):
    d1 = (f(x0 + h) - f(x0 - h)) / (2 * h)
    d2 = (f(x0 + h / 2) - f(x0 - h / 2)) / h
    return d1, (4 * d2 - d1) / 3


@pytest.mark.parametrize("w_depth,w_shade", [(1.0, 1.0), (0.0, 1.0)])
def test_grad64_central_differences(small, orc, w_depth, w_shade):
    """Central differences with h = 1e-3 (and their Richardson extrapolation
    with h/2) of the binary64 loss at fixed hit triangles: the closed-form
    gradient matches for the 12 largest vertex coordinates and all 12 free
    entries of M^-1.  The shading-only loss pins the shading term, which the
    depth term otherwise dominates."""
    w = small["w"]
    V = w["V"].astype(np.float64)
    minv = w["Minv"].astype(np.float64)
    args = (small["vid"], small["N"])
    rest = (w["light"], small["rays"], small["hits"]["prim"], small["target"], w_depth, w_shade)

    def loss_v(k):
        def f(x):
            Vx = V.copy()
            Vx.ravel()[k] = x
            return orc.diff_grad64(Vx, *args, minv, *rest)[0]
        return f

    def loss_m(c):
        def f(x):
            m = minv.copy()
            m[c] = x
            return orc.diff_grad64(V, *args, m, *rest)[0]
        return f

    _, gv, gm, _ = orc.diff_grad64(V, *args, minv, *rest)
    gvf = gv.ravel()
    scale_v = np.abs(gvf).max()
    assert scale_v > 0.1
    for k in np.argsort(-np.abs(gvf))[:12]:
        d1, dr = richardson(loss_v(k), V.ravel()[k], 1e-3)
        assert abs(d1 - gvf[k]) < 5e-4 * scale_v, (k, d1, gvf[k])
        assert abs(dr - gvf[k]) < 1e-6 * scale_v, (k, dr, gvf[k])
    scale_m = np.abs(gm).max()
    for c in range(16):
        if c % 4 == 3:
            assert gm[c] == 0.0  # bottom row of M^-1 is not a free parameter
            continue
        d1, dr = richardson(loss_m(c), minv[c], 1e-3)
        assert abs(d1 - gm[c]) < 3e-3 * scale_m, (c, d1, gm[c])
        # pixels crossing the shading terminator (the clamp's kink) under the
        # perturbation limit the extrapolation to ~2e-5 here
        assert abs(dr - gm[c]) < 1e-4 * scale_m, (c, dr, gm[c])


def test_grad64_shading_term_is_not_negligible(small, orc):
    """The shading-only gradient is substantial and differs in direction from
    the depth-only one, so the shading-only finite-difference pin has teeth."""
    w = small["w"]
    rest = (w["light"], small["rays"], small["hits"]["prim"], small["target"])
    _, g_s, _, _ = orc.diff_grad64(w["V"], small["vid"], small["N"], w["Minv"], *rest, 0.0, 1.0)
    _, g_d, _, _ = orc.diff_grad64(w["V"], small["vid"], small["N"], w["Minv"], *rest, 1.0, 0.0)
    cos = (g_s * g_d).sum() / np.linalg.norm(g_s) / np.linalg.norm(g_d)
    assert np.abs(g_s).max() > 0.1 and abs(cos) < 0.9
