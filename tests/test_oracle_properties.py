"""Checks of the oracle that do not come from the reference's own tests:
brute force (a second algorithm with no BVH), closed forms and invariants.
Each is chosen so that a plausible mistake in the restatement (wrong child
order, a dropped push, a wrong index in the partition, a stale box) fails.
"""
import numpy as np
import pytest

from bvhtracer_amd import scenes

f32 = np.float32


def soup_rays(orc, w=96, h=96):
    return orc.raygen(scenes.soup_camera(), w, h)


@pytest.mark.parametrize("faithful,max_leaf", [(True, 0), (False, 2), (False, 0)])
def test_traversal_equals_brute_force(orc, faithful, max_leaf):
    """Nearest t equals brute force over all triangles for every ray; the
    primitive id too, except where several triangles share the minimal t."""
    mesh = scenes.triangle_soup(2000, seed=11)
    nodes, used, perm = orc.bvh_build(mesh, faithful, max_leaf)
    rays = soup_rays(orc, 64, 64)
    hits = orc.bvh_trace(nodes, perm, rays)
    ref, ties = orc.brute_trace(perm, rays)
    assert np.array_equal(hits["t"].view(np.uint32), ref["t"].view(np.uint32))
    single = ties <= 1
    assert np.array_equal(hits["prim"][single], ref["prim"][single])
    assert (hits["prim"] != 0xFFFFFFFF).sum() > 200  # the test actually hits things


def test_traversal_brute_force_random_rays(orc):
    """Rays from inside the soup in random directions (incoherent), finite tmax."""
    mesh = scenes.triangle_soup(1500, seed=5)
    nodes, used, perm = orc.bvh_build(mesh, False, 2)
    rng = scenes.Pcg32(99, np.arange(3000, dtype=np.uint64))
    o = np.stack([rng.next_f32() for _ in range(3)], 1) * f32(10) - f32(5)
    d = np.stack([rng.next_f32() for _ in range(3)], 1) * f32(2) - f32(1)
    rays = np.zeros(3000, orc.RAY_DTYPE)
    rays["o"] = o
    rays["d"] = d
    rays["tmax"] = np.where(np.arange(3000) % 3 == 0, f32(2.5), np.finfo(np.float32).max)
    hits = orc.bvh_trace(nodes, perm, rays)
    ref, ties = orc.brute_trace(perm, rays)
    assert np.array_equal(hits["t"].view(np.uint32), ref["t"].view(np.uint32))
    assert np.array_equal(hits["prim"][ties <= 1], ref["prim"][ties <= 1])


def test_visit_counts_are_consistent(orc):
    """Every expanded interior node fetches one child pair; every leaf visit
    tests all its primitives: the counters bound the brute force work."""
    mesh = scenes.triangle_soup(3000, seed=3)
    nodes, used, perm = orc.bvh_build(mesh, False, 2)
    rays = soup_rays(orc, 32, 32)
    _, st = orc.bvh_trace(nodes, perm, rays, stats=True)
    internal = (nodes[:used]["count"] == 0).sum() - 1  # minus the dummy node 1
    assert (st["pairs"] >= 1).all()  # the root is always expanded (never culled)
    assert (st["pairs"] <= internal).all() and (st["tests"] <= mesh.shape[0]).all()


def two_pointer(flags):
    """The reference's in-place partition (bvh.rs:420-429) on a flag array."""
    a = list(range(len(flags)))
    i, j = 0, len(a) - 1
    while i <= j:
        if flags[a[i]]:
            i += 1
        else:
            a[i], a[j] = a[j], a[i]
            j -= 1
    return a, i


def closed_form(flags):
    """The permutation the GPU partition kernel computes from prefix counts."""
    n = len(flags)
    nl = int(sum(flags))
    a = n if nl == n else (nl if flags[nl] else nl + 1)
    lb = np.concatenate([[0], np.cumsum(flags)]).astype(int)
    kpos, sel = {}, {}
    for k in range(n):
        if k < a and not flags[k]:
            kpos[k - lb[k]] = k
        elif k >= a and flags[k]:
            sel[nl - lb[k] - 1] = n - 1 - k
    out = [None] * n
    for k in range(n):
        if k < a:
            dest = k if flags[k] else (n - 1 if k - lb[k] == 0 else n - 2 - sel[k - lb[k] - 1])
        else:
            dest = kpos[nl - lb[k] - 1] if flags[k] else k - 1
        assert out[dest] is None
        out[dest] = k
    return out, nl


def test_partition_closed_form_matches_two_pointer():
    """The closed-form permutation equals the sequential two-pointer loop on
    every flag pattern up to n = 10 and on random longer ones."""
    for n in range(1, 11):
        for bits in range(1 << n):
            flags = [(bits >> k) & 1 == 1 for k in range(n)]
            assert closed_form(flags) == two_pointer(flags), flags
    rng = np.random.default_rng(7)
    for _ in range(300):
        n = int(rng.integers(11, 400))
        flags = list(rng.random(n) < rng.random())
        assert closed_form(flags) == two_pointer(flags)


def test_worked_partition_example():
    # [R0, L1, R2, R3, L4] -> [L4, L1, R3, R2, R0]
    order, nl = two_pointer([False, True, False, False, True])
    assert order == [4, 1, 3, 2, 0] and nl == 2


def test_refit_equals_fresh_bounds(orc):
    """After deforming the vertices, refit boxes equal the exact min/max of
    the subtree's triangles (closed form), node by node."""
    mesh = scenes.icosphere(6, radius=10.0)
    nodes, used, perm = orc.bvh_build(mesh, False, 2)
    deformed = scenes.deform_wave(perm, frame=7)
    orc.bvh_refit(nodes, used, deformed)
    v = deformed.reshape(-1, 3, 3)

    def subtree_box(k):
        if nodes[k]["count"] > 0:
            lf, c = nodes[k]["left_first"], nodes[k]["count"]
            w = v[lf:lf + c].reshape(-1, 3)
            return w.min(0), w.max(0)
        a = subtree_box(nodes[k]["left_first"])
        b = subtree_box(nodes[k]["left_first"] + 1)
        return np.minimum(a[0], b[0]), np.maximum(a[1], b[1])

    for k in list(range(0, used, 7)) + [0]:
        if k == 1:
            continue
        lo, hi = subtree_box(k)
        assert np.array_equal(nodes[k]["bmin"], lo) and np.array_equal(nodes[k]["bmax"], hi)


def test_sah_mode_leaf_cap_and_ranges(orc):
    mesh = scenes.triangle_soup(5000, seed=2)
    nodes, used, perm = orc.bvh_build(mesh, False, 2)
    leaves = nodes[:used][nodes[:used]["count"] > 0]
    assert leaves["count"].max() <= 2
    # leaf ranges tile [0, N) exactly
    order = np.argsort(leaves["left_first"])
    lf, c = leaves["left_first"][order], leaves["count"][order]
    assert lf[0] == 0 and np.array_equal(lf[1:], (lf + c)[:-1]) and (lf + c)[-1] == 5000
    # the permutation is a permutation of the input rows
    assert np.array_equal(np.sort(perm.view("V36").ravel()), np.sort(mesh.view("V36").ravel()))


def test_transform_point_vector_closed_form(orc):
    m = np.arange(16, dtype=np.float32) * f32(0.5)  # column-major
    p = np.array([1.5, -2.0, 3.0], np.float32)
    M = m.reshape(4, 4).T.astype(np.float64)  # row-major
    exp_p = (M @ np.array([1.5, -2.0, 3.0, 1.0]))[:3]
    exp_v = (M @ np.array([1.5, -2.0, 3.0, 0.0]))[:3]
    assert np.array_equal(orc.transform_point(m, p), exp_p.astype(np.float32))
    assert np.array_equal(orc.transform_vector(m, p), exp_v.astype(np.float32))


def test_instance_bounds_contain_transformed_corners(orc):
    M, Minv = scenes.random_affines(4, seed=3)
    lo, hi = np.array([-1, -2, -3], np.float32), np.array([4, 5, 6], np.float32)
    for k in range(4):
        bmin, bmax = orc.instance_bounds(M[k], lo, hi)
        corners = np.array([[(hi if c & 1 else lo)[0], (hi if c & 2 else lo)[1], (hi if c & 4 else lo)[2]]
                            for c in range(8)], np.float32)
        w = np.stack([orc.transform_point(M[k], c) for c in corners])
        assert np.array_equal(bmin, w.min(0)) and np.array_equal(bmax, w.max(0))


def brute_tlas(orc, objs_minv, blases, obj_blas, rays):
    """Nearest hit over every object's triangles (no TLAS, no BVH)."""
    best_t = np.full(rays.shape[0], np.finfo(np.float32).max, np.float32)
    best_obj = np.full(rays.shape[0], -1)
    for k, minv in enumerate(objs_minv):
        mr = rays.copy()
        for i in range(rays.shape[0]):
            mr["o"][i] = orc.transform_point(minv, rays["o"][i])
            mr["d"][i] = orc.transform_vector(minv, rays["d"][i])
        h, _ = orc.brute_trace(blases[obj_blas[k]][1], mr)
        better = (h["prim"] != 0xFFFFFFFF) & (h["t"] < best_t)
        best_t[better] = h["t"][better]
        best_obj[better] = k
    return best_t, best_obj


def test_tlas_equals_brute_force_over_instances(orc):
    base = scenes.icosphere(4, radius=3.0, amp=0.1)
    nodes, used, perm = orc.bvh_build(base, False, 2)
    M, Minv = scenes.random_affines(6, seed=21, trans_range=8.0)
    bounds = np.stack([np.concatenate(orc.instance_bounds(M[k], nodes[0]["bmin"], nodes[0]["bmax"]))
                       for k in range(6)])
    tnodes, tused = orc.tlas_build(bounds)
    assert tused == 12
    inst = np.zeros(6, orc.INSTANCE_DTYPE)
    inst["minv"] = Minv
    cam = scenes.camera(eye=(0, 0, -40), forward=(0, 0, 1), right=(1, 0, 0), up=(0, 1, 0), fovy_deg=50)
    rays = orc.raygen(cam, 40, 40)
    hits = orc.tlas_trace(tnodes, inst, [(nodes, perm)], rays)
    bt, bo = brute_tlas(orc, Minv, [(nodes, perm)], [0] * 6, rays)
    found = hits["prim"] != 0xFFFFFFFF
    assert np.array_equal(found, bo >= 0)
    assert np.array_equal(hits["t"][found], bt[found])
    assert np.array_equal(hits["inst"][found].astype(int), bo[found])
    assert found.sum() > 50


def test_tlas_structure(orc):
    rng = np.random.default_rng(4)
    for n in (1, 2, 3, 7, 16, 33):
        lo = rng.uniform(-10, 10, (n, 3)).astype(np.float32)
        b = np.concatenate([lo, lo + rng.uniform(0.1, 3, (n, 3)).astype(np.float32)], 1)
        t, used = orc.tlas_build(b)
        assert used == 2 * n
        # every object appears in exactly one leaf reachable from the root;
        # interior boxes are the union of their children (tlas.rs:231-235)
        seen = []
        stack = [0]
        while stack:
            k = stack.pop()
            if t[k]["left_right"] == 0:
                seen.append(int(t[k]["blas"]))
                assert np.array_equal(t[k]["bmin"], b[t[k]["blas"], :3])
                continue
            l, r = t[k]["left_right"] >> 16, t[k]["left_right"] & 0xFFFF
            assert np.array_equal(t[k]["bmin"], np.minimum(t[l]["bmin"], t[r]["bmin"]))
            assert np.array_equal(t[k]["bmax"], np.maximum(t[l]["bmax"], t[r]["bmax"]))
            stack += [l, r]
        assert sorted(seen) == list(range(n))
