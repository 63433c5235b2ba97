"""The oracle's Tlas::rebuild and Tlas::intersect against scenes traced by
hand from bvhtracer/src/scene/tlas.rs:123-250 (tests/golden/
tlas_hand_trace.json; the derivation is written out in each case).  This
pins the merge walk (a = b; b = c), the replacement list[b] = list[cnt-1],
the first-index tie rule, an `a` left beyond the list, LeftRightIndex's
swapped decoding (left() = second argument of new()), the right-first
descent on equal distances and the root copy."""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "tlas_hand_trace.json")


def load():
    with open(FIXTURE) as f:
        return json.load(f)


def case_bounds(case):
    x = np.asarray(case["offsets_x"], np.float32)
    b = np.zeros((x.size, 6), np.float32)
    b[:, 0] = x
    b[:, 3] = x + 1
    b[:, 4:6] = 1
    return b


def expected_nodes(orc, case):
    nodes = np.zeros(len(case["nodes"]), orc.TNODE_DTYPE)
    for k, n in enumerate(case["nodes"]):
        nodes[k]["bmin"] = n["bmin"]
        nodes[k]["bmax"] = n["bmax"]
        nodes[k]["left_right"] = n["left_right"]
        nodes[k]["blas"] = n["blas"]
    return nodes


def translations(case):
    x = np.asarray(case["offsets_x"], np.float32)
    M = np.tile(np.eye(4, dtype=np.float32).reshape(16), (x.size, 1))
    Minv = M.copy()
    M[:, 12] = x
    Minv[:, 12] = -x
    return M, Minv


def hand_ray(orc, fx):
    rays = np.zeros(1, orc.RAY_DTYPE)
    rays["o"] = fx["ray"]["o"]
    rays["d"] = fx["ray"]["d"]
    rays["tmax"] = np.finfo(np.float32).max
    return rays


@pytest.mark.parametrize("name", [c["name"] for c in load()["cases"]])
def test_oracle_rebuild_equals_hand_trace(orc, name):
    fx = load()
    case = next(c for c in fx["cases"] if c["name"] == name)
    nodes, used = orc.tlas_build(case_bounds(case))
    assert used == case["nodes_used"]
    assert nodes.tobytes() == expected_nodes(orc, case).tobytes()
    # the same boxes through the object-bounds path (eight transformed
    # corners of the model box, scene_object.rs:60-75)
    M, _ = translations(case)
    tri = np.asarray(fx["model_triangle"], np.float32).reshape(1, 9)
    mb = tri.reshape(3, 3)
    b = np.stack([np.concatenate(orc.instance_bounds(M[k], mb.min(0), mb.max(0))) for k in range(M.shape[0])])
    assert b.tobytes() == case_bounds(case).tobytes()


@pytest.mark.parametrize("name", ["two_identical_objects", "three_identical_objects"])
def test_oracle_trace_takes_right_child_first_on_ties(orc, name):
    fx = load()
    case = next(c for c in fx["cases"] if c["name"] == name)
    tnodes = expected_nodes(orc, case)
    M, Minv = translations(case)
    tri = np.asarray(fx["model_triangle"], np.float32).reshape(1, 9)
    bnodes, bused, perm = orc.bvh_build(tri, True)
    assert bused == 1 or bnodes[0]["count"] == 1
    inst = np.zeros(M.shape[0], orc.INSTANCE_DTYPE)
    inst["minv"] = Minv
    hits, st = orc.tlas_trace(tnodes, inst, [(bnodes, perm)], hand_ray(orc, fx), stats=True)
    h = fx["hit"]
    assert hits[0]["t"] == np.float32(h["t"]) and hits[0]["u"] == np.float32(h["u"])
    assert hits[0]["v"] == np.float32(h["v"]) and hits[0]["prim"] == h["prim"]
    assert hits[0]["inst"] == case["hit_inst"]
    assert st[0]["tlas_pairs"] == case["stats"]["tlas_pairs"]
    assert st[0]["inst_visits"] == case["stats"]["inst_visits"]


def test_fixture_distinguishes_plausible_mistakes(orc):
    """Each plausible mistake in the merge or the decoding changes at least
    one hand-traced node array: a last-index tie rule, new(b, a) instead of
    new(a, b), a root that is the first node instead of the last merge, a
    merge with best(a) without waiting for a mutual match, and removal of
    entry b by shifting the list instead of moving the last entry into it.  The mistakes are planted in a pure-Python copy of the merge
    loop on the CPU; the unmodified copy must equal the fixture."""
    fx = load()

    def area(lo, hi):
        e = hi - lo
        return e[0] * e[1] + e[1] * e[2] + e[2] * e[0]

    def rebuild(bounds, tie_last=False, swap=False, root_first=False, greedy=False, shift=False):
        n = bounds.shape[0]
        lo = np.zeros((2 * n, 3), np.float32)
        hi = np.zeros((2 * n, 3), np.float32)
        lrw = np.zeros(2 * n, np.uint32)
        bl = np.zeros(2 * n, np.uint32)
        lst = list(range(1, n + 1))
        for i in range(n):
            lo[i + 1], hi[i + 1], bl[i + 1] = bounds[i, :3], bounds[i, 3:], i
        used, cnt = n + 1, n

        def best(a):
            s, bb = np.float32(np.finfo(np.float32).max), -1
            for b in range(cnt):
                if b != a:
                    ar = area(np.minimum(lo[lst[a]], lo[lst[b]]), np.maximum(hi[lst[a]], hi[lst[b]]))
                    if ar < s or (tie_last and ar == s):
                        s, bb = ar, b
            return bb

        a = 0
        b = best(a)
        while cnt > 1:
            c = best(b)
            if a == c or greedy:
                ia, ib = lst[a], lst[b]
                lrw[used] = (ib + (ia << 16)) if swap else (ia + (ib << 16))
                lo[used] = np.minimum(lo[ia], lo[ib])
                hi[used] = np.maximum(hi[ia], hi[ib])
                lst[a] = used
                used += 1
                if shift:
                    del lst[b]
                    lst.append(0)
                    a = a - 1 if b < a else a
                else:
                    lst[b] = lst[cnt - 1]
                cnt -= 1
                b = best(a)
            else:
                a, b = b, c
        r = 1 if root_first else lst[a]
        lo[0], hi[0], lrw[0], bl[0] = lo[r], hi[r], lrw[r], bl[r]
        out = np.zeros(2 * n, orc.TNODE_DTYPE)
        out["bmin"], out["bmax"], out["left_right"], out["blas"] = lo, hi, lrw, bl
        return out

    for case in fx["cases"]:
        assert rebuild(case_bounds(case)).tobytes() == expected_nodes(orc, case).tobytes(), case["name"]
    for kw in ({"tie_last": True}, {"swap": True}, {"root_first": True}, {"greedy": True}, {"shift": True}):
        changed = [c["name"] for c in fx["cases"]
                   if rebuild(case_bounds(c), **kw).tobytes() != expected_nodes(orc, c).tobytes()]
        assert changed, f"no hand-traced case detects {kw}"
