"""Shared helpers for the GPU parity tests."""
import numpy as np

import bvhtracer_amd as B


def hits_bits(h):
    """Hit records as raw 32-bit words (t, u, v compared bitwise)."""
    return np.ascontiguousarray(h).view(np.uint32).reshape(-1, 6)


def assert_hits_equal(gpu, ref, what=""):
    a, b = hits_bits(gpu), hits_bits(ref)
    if not np.array_equal(a, b):
        bad = np.nonzero((a != b).any(axis=1))[0]
        i = bad[0]
        raise AssertionError(f"{what}: {bad.size} of {a.shape[0]} hits differ; first at {i}: "
                             f"gpu={gpu[i]} oracle={ref[i]}")


def assert_nodes_equal(gpu, ref, used):
    """Node arrays: indices and counts exact, boxes equal as floats (+0 == -0),
    unused tail identical."""
    assert np.array_equal(gpu["left_first"][:used], ref["left_first"][:used])
    assert np.array_equal(gpu["count"][:used], ref["count"][:used])
    assert np.array_equal(gpu["bmin"][:used], ref["bmin"][:used])
    assert np.array_equal(gpu["bmax"][:used], ref["bmax"][:used])
    assert gpu[used:].tobytes() == ref[used:].tobytes()


def trace_gpu(ctx, bvh, rays, stats=False):
    return bvh.intersect(rays, stats=stats)


def upload(ctx, nodes, used, perm):
    return B.Bvh.from_arrays(nodes, used, perm, ctx=ctx)


def oracle_rays_to_product(rays):
    out = np.zeros(rays.shape[0], B.RAY_DTYPE)
    out.view(np.uint8)[:] = rays.view(np.uint8)
    return out


def comb_tree(n=150):
    """A valid hand-made BVH of depth n over n triangles in planes x = 1..n:
    node 0 and every left child are internal (triangles [0, k)), every right
    child is the leaf holding triangle k-1.  A ray along +x from x < 1 meets
    the left (deeper, nearer) child first at every level and pushes the right
    one, so the traversal stack grows to n - 1 entries (beyond any LDS stack)."""
    from oracle import NODE_DTYPE

    t = np.zeros((n, 9), np.float32)
    for k in range(n):
        x = float(k + 1)
        t[k] = [x, -1, -1, x, 2, -1, x, -1, 2]
    nodes = np.zeros(2 * n, NODE_DTYPE)

    def box(lo, hi):
        v = t[lo:hi].reshape(-1, 3)
        return v.min(0), v.max(0)

    idx, k = 0, n  # node idx covers triangles [0, k)
    nxt = 2
    while k > 1:
        nodes[idx]["bmin"], nodes[idx]["bmax"] = box(0, k)
        nodes[idx]["left_first"], nodes[idx]["count"] = nxt, 0
        l, r = nxt, nxt + 1
        nodes[r]["bmin"], nodes[r]["bmax"] = box(k - 1, k)
        nodes[r]["left_first"], nodes[r]["count"] = k - 1, 1
        idx, k, nxt = l, k - 1, nxt + 2
    nodes[idx]["bmin"], nodes[idx]["bmax"] = box(0, 1)
    nodes[idx]["left_first"], nodes[idx]["count"] = 0, 1
    return nodes, nxt, t
