"""CPU oracle for the ray/BVH traversal hot path -- TEST INFRASTRUCTURE ONLY.

A plain-C restatement (``bvht_oracle.c``) of the reference crate's CPU
algorithms, loaded through ctypes.  Only the test-suite, the smoke check and
the CPU-baseline leg of the benchmark use it, always as the checker or as the
timed CPU path; the product library never loads it.

Record layouts (numpy dtypes below) are byte-identical to the product's C ABI
so the same buffers can be handed to both.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

NODE_DTYPE = np.dtype([("bmin", "<f4", 3), ("left_first", "<u4"), ("bmax", "<f4", 3), ("count", "<u4")])
TNODE_DTYPE = np.dtype([("bmin", "<f4", 3), ("left_right", "<u4"), ("bmax", "<f4", 3), ("blas", "<u4")])
RAY_DTYPE = np.dtype([("o", "<f4", 3), ("tmax", "<f4"), ("d", "<f4", 3), ("pad", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("prim", "<u4"), ("inst", "<u4"), ("packed", "<u4")])
INSTANCE_DTYPE = np.dtype([("minv", "<f4", 16), ("blas", "<u4"), ("pad", "<u4", 3)])
STATS_DTYPE = np.dtype([("pairs", "<u4"), ("tests", "<u4"), ("tlas_pairs", "<u4"), ("inst_visits", "<u4")])
CAMERA_DTYPE = np.dtype([("tl", "<f4", 3), ("tr", "<f4", 3), ("bl", "<f4", 3), ("view_inv", "<f4", 16)])

BODY_DTYPE = np.dtype([
    ("inverse_mass", "<f4"), ("inverse_inertia_tensor", "<f4", 9), ("linear_drag", "<f4"), ("angular_drag", "<f4"),
    ("position", "<f4", 3), ("orientation", "<f4", 4), ("velocity", "<f4", 3), ("rotation", "<f4", 3),
    ("inverse_inertia_tensor_world", "<f4", 9), ("motion", "<f4"), ("is_awake", "<u4"), ("can_sleep", "<u4"),
    ("transform", "<f4", 16), ("force_accumulator", "<f4", 3), ("torque_accumulator", "<f4", 3),
    ("acceleration", "<f4", 3), ("last_frame_acceleration", "<f4", 3), ("pad", "<f4", 3)])
FORCE_DTYPE = np.dtype([("kind", "<u4"), ("body", "<u4"), ("other", "<u4"), ("pad", "<u4"), ("v0", "<f4", 3),
                        ("v1", "<f4", 3), ("spring_constant", "<f4"), ("rest_length", "<f4")])

assert NODE_DTYPE.itemsize == 32 and TNODE_DTYPE.itemsize == 32
assert BODY_DTYPE.itemsize == 272 and FORCE_DTYPE.itemsize == 48
assert RAY_DTYPE.itemsize == 32 and HIT_DTYPE.itemsize == 24 and INSTANCE_DTYPE.itemsize == 80

_lib = None


def build(force=False):
    """Compile liboracle.so with the committed Makefile."""
    import subprocess

    if force and os.path.exists(_LIB_PATH):
        os.remove(_LIB_PATH)
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run `make -C oracle`)")
        _lib = _load(_LIB_PATH)
    return _lib


_native = {}


def native_lib(out_dir):
    """The same oracle compiled with -march=native for the running host (the
    CPU baseline's timed build), made in ``out_dir`` on first use."""
    import subprocess

    path = os.path.join(out_dir, "liboracle_native.so")
    if path not in _native:
        subprocess.run(["make", "-s", "-B", "-C", _HERE, "native", f"NATIVE={path}"], check=True)
        _native[path] = _load(path)
    return _native[path]


def _load(path):
    L = ctypes.CDLL(path)
    P = ctypes.c_void_p
    u32, u64, i32, f32 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_float
    sig = {
        "orc_normalize": (None, [P, P]),
        "orc_ray_recip": (None, [P, P]),
        "orc_aabb_intersect": (i32, [P, P, P, P]),
        "orc_triangle_intersect": (i32, [P, P, P, P, P]),
        "orc_triangle_centroid": (None, [P, P]),
        "orc_pack_index": (u32, [u32, u32]),
        "orc_bvh_build": (i32, [P, u32, i32, u32, P, P]),
        "orc_bvh_refit": (i32, [P, u32, P]),
        "orc_bvh_depth": (u32, [P]),
        "orc_bvh_intersect": (i32, [P, P, P, P, P]),
        "orc_bvh_trace": (i32, [P, P, P, u64, P, P, i32]),
        "orc_brute_trace": (i32, [P, u32, P, u64, P, P]),
        "orc_transform_point": (None, [P, P, P]),
        "orc_transform_vector": (None, [P, P, P]),
        "orc_instance_bounds": (None, [P, P, P, P, P]),
        "orc_tlas_build": (i32, [P, u32, P, P]),
        "orc_tlas_intersect": (i32, [P, P, P, P, P, P, P]),
        "orc_tlas_trace": (i32, [P, P, P, P, P, u64, P, P, i32]),
        "orc_camera_ray": (None, [P, f32, f32, P]),
        "orc_raygen": (i32, [P, u32, u32, P]),
        "orc_mesh_normals": (None, [P, u32, P, u32, P]),
        "orc_diff_forward": (i32, [P, P, P, P, P, P, P, u64, P, P, i32]),
        "orc_diff_grad64": (i32, [P, u32, P, u32, P, P, P, P, P, u64, P, ctypes.c_double, ctypes.c_double,
                                  P, P, P, P]),
        "orc_powf": (ctypes.c_float, [ctypes.c_float, ctypes.c_float]),
        "orc_world_start_frame": (None, [P, u32]),
        "orc_world_run_physics": (None, [P, u32, P, u32, ctypes.c_float]),
        "orc_objects_from_bodies": (None, [P, P, P, u32, P, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f3(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.float32).reshape(3))


def make_ray(o, d, tmax=np.finfo(np.float32).max):
    r = np.zeros(1, RAY_DTYPE)
    r["o"][0] = _f3(o)
    r["d"][0] = _f3(d)
    r["tmax"][0] = np.float32(tmax)
    return r


def normalize(v):
    out = np.zeros(3, np.float32)
    lib().orc_normalize(_p(_f3(v)), _p(out))
    return out


def ray_recip(d):
    out = np.zeros(3, np.float32)
    lib().orc_ray_recip(_p(_f3(d)), _p(out))
    return out


def aabb_intersect(bmin, bmax, ray):
    t = np.zeros(1, np.float32)
    hit = lib().orc_aabb_intersect(_p(_f3(bmin)), _p(_f3(bmax)), _p(ray), _p(t))
    return float(t[0]) if hit else None


def triangle_intersect(tri, ray):
    tri = np.ascontiguousarray(np.asarray(tri, np.float32).reshape(9))
    out = np.zeros(3, np.float32)
    hit = lib().orc_triangle_intersect(_p(tri), _p(ray), _p(out[0:]), _p(out[1:]), _p(out[2:]))
    return (np.float32(out[0]), np.float32(out[1]), np.float32(out[2])) if hit else None


def centroid(tri):
    tri = np.ascontiguousarray(np.asarray(tri, np.float32).reshape(9))
    out = np.zeros(3, np.float32)
    lib().orc_triangle_centroid(_p(tri), _p(out))
    return out


def pack_index(inst, prim):
    return int(lib().orc_pack_index(inst, prim))


def bvh_build(tris, faithful=True, max_leaf=0):
    """Returns (nodes[2N], nodes_used, permuted_tris[N,9]); the input is not modified."""
    t = np.ascontiguousarray(np.asarray(tris, np.float32).reshape(-1, 9)).copy()
    n = t.shape[0]
    nodes = np.zeros(max(2 * n, 2), NODE_DTYPE)
    used = np.zeros(1, np.uint32)
    rc = lib().orc_bvh_build(_p(t), n, 1 if faithful else 0, max_leaf, _p(nodes), _p(used))
    if rc != 0:
        raise ValueError(f"orc_bvh_build failed rc={rc}")
    return nodes, int(used[0]), t


def bvh_refit(nodes, nodes_used, tris):
    tris = np.ascontiguousarray(np.asarray(tris, np.float32).reshape(-1, 9))
    rc = lib().orc_bvh_refit(_p(nodes), nodes_used, _p(tris))
    if rc != 0:
        raise ValueError(f"orc_bvh_refit failed rc={rc}")
    return nodes


def bvh_depth(nodes):
    return int(lib().orc_bvh_depth(_p(nodes)))


def bvh_trace(nodes, tris, rays, nthreads=1, stats=False, using=None):
    """Bvh::intersect for every ray on nthreads threads (contiguous ray
    chunks); ``using``: a library from native_lib() instead of the default."""
    tris = np.ascontiguousarray(np.asarray(tris, np.float32).reshape(-1, 9))
    rays = np.ascontiguousarray(rays)
    assert rays.dtype == RAY_DTYPE
    hits = np.zeros(rays.shape[0], HIT_DTYPE)
    st = np.zeros(rays.shape[0], STATS_DTYPE) if stats else None
    rc = (using or lib()).orc_bvh_trace(_p(nodes), _p(tris), _p(rays), rays.shape[0], _p(hits),
                             _p(st) if stats else None, nthreads)
    if rc != 0:
        raise ValueError(f"orc_bvh_trace failed rc={rc}")
    return (hits, st) if stats else hits


def brute_trace(tris, rays):
    """Nearest hit over all triangles in array order (no BVH); returns (hits, ties)."""
    tris = np.ascontiguousarray(np.asarray(tris, np.float32).reshape(-1, 9))
    rays = np.ascontiguousarray(rays)
    hits = np.zeros(rays.shape[0], HIT_DTYPE)
    ties = np.zeros(rays.shape[0], np.uint32)
    rc = lib().orc_brute_trace(_p(tris), tris.shape[0], _p(rays), rays.shape[0], _p(hits), _p(ties))
    if rc != 0:
        raise ValueError(f"orc_brute_trace failed rc={rc}")
    return hits, ties


def transform_point(m, p):
    m = np.ascontiguousarray(np.asarray(m, np.float32).reshape(16))
    out = np.zeros(3, np.float32)
    lib().orc_transform_point(_p(m), _p(_f3(p)), _p(out))
    return out


def transform_vector(m, v):
    m = np.ascontiguousarray(np.asarray(m, np.float32).reshape(16))
    out = np.zeros(3, np.float32)
    lib().orc_transform_vector(_p(m), _p(_f3(v)), _p(out))
    return out


def instance_bounds(m, model_bmin, model_bmax):
    m = np.ascontiguousarray(np.asarray(m, np.float32).reshape(16))
    bmin = np.zeros(3, np.float32)
    bmax = np.zeros(3, np.float32)
    lib().orc_instance_bounds(_p(m), _p(_f3(model_bmin)), _p(_f3(model_bmax)), _p(bmin), _p(bmax))
    return bmin, bmax


def tlas_build(bounds):
    """bounds: (N, 6) float32 rows (bmin, bmax).  Returns (tnodes[2N], nodes_used)."""
    b = np.ascontiguousarray(np.asarray(bounds, np.float32).reshape(-1, 6))
    n = b.shape[0]
    nodes = np.zeros(max(2 * n, 2), TNODE_DTYPE)
    used = np.zeros(1, np.uint32)
    rc = lib().orc_tlas_build(_p(b), n, _p(nodes), _p(used))
    if rc != 0:
        raise ValueError(f"orc_tlas_build failed rc={rc}")
    return nodes, int(used[0])


def _blas_ptr_arrays(blases):
    """blases: list of (nodes, tris).  Returns ctypes pointer arrays + keep-alive list."""
    keep = []
    nptr = (ctypes.c_void_p * len(blases))()
    tptr = (ctypes.c_void_p * len(blases))()
    for i, (nodes, tris) in enumerate(blases):
        tris = np.ascontiguousarray(np.asarray(tris, np.float32).reshape(-1, 9))
        nodes = np.ascontiguousarray(nodes)
        keep += [nodes, tris]
        nptr[i] = nodes.ctypes.data
        tptr[i] = tris.ctypes.data
    return nptr, tptr, keep


def tlas_trace(tnodes, instances, blases, rays, nthreads=1, stats=False):
    rays = np.ascontiguousarray(rays)
    instances = np.ascontiguousarray(instances)
    nptr, tptr, keep = _blas_ptr_arrays(blases)
    hits = np.zeros(rays.shape[0], HIT_DTYPE)
    st = np.zeros(rays.shape[0], STATS_DTYPE) if stats else None
    rc = lib().orc_tlas_trace(_p(tnodes), _p(instances), ctypes.cast(nptr, ctypes.c_void_p),
                              ctypes.cast(tptr, ctypes.c_void_p), _p(rays), rays.shape[0], _p(hits),
                              _p(st) if stats else None, nthreads)
    del keep
    if rc != 0:
        raise ValueError(f"orc_tlas_trace failed rc={rc}")
    return (hits, st) if stats else hits


def camera_ray(cam, u, v):
    out = np.zeros(1, RAY_DTYPE)
    lib().orc_camera_ray(_p(cam), np.float32(u), np.float32(v), _p(out))
    return out


def raygen(cam, width, height):
    rays = np.zeros(width * height, RAY_DTYPE)
    rc = lib().orc_raygen(_p(cam), width, height, _p(rays))
    if rc != 0:
        raise ValueError("orc_raygen failed")
    return rays


# ---- differentiable depth/shading checker (no reference counterpart) ----


def mesh_normals(V, F):
    """Area-weighted vertex normals in binary32, faces accumulated in order."""
    V = np.ascontiguousarray(V, np.float32).reshape(-1, 3)
    F = np.ascontiguousarray(F, np.uint32).reshape(-1, 3)
    N = np.zeros_like(V)
    lib().orc_mesh_normals(_p(V), V.shape[0], _p(F), F.shape[0], _p(N))
    return N


def diff_forward(nodes, tris, vid, normals, minv, light, rays, nthreads=1):
    """Binary32 forward: transform rays by minv (16 floats, column-major),
    trace the BLAS, shade.  Returns (hits, image (n, 2) = depth, shade)."""
    tris = np.ascontiguousarray(np.asarray(tris, np.float32).reshape(-1, 9))
    vid = np.ascontiguousarray(vid, np.uint32)
    normals = np.ascontiguousarray(normals, np.float32)
    minv = np.ascontiguousarray(minv, np.float32).reshape(16)
    light = np.ascontiguousarray(light, np.float32).reshape(3)
    rays = np.ascontiguousarray(rays)
    n = rays.shape[0]
    hits = np.zeros(n, HIT_DTYPE)
    image = np.zeros((n, 2), np.float32)
    rc = lib().orc_diff_forward(_p(np.ascontiguousarray(nodes)), _p(tris), _p(vid), _p(normals), _p(minv), _p(light),
                                _p(rays), n, _p(hits), _p(image), nthreads)
    if rc != 0:
        raise ValueError(f"orc_diff_forward failed rc={rc}")
    return hits, image


def diff_grad64(V, vid, N, minv, light, rays, prims, target, w_depth=1.0, w_shade=1.0):
    """Binary64 loss and gradient at fixed hit triangles.  Returns
    (loss, grad_V (nv, 3), grad_minv (16,), image (n, 2))."""
    V = np.ascontiguousarray(V, np.float64).reshape(-1, 3)
    vid = np.ascontiguousarray(vid, np.uint32).reshape(-1)
    N = np.ascontiguousarray(N, np.float64).reshape(-1, 3)
    minv = np.ascontiguousarray(minv, np.float64).reshape(16)
    light = np.ascontiguousarray(light, np.float64).reshape(3)
    rays = np.ascontiguousarray(rays)
    prims = np.ascontiguousarray(prims, np.uint32)
    target = np.ascontiguousarray(target, np.float32).reshape(-1, 2)
    n = rays.shape[0]
    loss = np.zeros(1, np.float64)
    gv = np.zeros_like(V)
    gm = np.zeros(16, np.float64)
    image = np.zeros((n, 2), np.float64)
    rc = lib().orc_diff_grad64(_p(V), V.shape[0], _p(vid), vid.shape[0] // 3, _p(N), _p(minv), _p(light), _p(rays),
                               _p(prims), n, _p(target), float(w_depth), float(w_shade), _p(loss), _p(gv), _p(gm),
                               _p(image))
    if rc != 0:
        raise ValueError(f"orc_diff_grad64 failed rc={rc}")
    return float(loss[0]), gv, gm, image


def powf(x, y):
    """f32::powf as the integrator evaluates it (binary64 restatement)."""
    return np.float32(lib().orc_powf(float(np.float32(x)), float(np.float32(y))))


def world_start_frame(bodies):
    """World::start_frame (physics/world.rs:44-49), in place."""
    assert bodies.dtype == BODY_DTYPE and bodies.flags.c_contiguous
    lib().orc_world_start_frame(_p(bodies), bodies.shape[0])
    return bodies


def world_run_physics(bodies, forces, duration):
    """World::run_physics (physics/world.rs:57-60), in place."""
    assert bodies.dtype == BODY_DTYPE and bodies.flags.c_contiguous
    forces = np.ascontiguousarray(forces, FORCE_DTYPE)
    lib().orc_world_run_physics(_p(bodies), bodies.shape[0], _p(forces) if forces.shape[0] else None,
                                forces.shape[0], float(np.float32(duration)))
    return bodies


def objects_from_bodies(bodies, object_body, init):
    """SceneObject::update_transform for each object: (M, M^-1) as (n, 16)
    column-major float32; rows of objects without a body stay zero."""
    object_body = np.ascontiguousarray(object_body, np.uint32)
    init = np.ascontiguousarray(init, np.float32).reshape(-1, 16)
    n = object_body.shape[0]
    m = np.zeros((n, 16), np.float32)
    minv = np.zeros((n, 16), np.float32)
    lib().orc_objects_from_bodies(_p(bodies), _p(object_body), _p(init), n, _p(m), _p(minv))
    return m, minv
