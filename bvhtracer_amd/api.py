"""Host-side mirror of the reference crate's traversal API over the C ABI.

Names and argument meaning follow bvhtracer (bvhtracer/src/lib.rs:16-27):

* ``BvhBuilder().build_for(mesh) -> Bvh``          (model/bvh.rs:509-541)
* ``Bvh.intersect(mesh, rays)``, ``Bvh.refit(mesh)``, ``Bvh.bounds()``,
  ``Bvh.nodes_used()``                             (model/bvh.rs:307-501)
* ``SceneObject`` with ``set_transform`` / ``bounds`` (scene/scene_object.rs)
* ``TlasBuilder().build_for(objects) -> Tlas``, ``Tlas.intersect(objects,
  rays)``, ``Tlas.rebuild(objects)``               (scene/tlas.rs:123-279)
* ``Ray.new`` / ``Ray.from_origin_dir``            (query/ray.rs:23-35)

The unit of work is a batch: ``intersect`` takes an array of rays (numpy
records of ``RAY_DTYPE``, or a ``DeviceArray`` already resident on the GPU)
and returns hit records.  ``intersect_one`` keeps the reference's per-ray
``Option<Intersection>`` shape (None on a miss).  Errors raise ``BvhtError``
instead of panicking.  There is no CPU path: every call runs on the GPU
through libbvht_hip.so.
"""
import ctypes
import weakref

import numpy as np

from . import _lib as L
from ._lib import (BUILD_REFERENCE, BUILD_SAH, CAMERA_DTYPE, HIT_DTYPE, INSTANCE_DTYPE, MISS, NODE_DTYPE,
                   ORDER_ROWS, ORDER_TILES, RAY_DTYPE, STATS_DTYPE, TNODE_DTYPE, BvhtError, check)

F32_MAX = np.float32(np.finfo(np.float32).max)


class DeviceArray:
    """A device allocation owned by a Context (freed on ``free``/GC)."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        self.owned = True
        p = ctypes.c_void_p()
        check(L.lib().bvht_device_alloc(ctx.handle, self.nbytes, ctypes.byref(p)))
        self.ptr = p
        ctx.adopt(self)

    @classmethod
    def view(cls, ctx, address, nbytes):
        """Non-owning handle on device memory owned by the library (e.g. a
        mesh's vertex buffer)."""
        v = cls.__new__(cls)
        v.ctx, v.nbytes, v.owned, v.ptr = ctx, int(nbytes), False, ctypes.c_void_p(address)
        return v

    @property
    def value(self):
        return self.ptr.value

    def offset(self, nbytes):
        return ctypes.c_void_p(self.ptr.value + int(nbytes))

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        check(L.lib().bvht_copy_to_device(self.ctx.handle, self.ptr, L.ptr(a), a.nbytes))
        return self

    def download(self, dtype, count=None):
        dtype = np.dtype(dtype)
        if count is None:
            count = self.nbytes // dtype.itemsize
        out = np.empty(count, dtype)
        check(L.lib().bvht_copy_to_host(self.ctx.handle, L.ptr(out), self.ptr, out.nbytes))
        return out

    def free(self):
        if self.owned and self.ptr is not None and self.ptr.value:
            L.lib().bvht_device_free(self.ctx.handle, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            if self.ptr is not None and self.ctx.handle:
                self.free()
        except Exception:
            pass


class Context:
    """A HIP device + stream (bvht_create).  One per host thread."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        check(L.lib().bvht_create(device, ctypes.byref(h)))
        self.handle = h
        self._children = weakref.WeakSet()

    def adopt(self, obj):
        """Register an object holding library resources of this context;
        close() releases it first (the C ABI requires objects to be destroyed
        before their context)."""
        self._children.add(obj)
        return obj

    def alloc(self, nbytes):
        return DeviceArray(self, nbytes)

    def to_device(self, arr):
        a = np.ascontiguousarray(arr)
        return DeviceArray(self, max(a.nbytes, 1)).upload(a)

    def sync(self):
        """Wait for the stream; raises BvhtError (BVHT_ESTACK) if a trace since
        the last fence dropped stack entries at the stack limit."""
        check(L.lib().bvht_sync(self.handle))

    def set_stack_limit(self, max_entries):
        """Cap the traversal stack of this context's traces (0: tree depth)."""
        check(L.lib().bvht_set_stack_limit(self.handle, int(max_entries)))

    def set_wave_schedule(self, enable=True):
        """Longest-first wave order for repeated TLAS traces of the same rays
        (bvht_set_wave_schedule); hits are the same either way."""
        check(L.lib().bvht_set_wave_schedule(self.handle, 1 if enable else 0))

    def event(self):
        e = ctypes.c_void_p()
        check(L.lib().bvht_event_create(self.handle, ctypes.byref(e)))
        return e

    def record(self, ev):
        check(L.lib().bvht_event_record(self.handle, ev))

    @staticmethod
    def elapsed_ms(a, b):
        ms = ctypes.c_float()
        check(L.lib().bvht_event_elapsed(a, b, ctypes.byref(ms)))
        return ms.value

    def close(self):
        if self.handle:
            for obj in list(self._children):
                obj.close() if not isinstance(obj, DeviceArray) else obj.free()
            L.lib().bvht_destroy(self.handle)
            self.handle = None


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class Ray:
    """Ray constructors (query/ray.rs:23-35) producing RAY_DTYPE records."""

    @staticmethod
    def new(origin, direction, t):
        r = np.zeros(1, RAY_DTYPE)
        r["o"][0] = np.asarray(origin, np.float32)
        r["d"][0] = np.asarray(direction, np.float32)
        r["tmax"][0] = np.float32(t)
        return r

    @staticmethod
    def from_origin_dir(origin, direction):
        return Ray.new(origin, direction, F32_MAX)


def _as_tris(mesh):
    a = np.asarray(mesh)
    if a.dtype != np.float32 or not a.flags.c_contiguous:
        raise TypeError("mesh must be a C-contiguous float32 array of shape (N, 9)")
    return a.reshape(-1, 9)


class Bvh:
    """A bottom-level acceleration structure resident on the GPU."""

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.handle = handle
        self.ctx.adopt(self)

    @classmethod
    def from_arrays(cls, nodes, nodes_used, mesh, ctx=None):
        """Adopt an existing node array + build-order triangles (bvht_blas_upload)."""
        ctx = ctx or default_context()
        nodes = np.ascontiguousarray(nodes, NODE_DTYPE)
        tris = np.ascontiguousarray(np.asarray(mesh, np.float32).reshape(-1, 9))
        h = ctypes.c_void_p()
        check(L.lib().bvht_blas_upload(ctx.handle, L.ptr(nodes), nodes.shape[0], nodes_used, L.ptr(tris),
                                       tris.shape[0], ctypes.byref(h)))
        return cls(ctx, h)

    def _info(self):
        vals = [ctypes.c_uint32() for _ in range(4)]
        check(L.lib().bvht_blas_info(self.handle, *[ctypes.byref(v) for v in vals]))
        return [v.value for v in vals]

    def nodes_used(self):
        return self._info()[2]

    def depth(self):
        return self._info()[3]

    def len_primitives(self):
        return self._info()[0]

    def bounds(self):
        out = np.zeros(6, np.float32)
        check(L.lib().bvht_blas_bounds(self.ctx.handle, self.handle, L.ptr(out)))
        return out[:3].copy(), out[3:].copy()

    def download(self):
        """(nodes[node_count], nodes_used, tris[N, 9]) as numpy arrays."""
        n, nc, used, _ = self._info()
        nodes = np.zeros(nc, NODE_DTYPE)
        tris = np.zeros((n, 9), np.float32)
        check(L.lib().bvht_blas_download(self.ctx.handle, self.handle, L.ptr(nodes), L.ptr(tris)))
        return nodes, used, tris

    def refit(self, mesh=None):
        """Bvh::refit (bvh.rs:469-493).  ``mesh``: new vertices in build order
        (host array or DeviceArray); None refits against the current ones."""
        if mesh is not None:
            if isinstance(mesh, DeviceArray):
                check(L.lib().bvht_blas_set_vertices_device(self.ctx.handle, self.handle, mesh.ptr,
                                                            self.len_primitives()))
            else:
                t = _as_tris(mesh)
                check(L.lib().bvht_blas_set_vertices(self.ctx.handle, self.handle, L.ptr(t), t.shape[0]))
        check(L.lib().bvht_blas_refit(self.ctx.handle, self.handle))

    def permutation(self):
        """perm[k] = input index of the triangle at build position k (GPU
        builds only; the reference keeps no permutation)."""
        out = np.zeros(self.len_primitives(), np.uint32)
        check(L.lib().bvht_blas_permutation(self.ctx.handle, self.handle, L.ptr(out)))
        return out

    def intersect_device(self, d_rays, n, d_hits, d_stats=None):
        """Enqueue a trace of n device-resident rays into device hits."""
        check(L.lib().bvht_trace_blas(self.ctx.handle, self.handle, d_rays.ptr, n, d_hits.ptr,
                                      d_stats.ptr if d_stats is not None else None))

    def intersect(self, mesh_or_rays, rays=None, stats=False):
        """Bvh::intersect for a batch of rays.  Accepts (mesh, rays) like the
        reference (the mesh must be the one the BVH was built over and is not
        re-read) or just (rays)."""
        rays = mesh_or_rays if rays is None else rays
        rays = np.ascontiguousarray(rays, RAY_DTYPE).reshape(-1)
        n = rays.shape[0]
        if n == 0:
            return (np.zeros(0, HIT_DTYPE), np.zeros(0, STATS_DTYPE)) if stats else np.zeros(0, HIT_DTYPE)
        d_rays = self.ctx.to_device(rays)
        d_hits = self.ctx.alloc(n * HIT_DTYPE.itemsize)
        d_stats = self.ctx.alloc(n * STATS_DTYPE.itemsize) if stats else None
        self.intersect_device(d_rays, n, d_hits, d_stats)
        hits = d_hits.download(HIT_DTYPE, n)
        if stats:
            return hits, d_stats.download(STATS_DTYPE, n)
        return hits

    def intersect_one(self, mesh, ray):
        """Reference shape: None on a miss, else the hit record."""
        h = self.intersect(mesh, ray)[0]
        return None if h["prim"] == MISS else h

    def close(self):
        if self.handle:
            L.lib().bvht_blas_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BvhBuilder:
    """BvhBuilder (bvh.rs:504-541).  ``mode`` BUILD_REFERENCE reproduces the
    reference exactly; BUILD_SAH uses empty-initialised bins and splits every
    node with more than ``max_leaf`` primitives."""

    def __init__(self, mode=BUILD_REFERENCE, max_leaf=0, ctx=None):
        self.mode = mode
        self.max_leaf = max_leaf
        self.ctx = ctx or default_context()

    def build_for(self, mesh):
        """Builds on the GPU; ``mesh`` ((N, 9) float32) is permuted in place
        into build order, as the reference permutes its slice."""
        t = _as_tris(mesh)
        h = ctypes.c_void_p()
        check(L.lib().bvht_blas_build(self.ctx.handle, L.ptr(t), t.shape[0], self.mode, self.max_leaf,
                                      ctypes.byref(h)))
        return Bvh(self.ctx, h)


class SceneObject:
    """A model instance with transform M and cached M^-1
    (scene/scene_object.rs, transform_component.rs)."""

    def __init__(self, model_index, transform=None, transform_inv=None):
        self.model = int(model_index)
        m = np.eye(4, dtype=np.float32) if transform is None else np.asarray(transform, np.float32).reshape(4, 4)
        self.m = m.reshape(16).copy()  # column-major m[col*4+row]
        if transform_inv is None:
            # cached inverse (TransformComponent3::new); float64 inverse of the float32 matrix
            colmajor = self.m.reshape(4, 4).astype(np.float64)  # [col][row]
            inv = np.linalg.inv(colmajor.T).T
            self.minv = inv.astype(np.float32).reshape(16)
        else:
            self.minv = np.asarray(transform_inv, np.float32).reshape(16).copy()

    def set_transform(self, transform, transform_inv=None):
        o = SceneObject(self.model, transform, transform_inv)
        self.m, self.minv = o.m, o.minv

    def record(self):
        r = np.zeros(1, INSTANCE_DTYPE)
        r["m"][0] = self.m
        r["minv"][0] = self.minv
        r["blas"][0] = self.model
        return r


def _instances(objects):
    return np.concatenate([o.record() for o in objects]) if objects else np.zeros(0, INSTANCE_DTYPE)


class Tlas:
    """Top-level structure over scene objects (scene/tlas.rs)."""

    def __init__(self, ctx, handle, n):
        self.ctx = ctx
        self.handle = handle
        self.ctx.adopt(self)
        self.n = n

    def rebuild(self, objects=None):
        """Scene::rebuild / Tlas::rebuild; if ``objects`` is given their
        transforms are set first (SceneObject::set_transform for each)."""
        if objects is not None:
            inst = _instances(objects)
            check(L.lib().bvht_tlas_set_objects(self.ctx.handle, self.handle, L.ptr(inst), 0, inst.shape[0]))
        check(L.lib().bvht_tlas_rebuild(self.ctx.handle, self.handle))

    def refresh_bounds(self):
        check(L.lib().bvht_tlas_refresh_bounds(self.ctx.handle, self.handle))

    def rebuild_async(self):
        """Tlas::rebuild without a host round trip (<= 64 objects)."""
        check(L.lib().bvht_tlas_rebuild_async(self.ctx.handle, self.handle))

    def download(self):
        nodes = np.zeros(2 * self.n, TNODE_DTYPE)
        used = ctypes.c_uint32()
        check(L.lib().bvht_tlas_download(self.ctx.handle, self.handle, L.ptr(nodes), ctypes.byref(used)))
        return nodes, used.value

    def intersect_device(self, d_rays, n, d_hits, d_stats=None):
        check(L.lib().bvht_trace_tlas(self.ctx.handle, self.handle, d_rays.ptr, n, d_hits.ptr,
                                      d_stats.ptr if d_stats is not None else None))

    def intersect(self, objects_or_rays, rays=None, stats=False):
        rays = objects_or_rays if rays is None else rays
        rays = np.ascontiguousarray(rays, RAY_DTYPE).reshape(-1)
        n = rays.shape[0]
        d_rays = self.ctx.to_device(rays)
        d_hits = self.ctx.alloc(max(n, 1) * HIT_DTYPE.itemsize)
        d_stats = self.ctx.alloc(max(n, 1) * STATS_DTYPE.itemsize) if stats else None
        self.intersect_device(d_rays, n, d_hits, d_stats)
        hits = d_hits.download(HIT_DTYPE, n)
        return (hits, d_stats.download(STATS_DTYPE, n)) if stats else hits

    def close(self):
        if self.handle:
            L.lib().bvht_tlas_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TlasBuilder:
    """TlasBuilder (tlas.rs:253-279): ``build_for(objects, models)`` where
    ``models`` are the Bvh handles the objects' ``model`` indices refer to."""

    def __init__(self, ctx=None):
        self.ctx = ctx or default_context()

    def build_for(self, objects, models):
        inst = _instances(objects)
        arr = (ctypes.c_void_p * len(models))(*[m.handle.value for m in models])
        h = ctypes.c_void_p()
        check(L.lib().bvht_tlas_create(self.ctx.handle, ctypes.cast(arr, ctypes.c_void_p), len(models), L.ptr(inst),
                                       inst.shape[0], ctypes.byref(h)))
        t = Tlas(self.ctx, h, inst.shape[0])
        t._models = list(models)  # keep the BLASes alive
        return t


class Batch:
    """Many scenes (Tlas objects) updated and traced with one launch per stage
    (bvht_batch_*).  Rays of all scenes live in one device buffer; scene s
    owns rays [offsets[s], offsets[s+1])."""

    def __init__(self, ctx, scenes):
        self.ctx = ctx
        self.scenes = list(scenes)
        arr = (ctypes.c_void_p * len(self.scenes))(*[t.handle.value for t in self.scenes])
        h = ctypes.c_void_p()
        check(L.lib().bvht_batch_create(ctx.handle, ctypes.cast(arr, ctypes.c_void_p), len(self.scenes),
                                        ctypes.byref(h)))
        self.handle = h
        self.ctx.adopt(self)

    def refit(self):
        check(L.lib().bvht_batch_refit(self.ctx.handle, self.handle))

    def rebuild(self):
        check(L.lib().bvht_batch_rebuild(self.ctx.handle, self.handle))

    def set_rays(self, offsets, rays_per_item=128):
        off = np.ascontiguousarray(offsets, np.uint32)
        assert off.shape[0] == len(self.scenes) + 1
        check(L.lib().bvht_batch_set_rays(self.ctx.handle, self.handle, L.ptr(off), rays_per_item))

    def config(self, wg, lds_depth, rays_per_item):
        check(L.lib().bvht_batch_config(self.ctx.handle, self.handle, wg, lds_depth, rays_per_item))

    def trace(self, d_rays, d_hits):
        check(L.lib().bvht_batch_trace(self.ctx.handle, self.handle, d_rays.ptr, d_hits.ptr))

    def close(self):
        if self.handle:
            L.lib().bvht_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """A hipGraph of work enqueued on a context (bvht_graph_*)."""

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.handle = handle
        self.ctx.adopt(self)

    @classmethod
    def capture(cls, ctx, enqueue, keep=()):
        """Capture everything ``enqueue()`` launches on ctx's stream.  The
        graph holds device addresses of what it launched on: it keeps
        ``enqueue`` (with everything its closure refers to) and the objects in
        ``keep`` alive until it is closed."""
        check(L.lib().bvht_graph_begin(ctx.handle))
        try:
            enqueue()
        finally:
            h = ctypes.c_void_p()
            rc = L.lib().bvht_graph_end(ctx.handle, ctypes.byref(h))
        check(rc)
        g = cls(ctx, h)
        g._keep = (enqueue, tuple(keep))
        return g

    def launch(self):
        check(L.lib().bvht_graph_launch(self.ctx.handle, self.handle))

    def close(self):
        if self.handle:
            L.lib().bvht_graph_destroy(self.handle)
            self.handle = None
            self._keep = ()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def deform_wave(ctx, bvh, d_orig, amp, kx, w, d_frame):
    """BLAS vertices = originals with y += amp*sin(kx*x + w*frame), frame = *d_frame."""
    check(L.lib().bvht_deform_wave(ctx.handle, bvh.handle, d_orig.ptr, amp, kx, w, d_frame.ptr))


def counter_add(ctx, d_value, v=1):
    check(L.lib().bvht_counter_add(ctx.handle, d_value.ptr, v))


def raygen(ctx, camera, width, height, order=ORDER_ROWS, d_rays=None):
    """Camera::get_ray_world for every pixel into device memory."""
    cam = np.ascontiguousarray(camera, CAMERA_DTYPE).reshape(1)
    n = width * height
    d_rays = d_rays or ctx.alloc(n * RAY_DTYPE.itemsize)
    check(L.lib().bvht_raygen(ctx.handle, L.ptr(cam), width, height, order, d_rays.ptr))
    return d_rays


def bounce(ctx, bvh, d_rays, d_hits, n, seed, d_out=None, d_count=None):
    d_out = d_out or ctx.alloc(max(n, 1) * RAY_DTYPE.itemsize)
    d_count = d_count or ctx.alloc(4)
    check(L.lib().bvht_bounce(ctx.handle, bvh.handle, d_rays.ptr, d_hits.ptr, n, seed, d_out.ptr, d_count.ptr))
    return d_out, d_count


def bounce_sorted(ctx, bvh, d_rays, d_hits, n, seed, d_out=None, d_src=None, d_count=None):
    """bounce() with the rays grouped by direction bin (ray order inside a
    bin); d_src[k] = index of the ray that spawned bounce ray k."""
    d_out = d_out or ctx.alloc(max(n, 1) * RAY_DTYPE.itemsize)
    d_src = d_src or ctx.alloc(max(n, 1) * 4)
    d_count = d_count or ctx.alloc(4)
    check(L.lib().bvht_bounce_sorted(ctx.handle, bvh.handle, d_rays.ptr, d_hits.ptr, n, seed, d_out.ptr, d_src.ptr,
                                     d_count.ptr))
    return d_out, d_src, d_count


def trace_camera_bounce(ctx, bvh, cam, width, height, order, seed, d_hits=None, d_out=None, d_src=None,
                        d_count=None):
    """raygen + Bvh trace + bounce_sorted of one camera frame in one pass
    (bvht_trace_camera_bounce): returns (d_hits, d_bounce_rays, d_src, d_count)."""
    n = width * height
    cam = np.ascontiguousarray(cam, CAMERA_DTYPE).reshape(1)
    d_hits = d_hits or ctx.alloc(max(n, 1) * HIT_DTYPE.itemsize)
    d_out = d_out or ctx.alloc(max(n, 1) * RAY_DTYPE.itemsize)
    d_src = d_src or ctx.alloc(max(n, 1) * 4)
    d_count = d_count or ctx.alloc(4)
    check(L.lib().bvht_trace_camera_bounce(ctx.handle, bvh.handle, L.ptr(cam), width, height, order, seed,
                                           d_hits.ptr, d_out.ptr, d_src.ptr, d_count.ptr))
    return d_hits, d_out, d_src, d_count


def bounce_tlas_sorted(ctx, tlas, d_rays, d_hits, n, seed, d_out=None, d_src=None, d_count=None):
    d_out = d_out or ctx.alloc(max(n, 1) * RAY_DTYPE.itemsize)
    d_src = d_src or ctx.alloc(max(n, 1) * 4)
    d_count = d_count or ctx.alloc(4)
    check(L.lib().bvht_bounce_tlas_sorted(ctx.handle, tlas.handle, d_rays.ptr, d_hits.ptr, n, seed, d_out.ptr,
                                          d_src.ptr, d_count.ptr))
    return d_out, d_src, d_count


def bounce_tlas(ctx, tlas, d_rays, d_hits, n, seed, d_out=None, d_count=None):
    d_out = d_out or ctx.alloc(max(n, 1) * RAY_DTYPE.itemsize)
    d_count = d_count or ctx.alloc(4)
    check(L.lib().bvht_bounce_tlas(ctx.handle, tlas.handle, d_rays.ptr, d_hits.ptr, n, seed, d_out.ptr,
                                   d_count.ptr))
    return d_out, d_count


# ------------------------------------------------ differentiable depth/shading


class Mesh:
    """Indexed mesh (V (n, 3) float32, F (m, 3) uint32) bound to a BLAS built
    from its face soup ``V[F].reshape(-1, 9)`` (face order), or to any BLAS
    whose triangle k is face ``face_of[k]``.  Owns the device vertices and the
    area-weighted vertex normals used by trace_forward / trace_adjoint."""

    def __init__(self, bvh, vertices, faces, face_of=None):
        self.ctx = bvh.ctx
        self.bvh = bvh
        V = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
        F = np.ascontiguousarray(faces, np.uint32).reshape(-1, 3)
        fo = None if face_of is None else np.ascontiguousarray(face_of, np.uint32)
        if fo is not None and fo.shape[0] != bvh.len_primitives():
            raise ValueError("face_of needs one entry per BLAS triangle")
        h = ctypes.c_void_p()
        check(L.lib().bvht_mesh_create(self.ctx.handle, bvh.handle, L.ptr(V), V.shape[0], L.ptr(F), F.shape[0],
                                       L.ptr(fo) if fo is not None else None, ctypes.byref(h)))
        self.handle = h
        self.ctx.adopt(self)
        self.n_vertices, self.n_faces = V.shape[0], F.shape[0]
        self.faces = F
        self.face_of = fo if fo is not None else bvh.permutation()
        pv, pn = ctypes.c_void_p(), ctypes.c_void_p()
        check(L.lib().bvht_mesh_buffers(h, ctypes.byref(pv), ctypes.byref(pn)))
        nb = 12 * self.n_vertices
        self.vertices = DeviceArray.view(self.ctx, pv.value, nb)
        self.normals = DeviceArray.view(self.ctx, pn.value, nb)

    def prim_vertices(self):
        """(n_tris, 3) vertex ids of each BLAS triangle (build order)."""
        return self.faces[self.face_of]

    def update(self, normals=True):
        """Publish changed vertices: gather into the BLAS, refit, and (if
        ``normals``) recompute the vertex normals."""
        check(L.lib().bvht_mesh_update(self.ctx.handle, self.handle, 1 if normals else 0))

    def set_vertices(self, V, normals=True):
        self.vertices.upload(np.ascontiguousarray(V, np.float32).reshape(-1, 3))
        self.update(normals)

    def sgd(self, d_grad, lr, max_step=0.0, along_normal=False):
        """v -= lr * g_v per vertex (along_normal: only the component along
        the vertex normal), the step clipped to max_step (<= 0: no clip)."""
        check(L.lib().bvht_mesh_sgd(self.ctx.handle, self.handle, d_grad.ptr, lr, max_step,
                                    L.SGD_ALONG_NORMAL if along_normal else 0))

    def download(self):
        """(vertices, normals) as (n, 3) float32 arrays."""
        return (self.vertices.download(np.float32).reshape(-1, 3), self.normals.download(np.float32).reshape(-1, 3))

    def close(self):
        if self.handle:
            L.lib().bvht_mesh_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _mat16(m):
    """16 floats, column-major m[col*4+row] (the SceneObject convention)."""
    return np.ascontiguousarray(np.asarray(m, np.float32).reshape(16))


def trace_forward(mesh, minv, light, d_rays, n, d_hits, d_image):
    """Trace n device rays through the mesh's object (inverse transform
    ``minv``: 16 floats, column-major) and shade:
    d_hits[n], d_image[2n] = {depth, shade}."""
    m = _mat16(minv)
    lt = np.ascontiguousarray(light, np.float32).reshape(3)
    check(L.lib().bvht_trace_forward(mesh.ctx.handle, mesh.handle, L.ptr(m), L.ptr(lt), d_rays.ptr, n, d_hits.ptr,
                                     d_image.ptr))


def trace_adjoint(mesh, minv, light, d_rays, n, d_hits, d_target, d_grad_vertices, d_grad_minv, d_loss,
                  w_depth=1.0, w_shade=1.0, accumulate=False):
    """dLoss/dvertex (float[3 n_vertices]), dLoss/dM^-1 (float[16],
    column-major) and the loss (double) into device buffers."""
    m = _mat16(minv)
    lt = np.ascontiguousarray(light, np.float32).reshape(3)
    check(L.lib().bvht_trace_adjoint(mesh.ctx.handle, mesh.handle, L.ptr(m), L.ptr(lt), d_rays.ptr, n, d_hits.ptr,
                                     d_target.ptr, w_depth, w_shade, 1 if accumulate else 0, d_grad_vertices.ptr,
                                     d_grad_minv.ptr, d_loss.ptr))


def transform_gradient(minv, grad_minv):
    """dLoss/dM from dLoss/dM^-1 (both flat column-major): -(M^-1)^T G (M^-1)^T,
    returned flat column-major (float64)."""
    A = np.asarray(minv, np.float64).reshape(4, 4).T  # A[r, c]
    G = np.asarray(grad_minv, np.float64).reshape(4, 4).T
    return np.ascontiguousarray((-(A.T @ G @ A.T)).T.reshape(16))
