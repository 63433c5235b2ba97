"""Host mirror of the reference's model and scene types over the GPU library.

Model / ModelInstance / ModelBuilder (bvhtracer/src/model/model.rs:15-150),
SceneObjectBuilder (scene/scene_object.rs:92-135), Scene / SceneBuilder
(scene/scene.rs:7-96) and Triangle (geometry/triangle.rs:9-108) with the same
names and argument meaning.  Every query runs on the GPU (libbvht_hip.so);
these classes only hold host-side bookkeeping (meshes, transforms, object
lists) and device handles.
"""
import numpy as np

from . import _lib as L
from .api import (BUILD_REFERENCE, HIT_DTYPE, INSTANCE_DTYPE, MISS, RAY_DTYPE, BvhBuilder, SceneObject, TlasBuilder,
                  _as_tris, default_context)


class ModelInstance:
    """A shared model: mesh (triangles in build order) + its BVH on the GPU.
    Python references give the sharing of Rc<RefCell<Model>>
    (model.rs:15-58): every SceneObject built from one ModelInstance uses
    the same mesh and the same BLAS."""

    def __init__(self, mesh, bvh, attributes=None):
        self._mesh = mesh
        self.bvh = bvh
        self._attributes = attributes  # the decoded Mesh, attributes in build order

    def mesh(self):
        """The Mesh (texture coordinates and normals permuted with the
        triangles), when the model was built from one."""
        return self._attributes

    def set_attributes(self, normals=None, tex_coords=None, texture=None):
        """Upload shading attributes (build order): normals (n, 3, 3),
        tex_coords (n, 3, 2), texture (H, W, 3) uint8."""
        nrm = None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(-1)
        tc = None if tex_coords is None else np.ascontiguousarray(tex_coords, np.float32).reshape(-1)
        tx = None if texture is None else np.ascontiguousarray(texture, np.uint8)
        L.check(L.lib().bvht_blas_set_attributes(self.bvh.ctx.handle, self.bvh.handle,
                                                 L.ptr(nrm) if nrm is not None else None,
                                                 L.ptr(tc) if tc is not None else None,
                                                 L.ptr(tx) if tx is not None else None,
                                                 0 if tx is None else tx.shape[1], 0 if tx is None else tx.shape[0]))
        self._texture = tx

    def texture(self):
        return getattr(self, "_texture", None)

    def intersect(self, rays):
        """Model::intersect (model.rs:69-71) for a batch of model-space rays."""
        return self.bvh.intersect(rays)

    def intersect_device(self, d_rays, n, d_hits, d_stats=None):
        self.bvh.intersect_device(d_rays, n, d_hits, d_stats)

    def refit(self):
        """Model::refit (model.rs:73-75): Bvh::refit against the current
        primitives (edit them through primitives_mut())."""
        self.bvh.refit(self._mesh)

    def bounds(self):
        """Model-space bounds (model.rs:77-80): (bmin, bmax)."""
        return self.bvh.bounds()

    def model(self):
        return self

    def primitives(self):
        view = self._mesh.view()
        view.flags.writeable = False
        return view

    def primitives_mut(self):
        return self._mesh

    def len(self):
        return self._mesh.shape[0]

    def len_primitives(self):
        return self._mesh.shape[0]


class ModelBuilder:
    """ModelBuilder (model.rs:113-150).  build() builds the BVH on the GPU and
    permutes the mesh's triangles into build order, as
    BvhBuilder::build_for(mesh.primitives_mut()) does.  with_build() selects
    the builder (BUILD_REFERENCE reproduces the reference exactly)."""

    def __init__(self, ctx=None):
        self.ctx = ctx or default_context()
        self.mesh = np.zeros((0, 9), np.float32)
        self.attributes = None
        self.mode, self.max_leaf = BUILD_REFERENCE, 0

    def with_mesh(self, mesh):
        """``mesh``: a decoded Mesh (bvhtracer_amd.mesh) or an (N, 9) float32
        triangle array."""
        if hasattr(mesh, "primitives_mut"):
            self.attributes = mesh
            self.mesh = mesh.primitives_mut()
        else:
            self.attributes = None
            self.mesh = _as_tris(mesh)
        return self

    def with_build(self, mode, max_leaf=0):
        self.mode, self.max_leaf = mode, max_leaf
        return self

    def with_texture(self, texture):
        """ModelBuilder::with_texture (model.rs:135-139): an (H, W, 3) uint8
        RGB image (TextureMaterial)."""
        self.texture = np.ascontiguousarray(texture, np.uint8)
        if self.texture.ndim != 3 or self.texture.shape[2] != 3:
            raise ValueError("texture must be (H, W, 3) uint8 RGB")
        return self

    def build(self):
        bvh = BvhBuilder(self.mode, self.max_leaf, ctx=self.ctx).build_for(self.mesh)
        attrs = getattr(self, "attributes", None)
        if attrs is not None:
            attrs.permute(bvh.permutation())
        model = ModelInstance(self.mesh, bvh, attrs)
        tex = getattr(self, "texture", None)
        if attrs is not None or tex is not None:
            model.set_attributes(attrs.normals() if attrs is not None else None,
                                 attrs.tex_coords() if attrs is not None else None, tex)
        return model


class SceneObjectBuilder:
    """SceneObjectBuilder (scene_object.rs:92-135): a ModelInstance placed by
    a transform M (column-major 4x4; identity by default), following the
    rigid body ``rigid_body`` (a RigidBodyInstance of the scene's World) if
    given: SceneObject::update_transform then sets M = body transform x this
    transform."""

    def __init__(self, model, rigid_body=None):
        self.model = model
        self.rigid_body = rigid_body
        self.transform = None

    def with_transform(self, transform):
        self.transform = np.asarray(transform, np.float32).reshape(16)
        return self

    def build(self):
        obj = SceneObject(-1, self.transform)
        obj.instance = self.model
        obj.rigid_body = self.rigid_body
        obj.transform_init = obj.m.copy()
        return obj


class Scene:
    """Scene (scene.rs:7-58): objects + TLAS + active camera record + World.
    After run_physics the objects' transforms live on the device; the host
    records are refreshed when an object is read (get_unchecked)."""

    def __init__(self, tlas, objects, camera, physics=None):
        self.tlas = tlas
        self.objects = objects
        self.camera = camera
        self.physics = physics
        self._bound = False
        self._device_ahead = False

    def _sync(self):
        if self._device_ahead:
            n = len(self.objects)
            recs = np.zeros(n, INSTANCE_DTYPE)
            L.check(L.lib().bvht_tlas_get_objects(self.tlas.ctx.handle, self.tlas.handle, L.ptr(recs), 0, n))
            for o, r in zip(self.objects, recs):
                o.m, o.minv = r["m"].copy(), r["minv"].copy()
            self._device_ahead = False

    def get_unchecked(self, index):
        self._sync()
        return self.objects[index]

    def get_mut_unchecked(self, index):
        self._sync()
        return self.objects[index]

    def active_camera(self):
        return self.camera

    def intersect(self, rays):
        """Scene::intersect (scene.rs:32-34) for a batch of world-space rays."""
        return self.tlas.intersect(rays)

    def intersect_device(self, d_rays, n, d_hits, d_stats=None):
        self.tlas.intersect_device(d_rays, n, d_hits, d_stats)

    def rebuild(self):
        """Scene::rebuild (scene.rs:36-38): object transforms and world bounds
        from the current models' bounds, then Tlas::rebuild."""
        if self._device_ahead:
            self.tlas.rebuild()
        else:
            self.tlas.rebuild(self.objects)

    def _bind(self):
        if not self._bound:
            body = np.array([MISS if getattr(o, "rigid_body", None) is None else o.rigid_body.index
                             for o in self.objects], np.uint32)
            init = np.stack([getattr(o, "transform_init", o.m) for o in self.objects]).astype(np.float32)
            for o in self.objects:
                rb = getattr(o, "rigid_body", None)
                if rb is not None and rb.world is not self.physics:
                    raise ValueError("a scene object follows a body of another world")
            L.check(L.lib().bvht_tlas_bind_world(self.tlas.ctx.handle, self.tlas.handle, self.physics.device(),
                                                 L.ptr(body), L.ptr(init)))
            self._bound = True

    def run_physics(self, elapsed):
        """Scene::run_physics (scene.rs:40-45) on the device: World::run_physics,
        then SceneObject::update_transform of every object with a body."""
        if self.physics is None or len(self.physics) == 0:
            return
        self._bind()
        L.check(L.lib().bvht_scene_run_physics(self.tlas.ctx.handle, self.tlas.handle, np.float32(elapsed)))
        self._device_ahead = True

    def run(self, elapsed):
        """Scene::run (scene.rs:47-50)."""
        self.run_physics(elapsed)
        self.rebuild()


class SceneBuilder:
    """SceneBuilder (scene.rs:61-96).  Objects built by SceneObjectBuilder
    share BLASes by ModelInstance identity."""

    def __init__(self, camera=None, ctx=None):
        self.camera = camera
        self.objects = []
        self.physics = None
        self.ctx = ctx or default_context()

    def with_object(self, obj):
        self.objects.append(obj)
        return self

    def with_objects(self, objects):
        self.objects = list(objects)
        return self

    def with_physics(self, physics):
        """SceneBuilder::with_physics (scene.rs:79-83): a physics.World."""
        self.physics = physics
        return self

    def build(self):
        models, index = [], {}
        for o in self.objects:
            mi = getattr(o, "instance", None)
            if mi is None:
                raise ValueError("scene objects must come from SceneObjectBuilder (they name their model)")
            if id(mi) not in index:
                index[id(mi)] = len(models)
                models.append(mi)
            o.model = index[id(mi)]
        tlas = TlasBuilder(ctx=self.ctx).build_for(self.objects, [m.bvh for m in models])
        scene = Scene(tlas, self.objects, self.camera, self.physics)
        scene.models = models
        return scene


class Triangle:
    """Triangle (geometry/triangle.rs:9-108).  intersect() tests triangle
    i against ray i on the GPU for arrays of triangles ((N, 9) float32) and
    rays; a single triangle is tested against every ray."""

    def __init__(self, vertex0, vertex1, vertex2):
        self.tri = np.concatenate([np.asarray(v, np.float32).reshape(3) for v in (vertex0, vertex1, vertex2)])

    def intersect(self, rays, ctx=None):
        """(t, u, v) per ray, or None where it misses (Triangle::intersect)."""
        rays = np.ascontiguousarray(rays, RAY_DTYPE).reshape(-1)
        hits = intersect_pairs(np.tile(self.tri, (rays.shape[0], 1)), rays, ctx)
        return [None if h["prim"] == MISS else (float(h["t"]), float(h["u"]), float(h["v"])) for h in hits]


def intersect_pairs(tris, rays, ctx=None):
    """Triangle::intersect / intersect_mut for n (triangle, ray) pairs on the
    GPU: hit i = {min(ray.t, t), u, v, prim = i} or a miss record."""
    ctx = ctx or default_context()
    tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
    rays = np.ascontiguousarray(rays, RAY_DTYPE).reshape(-1)
    if tris.shape[0] != rays.shape[0]:
        raise ValueError("one triangle per ray")
    n = rays.shape[0]
    if n == 0:
        return np.zeros(0, HIT_DTYPE)
    d_t, d_r, d_h = ctx.to_device(tris), ctx.to_device(rays), ctx.alloc(n * HIT_DTYPE.itemsize)
    L.check(L.lib().bvht_intersect_pairs(ctx.handle, d_t.ptr, d_r.ptr, n, d_h.ptr))
    return d_h.download(HIT_DTYPE, n)
