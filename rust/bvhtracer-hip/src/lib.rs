//! The reference crate's traversal API (bvhtracer/src/lib.rs:16-27) over
//! libbvht_hip.so.  Types keep the reference's names and record layouts:
//! `Triangle` is three vertices (36 bytes, geometry/triangle.rs:9-16),
//! `Ray` carries origin, direction, reciprocal direction and t
//! (query/ray.rs:8-35), `Intersection` the accepted ray, (t, u, v) and the
//! packed instance/primitive word (query/intersection.rs).  `intersect` on
//! one ray is a batch of one (correct, launch-bound); frame loops call the
//! `*_batch` methods on device buffers.  Failures the reference panics on
//! (empty mesh, empty TLAS) come back as `Error`; `expect` reproduces the
//! panic.
use bvhtracer_hip_sys as sys;
use std::ffi::CStr;
use std::marker::PhantomData;
use std::os::raw::c_void;
use std::rc::Rc;

#[derive(Debug, Clone, PartialEq)]
pub struct Error {
    pub code: i32,
    pub message: String,
}

pub type Result<T> = std::result::Result<T, Error>;

fn check(rc: i32) -> Result<()> {
    if rc == sys::BVHT_OK {
        return Ok(());
    }
    let message = unsafe { CStr::from_ptr(sys::bvht_last_error()) }.to_string_lossy().into_owned();
    Err(Error { code: rc, message })
}

/// A HIP device and stream (`bvht_create`).  Objects made with a context keep
/// it alive, so it is destroyed after them.
pub struct Context {
    raw: sys::bvht_ctx,
}

impl Context {
    pub fn new(device: i32) -> Result<Rc<Context>> {
        let mut raw = std::ptr::null_mut();
        check(unsafe { sys::bvht_create(device, &mut raw) })?;
        Ok(Rc::new(Context { raw }))
    }
    pub fn sync(&self) -> Result<()> {
        check(unsafe { sys::bvht_sync(self.raw) })
    }
    pub fn raw(&self) -> sys::bvht_ctx {
        self.raw
    }
}

impl Drop for Context {
    fn drop(&mut self) {
        unsafe { sys::bvht_destroy(self.raw) };
    }
}

/// Device memory of `len` records of T.
pub struct DeviceBuffer<T: Copy> {
    ctx: Rc<Context>,
    ptr: *mut c_void,
    len: usize,
    _t: PhantomData<T>,
}

impl<T: Copy> DeviceBuffer<T> {
    pub fn new(ctx: &Rc<Context>, len: usize) -> Result<Self> {
        let mut ptr = std::ptr::null_mut();
        check(unsafe { sys::bvht_device_alloc(ctx.raw, len.max(1) * std::mem::size_of::<T>(), &mut ptr) })?;
        Ok(DeviceBuffer { ctx: ctx.clone(), ptr, len, _t: PhantomData })
    }
    pub fn from_slice(ctx: &Rc<Context>, data: &[T]) -> Result<Self> {
        let b = Self::new(ctx, data.len())?;
        check(unsafe {
            sys::bvht_copy_to_device(ctx.raw, b.ptr, data.as_ptr() as *const c_void, std::mem::size_of_val(data))
        })?;
        Ok(b)
    }
    /// Blocking read-back; BVHT_ESTACK if a trace since the last fence
    /// dropped stack entries (only under a stack limit).
    pub fn to_vec(&self) -> Result<Vec<T>> {
        let mut out = Vec::<T>::with_capacity(self.len);
        check(unsafe {
            sys::bvht_copy_to_host(self.ctx.raw, out.as_mut_ptr() as *mut c_void, self.ptr,
                                   self.len * std::mem::size_of::<T>())
        })?;
        unsafe { out.set_len(self.len) };
        Ok(out)
    }
    pub fn len(&self) -> usize {
        self.len
    }
    pub fn is_empty(&self) -> bool {
        self.len == 0
    }
    pub fn as_ptr(&self) -> *const T {
        self.ptr as *const T
    }
    pub fn as_mut_ptr(&mut self) -> *mut T {
        self.ptr as *mut T
    }
}

impl<T: Copy> Drop for DeviceBuffer<T> {
    fn drop(&mut self) {
        unsafe { sys::bvht_device_free(self.ctx.raw, self.ptr) };
    }
}

pub type Vector3 = [f32; 3];

/// Triangle<f32> (geometry/triangle.rs:9-16): vertices, 36 bytes.
#[repr(C)]
#[derive(Copy, Clone, Debug, PartialEq)]
pub struct Triangle {
    pub vertex0: Vector3,
    pub vertex1: Vector3,
    pub vertex2: Vector3,
}

/// Ray<f32> (query/ray.rs:8-35): the reciprocal direction is IEEE 1/d.
#[derive(Copy, Clone, Debug, PartialEq)]
pub struct Ray {
    pub origin: Vector3,
    pub direction: Vector3,
    pub recip_direction: Vector3,
    pub t: f32,
}

impl Ray {
    pub fn new(origin: Vector3, direction: Vector3, t: f32) -> Ray {
        let recip_direction = [1.0 / direction[0], 1.0 / direction[1], 1.0 / direction[2]];
        Ray { origin, direction, recip_direction, t }
    }
    pub fn from_origin_dir(origin: Vector3, direction: Vector3) -> Ray {
        Ray::new(origin, direction, f32::MAX)
    }
    fn record(&self) -> sys::bvht_ray {
        sys::bvht_ray { o: self.origin, tmax: self.t, d: self.direction, pad: 0.0 }
    }
}

#[derive(Copy, Clone, Debug, PartialEq)]
pub struct SurfaceInteraction {
    pub t: f32,
    pub u: f32,
    pub v: f32,
}

/// InstancePrimitiveIndex (query/intersection.rs:41-67): the reference's
/// packed word, instance bits 0 as the reference reports them; `primitive`
/// and `instance` are the true indices, which the packed word loses beyond
/// 2^20 primitives.
#[derive(Copy, Clone, Debug, PartialEq)]
pub struct InstancePrimitiveIndex {
    pub packed: u32,
    pub primitive: u32,
    pub instance: u32,
}

#[derive(Copy, Clone, Debug, PartialEq)]
pub struct Intersection {
    pub ray: Ray,
    pub interaction: SurfaceInteraction,
    pub instance_primitive: InstancePrimitiveIndex,
}

fn intersection_of(ray: &Ray, hit: &sys::bvht_hit) -> Option<Intersection> {
    if hit.prim == sys::BVHT_MISS {
        return None;
    }
    let mut r = *ray;
    r.t = hit.t;
    Some(Intersection {
        ray: r,
        interaction: SurfaceInteraction { t: hit.t, u: hit.u, v: hit.v },
        instance_primitive: InstancePrimitiveIndex { packed: hit.packed, primitive: hit.prim, instance: hit.inst },
    })
}

/// Bvh (model/bvh.rs:227-233) resident on the device.
pub struct Bvh {
    ctx: Rc<Context>,
    raw: sys::bvht_blas,
}

/// BvhBuilder (bvh.rs:505-541): builds on the device and permutes the
/// caller's triangles in place exactly as the reference does.
pub struct BvhBuilder {
    ctx: Rc<Context>,
}

impl BvhBuilder {
    pub fn new(ctx: &Rc<Context>) -> BvhBuilder {
        BvhBuilder { ctx: ctx.clone() }
    }
    pub fn build_for(self, mesh: &mut [Triangle]) -> Result<Bvh> {
        let mut raw = std::ptr::null_mut();
        check(unsafe {
            sys::bvht_blas_build(self.ctx.raw, mesh.as_mut_ptr() as *mut f32, mesh.len() as u32,
                                 sys::BVHT_BUILD_REFERENCE, 0, &mut raw)
        })?;
        Ok(Bvh { ctx: self.ctx, raw })
    }
}

impl Bvh {
    /// Bvh::intersect (bvh.rs:307-309) for one ray.
    pub fn intersect(&self, _mesh: &[Triangle], ray: &Ray) -> Result<Option<Intersection>> {
        let rays = DeviceBuffer::from_slice(&self.ctx, &[ray.record()])?;
        let mut hits = DeviceBuffer::<sys::bvht_hit>::new(&self.ctx, 1)?;
        self.intersect_batch(&rays, &mut hits)?;
        Ok(intersection_of(ray, &hits.to_vec()?[0]))
    }
    /// The batch form: one launch for every ray of the buffer, enqueued on
    /// the context's stream (a read-back or `Context::sync` is the fence).
    pub fn intersect_batch(&self, rays: &DeviceBuffer<sys::bvht_ray>,
                           hits: &mut DeviceBuffer<sys::bvht_hit>) -> Result<()> {
        assert!(hits.len() >= rays.len());
        check(unsafe {
            sys::bvht_trace_blas(self.ctx.raw, self.raw, rays.as_ptr(), rays.len() as u32, hits.as_mut_ptr(),
                                 std::ptr::null_mut())
        })
    }
    /// Bvh::refit (bvh.rs:469-493) after Model::primitives_mut changed the
    /// vertices (build order).
    pub fn refit(&mut self, mesh: &[Triangle]) -> Result<()> {
        check(unsafe {
            sys::bvht_blas_set_vertices(self.ctx.raw, self.raw, mesh.as_ptr() as *const f32, mesh.len() as u32)
        })?;
        check(unsafe { sys::bvht_blas_refit(self.ctx.raw, self.raw) })
    }
    /// Bvh::bounds (bvh.rs:499-501): (min, max) of the root box.
    pub fn bounds(&self) -> Result<(Vector3, Vector3)> {
        let mut b = [0f32; 6];
        check(unsafe { sys::bvht_blas_bounds(self.ctx.raw, self.raw, b.as_mut_ptr()) })?;
        Ok(([b[0], b[1], b[2]], [b[3], b[4], b[5]]))
    }
    /// Bvh::nodes_used (bvh.rs:313).
    pub fn nodes_used(&self) -> Result<usize> {
        let (mut tris, mut nodes, mut used, mut depth) = (0u32, 0u32, 0u32, 0u32);
        check(unsafe { sys::bvht_blas_info(self.raw, &mut tris, &mut nodes, &mut used, &mut depth) })?;
        Ok(used as usize)
    }
    pub fn raw(&self) -> sys::bvht_blas {
        self.raw
    }
}

impl Drop for Bvh {
    fn drop(&mut self) {
        unsafe { sys::bvht_blas_destroy(self.raw) };
    }
}

/// SceneObject (scene/scene_object.rs:11-18): a model by index, M and its
/// cached inverse (column-major).
#[derive(Copy, Clone, Debug, PartialEq)]
pub struct SceneObject {
    pub model: u32,
    pub transform: [f32; 16],
    pub transform_inv: [f32; 16],
}

impl SceneObject {
    fn record(&self) -> sys::bvht_instance {
        sys::bvht_instance { m: self.transform, minv: self.transform_inv, blas: self.model, pad: [0; 3] }
    }
}

/// Tlas (scene/tlas.rs:111-250) over scene objects; the models outlive it.
pub struct Tlas<'m> {
    ctx: Rc<Context>,
    raw: sys::bvht_tlas,
    _models: PhantomData<&'m Bvh>,
}

pub struct TlasBuilder {
    ctx: Rc<Context>,
}

impl TlasBuilder {
    pub fn new(ctx: &Rc<Context>) -> TlasBuilder {
        TlasBuilder { ctx: ctx.clone() }
    }
    /// TlasBuilder::build_for (tlas.rs:273-279) with the objects' world
    /// bounds from their models (SceneObjectBuilder::with_transform).
    pub fn build_for<'m>(self, objects: &[SceneObject], models: &'m [Bvh]) -> Result<Tlas<'m>> {
        let blases: Vec<sys::bvht_blas> = models.iter().map(|m| m.raw).collect();
        let recs: Vec<sys::bvht_instance> = objects.iter().map(|o| o.record()).collect();
        let mut raw = std::ptr::null_mut();
        check(unsafe {
            sys::bvht_tlas_create(self.ctx.raw, blases.as_ptr(), blases.len() as u32, recs.as_ptr(),
                                  recs.len() as u32, &mut raw)
        })?;
        Ok(Tlas { ctx: self.ctx, raw, _models: PhantomData })
    }
}

impl<'m> Tlas<'m> {
    /// Tlas::intersect / Scene::intersect (tlas.rs:123-177) for one ray.
    pub fn intersect(&self, ray: &Ray) -> Result<Option<Intersection>> {
        let rays = DeviceBuffer::from_slice(&self.ctx, &[ray.record()])?;
        let mut hits = DeviceBuffer::<sys::bvht_hit>::new(&self.ctx, 1)?;
        self.intersect_batch(&rays, &mut hits)?;
        Ok(intersection_of(ray, &hits.to_vec()?[0]))
    }
    pub fn intersect_batch(&self, rays: &DeviceBuffer<sys::bvht_ray>,
                           hits: &mut DeviceBuffer<sys::bvht_hit>) -> Result<()> {
        assert!(hits.len() >= rays.len());
        check(unsafe {
            sys::bvht_trace_tlas(self.ctx.raw, self.raw, rays.as_ptr(), rays.len() as u32, hits.as_mut_ptr(),
                                 std::ptr::null_mut())
        })
    }
    /// SceneObject::set_transform for objects [first, first + n), then
    /// nothing else: as in the reference, the tree changes on `rebuild`.
    pub fn set_objects(&mut self, first: u32, objects: &[SceneObject]) -> Result<()> {
        let recs: Vec<sys::bvht_instance> = objects.iter().map(|o| o.record()).collect();
        check(unsafe {
            sys::bvht_tlas_set_objects(self.ctx.raw, self.raw, recs.as_ptr(), first, recs.len() as u32)
        })
    }
    /// Tlas::rebuild / Scene::rebuild (tlas.rs:204-250).
    pub fn rebuild(&mut self) -> Result<()> {
        check(unsafe { sys::bvht_tlas_rebuild(self.ctx.raw, self.raw) })
    }
}

impl<'m> Drop for Tlas<'m> {
    fn drop(&mut self) {
        unsafe { sys::bvht_tlas_destroy(self.raw) };
    }
}
