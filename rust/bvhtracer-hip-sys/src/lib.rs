//! Raw bindings to libbvht_hip.so: every entry point of include/bvht.h, the
//! C ABI of the gfx950 (MI355X) traversal library that replaces the
//! reference crate's ray/BVH path (bvhtracer/src/lib.rs:16-27).  Record
//! layouts are byte-identical to the C structs; handles are opaque pointers.
//! Each entry point's meaning, and the reference API it replaces, is
//! documented in include/bvht.h.  The safe, crate-shaped API is the
//! bvhtracer-hip crate beside this one.
#![allow(non_camel_case_types)]
use std::os::raw::{c_char, c_void};

pub const BVHT_OK: i32 = 0;
pub const BVHT_EINVAL: i32 = -1;
pub const BVHT_ENOMEM: i32 = -2;
pub const BVHT_EDEVICE: i32 = -3;
pub const BVHT_ESTACK: i32 = -4;
pub const BVHT_MISS: u32 = 0xFFFF_FFFF;
pub const BVHT_ORDER_ROWS: i32 = 0;
pub const BVHT_ORDER_TILES: i32 = 1;
pub const BVHT_BUILD_REFERENCE: i32 = 0;
pub const BVHT_BUILD_SAH: i32 = 1;
pub const BVHT_NOALLOC_MAX_PIXELS: u32 = 4_194_304;
pub const BVHT_ACC_DEPTH: i32 = 0;
pub const BVHT_ACC_INTERSECTION: i32 = 1;
pub const BVHT_ACC_UV: i32 = 2;
pub const BVHT_ACC_NORMAL: i32 = 3;
pub const BVHT_ACC_TEXTURE: i32 = 4;
pub const BVHT_SHADE_RADIANCE_TO_RGB: i32 = 0;
pub const BVHT_SHADE_INTERSECTION: i32 = 1;
pub const BVHT_SHADE_DEPTH_MAPPING: i32 = 2;
pub const BVHT_SGD_ALONG_NORMAL: i32 = 1;
pub const BVHT_FORCE_NULL: u32 = 0;
pub const BVHT_FORCE_GRAVITY: u32 = 1;
pub const BVHT_FORCE_SPRING: u32 = 2;

/// 32-byte BLAS node (bvh.rs:129-134).
#[repr(C)]
#[derive(Copy, Clone, Default, Debug, PartialEq)]
pub struct bvht_node {
    pub bmin: [f32; 3],
    pub left_first: u32,
    pub bmax: [f32; 3],
    pub count: u32,
}
/// 32-byte TLAS node (tlas.rs:41-46); left child in the high 16 bits.
#[repr(C)]
#[derive(Copy, Clone, Default, Debug, PartialEq)]
pub struct bvht_tnode {
    pub bmin: [f32; 3],
    pub left_right: u32,
    pub bmax: [f32; 3],
    pub blas: u32,
}
#[repr(C)]
#[derive(Copy, Clone, Default, Debug, PartialEq)]
pub struct bvht_ray {
    pub o: [f32; 3],
    pub tmax: f32,
    pub d: [f32; 3],
    pub pad: f32,
}
#[repr(C)]
#[derive(Copy, Clone, Default, Debug, PartialEq)]
pub struct bvht_hit {
    pub t: f32,
    pub u: f32,
    pub v: f32,
    pub prim: u32,
    pub inst: u32,
    pub packed: u32,
}
/// Scene object: M and its cached inverse, column-major, and the BLAS index.
#[repr(C)]
#[derive(Copy, Clone, Debug, PartialEq)]
pub struct bvht_instance {
    pub m: [f32; 16],
    pub minv: [f32; 16],
    pub blas: u32,
    pub pad: [u32; 3],
}
#[repr(C)]
#[derive(Copy, Clone, Debug, PartialEq)]
pub struct bvht_camera {
    pub tl: [f32; 3],
    pub tr: [f32; 3],
    pub bl: [f32; 3],
    pub view_inv: [f32; 16],
}
#[repr(C)]
#[derive(Copy, Clone, Default, Debug, PartialEq)]
pub struct bvht_stats {
    pub pairs: u32,
    pub tests: u32,
    pub tlas_pairs: u32,
    pub inst_visits: u32,
}
#[repr(C)]
#[derive(Copy, Clone, Default, Debug, PartialEq)]
pub struct bvht_accum_params {
    pub hit_value: [f32; 3],
    pub miss_value: [f32; 3],
}
#[repr(C)]
#[derive(Copy, Clone, Default, Debug, PartialEq)]
pub struct bvht_shader_params {
    pub scale: f32,
    pub offset: f32,
    pub hit_rgba: [u8; 4],
    pub miss_rgba: [u8; 4],
}
/// RigidBody, field for field (physics/rigid_body.rs:66-86); 272 bytes.
#[repr(C)]
#[derive(Copy, Clone, Debug, PartialEq)]
pub struct bvht_body {
    pub inverse_mass: f32,
    pub inverse_inertia_tensor: [f32; 9],
    pub linear_drag: f32,
    pub angular_drag: f32,
    pub position: [f32; 3],
    pub orientation: [f32; 4],
    pub velocity: [f32; 3],
    pub rotation: [f32; 3],
    pub inverse_inertia_tensor_world: [f32; 9],
    pub motion: f32,
    pub is_awake: u32,
    pub can_sleep: u32,
    pub transform: [f32; 16],
    pub force_accumulator: [f32; 3],
    pub torque_accumulator: [f32; 3],
    pub acceleration: [f32; 3],
    pub last_frame_acceleration: [f32; 3],
    pub pad: [f32; 3],
}
/// One ForceRegistry entry (force_registry.rs:31-78).
#[repr(C)]
#[derive(Copy, Clone, Default, Debug, PartialEq)]
pub struct bvht_force {
    pub kind: u32,
    pub body: u32,
    pub other: u32,
    pub pad: u32,
    pub v0: [f32; 3],
    pub v1: [f32; 3],
    pub spring_constant: f32,
    pub rest_length: f32,
}

pub enum bvht_ctx_s {}
pub type bvht_ctx = *mut bvht_ctx_s;
pub enum bvht_blas_s {}
pub type bvht_blas = *mut bvht_blas_s;
pub enum bvht_tlas_s {}
pub type bvht_tlas = *mut bvht_tlas_s;
pub enum bvht_batch_s {}
pub type bvht_batch = *mut bvht_batch_s;
pub enum bvht_graph_s {}
pub type bvht_graph = *mut bvht_graph_s;
pub enum bvht_mesh_s {}
pub type bvht_mesh = *mut bvht_mesh_s;
pub enum bvht_world_s {}
pub type bvht_world = *mut bvht_world_s;

const _: () = {
    assert!(std::mem::size_of::<bvht_node>() == 32);
    assert!(std::mem::size_of::<bvht_tnode>() == 32);
    assert!(std::mem::size_of::<bvht_ray>() == 32);
    assert!(std::mem::size_of::<bvht_hit>() == 24);
    assert!(std::mem::size_of::<bvht_instance>() == 144);
    assert!(std::mem::size_of::<bvht_camera>() == 100);
    assert!(std::mem::size_of::<bvht_body>() == 272);
    assert!(std::mem::size_of::<bvht_force>() == 48);
};

#[link(name = "bvht_hip")]
extern "C" {
pub fn bvht_create(device: i32, out: *mut bvht_ctx) -> i32;
    pub fn bvht_destroy(ctx: bvht_ctx) -> i32;
    pub fn bvht_last_error() -> *const c_char;
    pub fn bvht_sync(ctx: bvht_ctx) -> i32;
    pub fn bvht_stream(ctx: bvht_ctx) -> *mut c_void;
    pub fn bvht_device_alloc(ctx: bvht_ctx, bytes: usize, out: *mut *mut c_void) -> i32;
    pub fn bvht_device_free(ctx: bvht_ctx, ptr: *mut c_void) -> i32;
    pub fn bvht_copy_to_device(ctx: bvht_ctx, dst: *mut c_void, src: *const c_void, bytes: usize) -> i32;
    pub fn bvht_copy_to_host(ctx: bvht_ctx, dst: *mut c_void, src: *const c_void, bytes: usize) -> i32;
    pub fn bvht_event_create(ctx: bvht_ctx, out: *mut *mut c_void) -> i32;
    pub fn bvht_event_record(ctx: bvht_ctx, ev: *mut c_void) -> i32;
    pub fn bvht_event_elapsed(start: *mut c_void, end: *mut c_void, ms: *mut f32) -> i32;
    pub fn bvht_event_destroy(ev: *mut c_void) -> i32;
    pub fn bvht_stream_copy(ctx: bvht_ctx, d_dst: *mut c_void, d_src: *const c_void, bytes: usize) -> i32;
    pub fn bvht_blas_build(ctx: bvht_ctx, tris: *mut f32, n: u32, mode: i32, max_leaf: u32, out: *mut bvht_blas) -> i32;
    pub fn bvht_blas_upload(ctx: bvht_ctx, nodes: *const bvht_node, node_count: u32, nodes_used: u32,
        tris: *const f32, n: u32, out: *mut bvht_blas) -> i32;
    pub fn bvht_blas_download(ctx: bvht_ctx, blas: bvht_blas, nodes: *mut bvht_node, tris: *mut f32) -> i32;
    pub fn bvht_blas_info(blas: bvht_blas, tri_count: *mut u32, node_count: *mut u32, nodes_used: *mut u32,
        depth: *mut u32) -> i32;
    pub fn bvht_blas_set_vertices(ctx: bvht_ctx, blas: bvht_blas, tris: *const f32, n: u32) -> i32;
    pub fn bvht_blas_set_vertices_device(ctx: bvht_ctx, blas: bvht_blas, d_tris: *const f32, n: u32) -> i32;
    pub fn bvht_blas_permutation(ctx: bvht_ctx, blas: bvht_blas, perm: *mut u32) -> i32;
    pub fn bvht_blas_refit(ctx: bvht_ctx, blas: bvht_blas) -> i32;
    pub fn bvht_blas_bounds(ctx: bvht_ctx, blas: bvht_blas, out6: *mut f32) -> i32;
    pub fn bvht_blas_destroy(blas: bvht_blas) -> i32;
    pub fn bvht_intersect_pairs(ctx: bvht_ctx, d_tris: *const f32, d_rays: *const bvht_ray, n: u32,
        d_hits: *mut bvht_hit) -> i32;
    pub fn bvht_trace_blas(ctx: bvht_ctx, blas: bvht_blas, d_rays: *const bvht_ray, n: u32, d_hits: *mut bvht_hit,
        d_stats: *mut bvht_stats) -> i32;
    pub fn bvht_trace_blas_count(ctx: bvht_ctx, blas: bvht_blas, d_rays: *const bvht_ray, d_count: *const u32,
        max_n: u32, d_hits: *mut bvht_hit, d_stats: *mut bvht_stats) -> i32;
    pub fn bvht_set_trace_config(ctx: bvht_ctx, wg: i32, lds_depth: i32) -> i32;
    pub fn bvht_get_trace_config(ctx: bvht_ctx, wg: *mut i32, lds_depth: *mut i32) -> i32;
    pub fn bvht_set_stack_limit(ctx: bvht_ctx, max_entries: u32) -> i32;
    pub fn bvht_set_wave_schedule(ctx: bvht_ctx, enable: i32) -> i32;
    pub fn bvht_raygen(ctx: bvht_ctx, cam: *const bvht_camera, width: u32, height: u32, order: i32,
        d_rays: *mut bvht_ray) -> i32;
    pub fn bvht_bounce(ctx: bvht_ctx, blas: bvht_blas, d_rays: *const bvht_ray, d_hits: *const bvht_hit, n: u32,
        seed: u64, d_out: *mut bvht_ray, d_count: *mut u32) -> i32;
    pub fn bvht_bounce_sorted(ctx: bvht_ctx, blas: bvht_blas, d_rays: *const bvht_ray, d_hits: *const bvht_hit,
        n: u32, seed: u64, d_out: *mut bvht_ray, d_src: *mut u32, d_count: *mut u32) -> i32;
    pub fn bvht_trace_camera_bounce(ctx: bvht_ctx, blas: bvht_blas, cam: *const bvht_camera, width: u32,
        height: u32, order: i32, seed: u64, d_hits: *mut bvht_hit, d_out: *mut bvht_ray, d_src: *mut u32,
        d_count: *mut u32) -> i32;
    pub fn bvht_trace_camera_bounce_noalloc(ctx: bvht_ctx, blas: bvht_blas, cam: *const bvht_camera, width: u32,
        height: u32, order: i32, seed: u64, d_hits: *mut bvht_hit, d_out: *mut bvht_ray, d_src: *mut u32,
        d_count: *mut u32) -> i32;
    pub fn bvht_depth(ctx: bvht_ctx, d_hits: *const bvht_hit, n: u32, d_depth: *mut f32) -> i32;
    pub fn bvht_accumulate(ctx: bvht_ctx, d_hits: *const bvht_hit, width: u32, height: u32, order: i32, kind: i32,
        params: *const bvht_accum_params, d_accum: *mut f32) -> i32;
    pub fn bvht_blas_set_attributes(ctx: bvht_ctx, blas: bvht_blas, normals: *const f32, tex_coords: *const f32,
        texture: *const u8, tex_width: u32, tex_height: u32) -> i32;
    pub fn bvht_accumulate_blas(ctx: bvht_ctx, blas: bvht_blas, d_hits: *const bvht_hit, width: u32, height: u32,
        order: i32, kind: i32, d_accum: *mut f32) -> i32;
    pub fn bvht_accumulate_tlas(ctx: bvht_ctx, tlas: bvht_tlas, d_hits: *const bvht_hit, width: u32, height: u32,
        order: i32, kind: i32, d_accum: *mut f32) -> i32;
    pub fn bvht_shade(ctx: bvht_ctx, d_accum: *const f32, n_pixels: u32, kind: i32,
        params: *const bvht_shader_params, d_rgba: *mut u8) -> i32;
    pub fn bvht_tlas_create(ctx: bvht_ctx, blases: *const bvht_blas, blas_count: u32,
        objects: *const bvht_instance, n: u32, out: *mut bvht_tlas) -> i32;
    pub fn bvht_tlas_set_objects(ctx: bvht_ctx, tlas: bvht_tlas, objects: *const bvht_instance, first: u32,
        count: u32) -> i32;
    pub fn bvht_tlas_get_objects(ctx: bvht_ctx, tlas: bvht_tlas, out: *mut bvht_instance, first: u32, count: u32) -> i32;
    pub fn bvht_tlas_refresh_bounds(ctx: bvht_ctx, tlas: bvht_tlas) -> i32;
    pub fn bvht_tlas_rebuild(ctx: bvht_ctx, tlas: bvht_tlas) -> i32;
    pub fn bvht_tlas_rebuild_async(ctx: bvht_ctx, tlas: bvht_tlas) -> i32;
    pub fn bvht_tlas_download(ctx: bvht_ctx, tlas: bvht_tlas, nodes: *mut bvht_tnode, nodes_used: *mut u32) -> i32;
    pub fn bvht_tlas_destroy(tlas: bvht_tlas) -> i32;
    pub fn bvht_trace_tlas(ctx: bvht_ctx, tlas: bvht_tlas, d_rays: *const bvht_ray, n: u32, d_hits: *mut bvht_hit,
        d_stats: *mut bvht_stats) -> i32;
    pub fn bvht_trace_tlas_count(ctx: bvht_ctx, tlas: bvht_tlas, d_rays: *const bvht_ray, d_count: *const u32,
        max_n: u32, d_hits: *mut bvht_hit, d_stats: *mut bvht_stats) -> i32;
    pub fn bvht_bounce_tlas(ctx: bvht_ctx, tlas: bvht_tlas, d_rays: *const bvht_ray, d_hits: *const bvht_hit,
        n: u32, seed: u64, d_out: *mut bvht_ray, d_count: *mut u32) -> i32;
    pub fn bvht_bounce_tlas_sorted(ctx: bvht_ctx, tlas: bvht_tlas, d_rays: *const bvht_ray,
        d_hits: *const bvht_hit, n: u32, seed: u64, d_out: *mut bvht_ray, d_src: *mut u32, d_count: *mut u32) -> i32;
    pub fn bvht_batch_create(ctx: bvht_ctx, scenes: *const bvht_tlas, n_scenes: u32, out: *mut bvht_batch) -> i32;
    pub fn bvht_batch_destroy(batch: bvht_batch) -> i32;
    pub fn bvht_batch_refit(ctx: bvht_ctx, batch: bvht_batch) -> i32;
    pub fn bvht_batch_rebuild(ctx: bvht_ctx, batch: bvht_batch) -> i32;
    pub fn bvht_batch_set_rays(ctx: bvht_ctx, batch: bvht_batch, ray_offsets: *const u32, rays_per_item: u32) -> i32;
    pub fn bvht_batch_config(ctx: bvht_ctx, batch: bvht_batch, wg: i32, lds_depth: i32, rays_per_item: i32) -> i32;
    pub fn bvht_batch_trace(ctx: bvht_ctx, batch: bvht_batch, d_rays: *const bvht_ray, d_hits: *mut bvht_hit) -> i32;
    pub fn bvht_deform_wave(ctx: bvht_ctx, blas: bvht_blas, d_orig: *const f32, amp: f32, kx: f32, w: f32,
        d_frame: *const u32) -> i32;
    pub fn bvht_counter_add(ctx: bvht_ctx, d_value: *mut u32, v: u32) -> i32;
    pub fn bvht_mesh_create(ctx: bvht_ctx, blas: bvht_blas, vertices: *const f32, n_vertices: u32,
        faces: *const u32, n_faces: u32, face_of: *const u32, out: *mut bvht_mesh) -> i32;
    pub fn bvht_mesh_destroy(mesh: bvht_mesh) -> i32;
    pub fn bvht_mesh_buffers(mesh: bvht_mesh, d_vertices: *mut *mut f32, d_normals: *mut *mut f32) -> i32;
    pub fn bvht_mesh_update(ctx: bvht_ctx, mesh: bvht_mesh, normals: i32) -> i32;
    pub fn bvht_mesh_sgd(ctx: bvht_ctx, mesh: bvht_mesh, d_grad: *const f32, lr: f32, max_step: f32, flags: i32) -> i32;
    pub fn bvht_trace_forward(ctx: bvht_ctx, mesh: bvht_mesh, minv: *const f32, light: *const f32,
        d_rays: *const bvht_ray, n: u32, d_hits: *mut bvht_hit, d_image: *mut f32) -> i32;
    pub fn bvht_trace_adjoint(ctx: bvht_ctx, mesh: bvht_mesh, minv: *const f32, light: *const f32,
        d_rays: *const bvht_ray, n: u32, d_hits: *const bvht_hit, d_target: *const f32, w_depth: f32, w_shade: f32,
        accumulate: i32, d_grad_vertices: *mut f32, d_grad_minv: *mut f32, d_loss: *mut f64) -> i32;
    pub fn bvht_world_create(ctx: bvht_ctx, bodies: *const bvht_body, n_bodies: u32, forces: *const bvht_force,
        n_forces: u32, out: *mut bvht_world) -> i32;
    pub fn bvht_world_destroy(ctx: bvht_ctx, world: bvht_world) -> i32;
    pub fn bvht_world_set_bodies(ctx: bvht_ctx, world: bvht_world, bodies: *const bvht_body, first: u32, count: u32) -> i32;
    pub fn bvht_world_get_bodies(ctx: bvht_ctx, world: bvht_world, out: *mut bvht_body, first: u32, count: u32) -> i32;
    pub fn bvht_world_start_frame(ctx: bvht_ctx, world: bvht_world) -> i32;
    pub fn bvht_world_run_physics(ctx: bvht_ctx, world: bvht_world, duration: f32) -> i32;
    pub fn bvht_tlas_bind_world(ctx: bvht_ctx, tlas: bvht_tlas, world: bvht_world, object_body: *const u32,
        transform_init: *const f32) -> i32;
    pub fn bvht_scene_run_physics(ctx: bvht_ctx, tlas: bvht_tlas, elapsed: f32) -> i32;
    pub fn bvht_graph_begin(ctx: bvht_ctx) -> i32;
    pub fn bvht_graph_end(ctx: bvht_ctx, out: *mut bvht_graph) -> i32;
    pub fn bvht_graph_launch(ctx: bvht_ctx, graph: bvht_graph) -> i32;
    pub fn bvht_graph_destroy(graph: bvht_graph) -> i32;
}
