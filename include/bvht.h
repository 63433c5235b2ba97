/*
 * bvht.h -- C ABI of the MI355X (gfx950) ray/BVH traversal library
 * (libbvht_hip.so).  Drop-in boundary for the traversal hot path of the
 * reference crate lambdaxymox/bvhtracer (Rust, CPU-only).  The reference has
 * no FFI of its own; every entry point below names the reference API it
 * replaces (paths relative to the reference repository root).
 *
 * Conventions
 *   - Every function returns int32_t status: BVHT_OK (0) or a negative code;
 *     bvht_last_error() gives a message.  Nothing panics, aborts or throws
 *     across the ABI (the reference panics on empty meshes and TLASes,
 *     bvhtracer/src/model/bvh.rs:530-535, bvhtracer/src/scene/tlas.rs:248).
 *   - The unit of work is a batch of rays.  Ray and hit buffers passed to
 *     trace entry points are DEVICE pointers (HBM-resident); triangle and
 *     node arrays passed to build/upload/download are HOST pointers.
 *   - All work is enqueued on the context's HIP stream; bvht_sync() or a
 *     read-back is the fence.  One context per host thread.  A tree (BLAS,
 *     TLAS) is only read by traces, so several contexts may trace it at the
 *     same time; each context has its own scratch and traversal-stack spill
 *     buffers, grown on demand (a graph captured on a context keeps their
 *     addresses: size them with one eager call first).  Destroy every
 *     object made with a context (BLAS, TLAS, batch, mesh, world, graph,
 *     device allocation) before the context itself, and a world after the
 *     TLASes bound to it.
 *   - Arithmetic is IEEE binary32 in the reference's operation order, no FMA
 *     contraction, correctly rounded division: hits are bit-identical to the
 *     reference algorithm on the same BVH and rays.
 */
#ifndef BVHT_H
#define BVHT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BVHT_OK 0
#define BVHT_EINVAL (-1)   /* bad argument, empty input, size limit */
#define BVHT_ENOMEM (-2)   /* device or host allocation failed */
#define BVHT_EDEVICE (-3)  /* HIP runtime error */
#define BVHT_ESTACK (-4)   /* traversal stack bound exceeded */

#define BVHT_MISS 0xFFFFFFFFu

/* ray orders produced by bvht_raygen */
#define BVHT_ORDER_ROWS 0  /* index y*W + x */
#define BVHT_ORDER_TILES 1 /* 8x8 tiles in PathTracer order: index ((ty*W/8 + tx)*64 + (y%8)*8 + x%8) */

/* BLAS build modes */
#define BVHT_BUILD_REFERENCE 0 /* BvhBuilder::build_for exactly (bvh.rs:333-541) */
#define BVHT_BUILD_SAH 1       /* binned SAH with empty-initialised bins and a leaf cap */

/* ---- records (byte layouts shared with hosts on both sides) ---------- */

/* 32-byte BLAS node, bvh.rs:129-134: aabb, left_first (first primitive of a
 * leaf / left child of a branch; the right child is left_first + 1), count
 * (leaf iff > 0). */
typedef struct {
    float bmin[3];
    uint32_t left_first;
    float bmax[3];
    uint32_t count;
} bvht_node;

/* 32-byte TLAS node, tlas.rs:41-46: aabb, LeftRightIndex word (leaf iff 0;
 * left child = high 16 bits, right child = low 16 bits), scene-object id. */
typedef struct {
    float bmin[3];
    uint32_t left_right;
    float bmax[3];
    uint32_t blas;
} bvht_tnode;

/* Ray, query/ray.rs:8-35: origin, t (f32::MAX for camera rays), direction.
 * The reciprocal direction is recomputed on device (IEEE 1/d). */
typedef struct {
    float o[3];
    float tmax;
    float d[3];
    float pad;
} bvht_ray;

/* Hit: t,u,v of the accepted SurfaceInteraction (query/intersection.rs:8-26),
 * the true primitive index (position in the build-permuted triangle array)
 * and scene-object index, and the reference's packed InstancePrimitiveIndex
 * word (instance bits always 0 as in the reference, intersection.rs:41-47).
 * Miss: t = FLT_MAX, u = v = 0, prim = inst = packed = BVHT_MISS. */
typedef struct {
    float t, u, v;
    uint32_t prim;
    uint32_t inst;
    uint32_t packed;
} bvht_hit;

/* Scene object (scene_object.rs:11-18): transform M and its cached inverse
 * (column-major 4x4, m[col*4+row], transform_component.rs:7-27) and the
 * BLAS index (into the array passed to bvht_tlas_create). */
typedef struct {
    float m[16];
    float minv[16];
    uint32_t blas;
    uint32_t pad[3];
} bvht_instance;

/* Pinhole camera (camera.rs:994-1010): eye-space viewport corners at the
 * near plane and the eye-to-world matrix (column-major). */
typedef struct {
    float tl[3];
    float tr[3];
    float bl[3];
    float view_inv[16];
} bvht_camera;

/* Per-ray work counters (optional instrumentation of the trace kernels). */
typedef struct {
    uint32_t pairs;       /* BLAS interior nodes expanded (child-pair fetches) */
    uint32_t tests;       /* triangle tests */
    uint32_t tlas_pairs;  /* TLAS interior nodes expanded */
    uint32_t inst_visits; /* TLAS leaves visited (instance transforms) */
} bvht_stats;

typedef struct bvht_ctx_s* bvht_ctx;
typedef struct bvht_blas_s* bvht_blas;
typedef struct bvht_tlas_s* bvht_tlas;
typedef struct bvht_batch_s* bvht_batch;
typedef struct bvht_graph_s* bvht_graph;

/* ---- context ---------------------------------------------------------- */

/* Create a context on HIP device `device` with its own stream.  (A build's
 * copy back to the host runs on one more stream per device, shared by all
 * contexts and made on first use.)  One host thread at a time may use a
 * context. */
int32_t bvht_create(int32_t device, bvht_ctx* out);
/* Destroy the BLASes, TLASes, batches and graphs made on the context first.
 * The context keeps, until this call, its build workspace, its pinned copy
 * staging with three helper threads for the host side of large copies, and
 * the device buffers of destroyed BLASes (up to 1 GiB, reused by the next
 * BLAS built or uploaded on it). */
int32_t bvht_destroy(bvht_ctx ctx);
/* Message for the last failing call on this thread ("" if none). */
const char* bvht_last_error(void);
/* Wait for the context's stream.  BVHT_ESTACK if a trace enqueued since the
 * last fence dropped traversal stack entries at its stack bound (only
 * possible under bvht_set_stack_limit; its hits may be wrong); the status is
 * then cleared. */
int32_t bvht_sync(bvht_ctx ctx);
/* The context's hipStream_t, for callers that order their own work on it. */
void* bvht_stream(bvht_ctx ctx);
/* Device memory helpers for hosts without their own allocator. */
int32_t bvht_device_alloc(bvht_ctx ctx, size_t bytes, void** out);
int32_t bvht_device_free(bvht_ctx ctx, void* ptr);
int32_t bvht_copy_to_device(bvht_ctx ctx, void* dst, const void* src, size_t bytes);
/* Blocking read-back; a fence like bvht_sync, with the same BVHT_ESTACK
 * report (the copy itself has completed when that is returned). */
int32_t bvht_copy_to_host(bvht_ctx ctx, void* dst, const void* src, size_t bytes);

/* HIP events on the context's stream, for timing enqueued work. */
int32_t bvht_event_create(bvht_ctx ctx, void** out);
int32_t bvht_event_record(bvht_ctx ctx, void* ev);
/* Waits for `end`, then *ms = elapsed milliseconds between the two events. */
int32_t bvht_event_elapsed(void* start, void* end, float* ms);
int32_t bvht_event_destroy(void* ev);

/* Plain device-to-device copy of `bytes` (multiple of 16, 16-byte aligned
 * pointers) by a streaming kernel.  Measurement aid with no reference
 * counterpart: gives the achievable HBM copy bandwidth on the running GPU
 * and calibrates byte counters read from hardware profilers. */
int32_t bvht_stream_copy(bvht_ctx ctx, void* d_dst, const void* d_src, size_t bytes);

/* ---- BLAS ------------------------------------------------------------- */

/* BvhBuilder::new().build_for(&mut [Triangle<f32>]) (bvh.rs:509-541) on the
 * GPU.  tris: host float[9*n] (v0,v1,v2 per triangle), permuted IN PLACE into
 * the build order like the reference's slice.  mode BVHT_BUILD_REFERENCE
 * reproduces the reference node array and permutation (max_leaf ignored);
 * BVHT_BUILD_SAH starts bins empty and splits every node with more than
 * max_leaf primitives (median split when no plane separates them). */
int32_t bvht_blas_build(bvht_ctx ctx, float* tris, uint32_t n, int32_t mode, uint32_t max_leaf,
                        bvht_blas* out);
/* Adopt an existing BVH (node array of node_count >= nodes_used nodes, and
 * its triangle array in build order).  Used to trace a CPU-built BVH. */
int32_t bvht_blas_upload(bvht_ctx ctx, const bvht_node* nodes, uint32_t node_count,
                         uint32_t nodes_used, const float* tris, uint32_t n, bvht_blas* out);
/* Copy back nodes[node_count] (2n for built BLASes) and tris[9n]; either may be NULL. */
int32_t bvht_blas_download(bvht_ctx ctx, bvht_blas blas, bvht_node* nodes, float* tris);
int32_t bvht_blas_info(bvht_blas blas, uint32_t* tri_count, uint32_t* node_count,
                       uint32_t* nodes_used, uint32_t* depth);
/* Replace the triangle vertices (host float[9n], build order), as
 * Model::primitives_mut() does before ModelInstance::refit (model.rs:31-33,87-89). */
int32_t bvht_blas_set_vertices(bvht_ctx ctx, bvht_blas blas, const float* tris, uint32_t n);
/* Same from device memory (float[9n]). */
int32_t bvht_blas_set_vertices_device(bvht_ctx ctx, bvht_blas blas, const float* d_tris, uint32_t n);
/* The build's triangle permutation (host uint32[n]): perm[k] = index, in the
 * array passed to bvht_blas_build, of the triangle now at build position k.
 * The reference reorders the caller's array in place and keeps no
 * permutation (bvh.rs:420-429); this is what per-triangle attributes are
 * permuted with.  BLASes adopted by bvht_blas_upload have none (EINVAL). */
int32_t bvht_blas_permutation(bvht_ctx ctx, bvht_blas blas, uint32_t* perm);
/* Bvh::refit (bvh.rs:469-493) on the GPU. */
int32_t bvht_blas_refit(bvht_ctx ctx, bvht_blas blas);
/* Bvh::bounds (bvh.rs:499-501): out[6] = bmin, bmax of node 0 (host). */
int32_t bvht_blas_bounds(bvht_ctx ctx, bvht_blas blas, float* out6);
int32_t bvht_blas_destroy(bvht_blas blas);

/* ---- traversal -------------------------------------------------------- */

/* Triangle::intersect / intersect_mut (geometry/triangle.rs:41-108) for n
 * independent (triangle, ray) pairs: d_hits[i] = {min(ray.t, t), u, v,
 * prim = i} when ray i hits triangle i (d_tris float[9n]), else a miss. */
int32_t bvht_intersect_pairs(bvht_ctx ctx, const float* d_tris, const bvht_ray* d_rays, uint32_t n,
                             bvht_hit* d_hits);

/* Bvh::intersect (bvh.rs:307-309, intersect_subtree :242-305) for a batch:
 * d_rays[n] -> d_hits[n].  d_stats (device, may be NULL) receives per-ray
 * work counters. */
int32_t bvht_trace_blas(bvht_ctx ctx, bvht_blas blas, const bvht_ray* d_rays, uint32_t n,
                        bvht_hit* d_hits, bvht_stats* d_stats);

/* Same, with the ray count read from device memory (*d_count, clamped to
 * max_n) so a frame can be enqueued without a host round trip. */
int32_t bvht_trace_blas_count(bvht_ctx ctx, bvht_blas blas, const bvht_ray* d_rays,
                              const uint32_t* d_count, uint32_t max_n, bvht_hit* d_hits,
                              bvht_stats* d_stats);

/* Trace kernel configuration for this context: workgroup size (64, 128 or
 * 256 lanes) and per-lane LDS stack entries (8, 16, 24 or 32; deeper stacks
 * spill to HBM).  wg = 0 restores the default (64 lanes, LDS depth from the
 * tree depth, at most 24).  Results never depend on the configuration; the autotuner
 * (bvhtracer_amd/autotune.py) picks the fastest. */
int32_t bvht_set_trace_config(bvht_ctx ctx, int32_t wg, int32_t lds_depth);
int32_t bvht_get_trace_config(bvht_ctx ctx, int32_t* wg, int32_t* lds_depth);

/* Upper bound on the traversal stack entries of this context's traces (no
 * reference counterpart; the reference's stack is a growable Vec,
 * bvh.rs:244).  0, the default, sizes the stack from the tree depth, which is
 * always enough.  A smaller bound caps the spill memory a trace allocates; a
 * ray that needs more drops the pushes beyond it (every access stays in
 * bounds) and the next fence returns BVHT_ESTACK. */
int32_t bvht_set_stack_limit(bvht_ctx ctx, uint32_t max_entries);

/* Longest-first wave order for repeated traces (no reference counterpart).
 * With enable != 0, each bvht_trace_blas / bvht_trace_tlas call on this
 * context records the duration of every 64-ray group, and a later call with
 * the same tree, ray buffer and ray count starts the groups longest first
 * (the order is rebuilt after the first such call and then every 8 calls).
 * That pays off when consecutive calls trace nearly the same rays (a fixed
 * camera over a slowly changing scene); rays that change from call to call
 * (bounce rays) gain nothing.  Hits are the same either way.  Off by
 * default; the _count variants and traces that collect statistics are never
 * reordered.  A call that would have to grow the order buffers while a graph
 * captured on the context lives is refused (BVHT_EINVAL). */
int32_t bvht_set_wave_schedule(bvht_ctx ctx, int32_t enable);

/* ---- frame loop ------------------------------------------------------- */

/* Camera::get_ray_world for every pixel (camera.rs:994-1010) with the
 * per-pixel mapping u = x/W, v = y/H of PathTracer::evaluate
 * (renderer.rs:345-368).  order BVHT_ORDER_TILES writes the rays in the
 * PathTracer's 8x8-tile order (renderer.rs:353-361; W and H multiples of 8),
 * so one wave64 traces one tile. */
int32_t bvht_raygen(bvht_ctx ctx, const bvht_camera* cam, uint32_t width, uint32_t height,
                    int32_t order, bvht_ray* d_rays);
/* One cosine-weighted diffuse bounce per hit (no reference counterpart):
 * origin = o + t d + 1e-4 n, n the geometric normal of the hit triangle
 * flipped against d, direction from a PCG32 stream seeded by (seed, ray
 * index).  Misses produce no ray; bounce rays are packed in ray order.
 * *d_count (device uint32) receives the number of bounce rays.  The BLAS
 * provides the triangles; for a TLAS use bvht_bounce_tlas. */
int32_t bvht_bounce(bvht_ctx ctx, bvht_blas blas, const bvht_ray* d_rays, const bvht_hit* d_hits,
                    uint32_t n, uint64_t seed, bvht_ray* d_out, uint32_t* d_count);

/* The same bounce rays grouped by direction (24 bins: dominant axis and its
 * sign x the signs of the other two components), ray order inside a bin, so
 * that a wave's rays take the same child order more often.  d_src[k]
 * (device uint32, may be NULL) receives the index of the ray whose hit
 * spawned bounce ray k.  The rays and their hits are those of bvht_bounce;
 * only the order differs. */
int32_t bvht_bounce_sorted(bvht_ctx ctx, bvht_blas blas, const bvht_ray* d_rays, const bvht_hit* d_hits,
                           uint32_t n, uint64_t seed, bvht_ray* d_out, uint32_t* d_src, uint32_t* d_count);

/* bvht_raygen + bvht_trace_blas + bvht_bounce_sorted for the width x height
 * camera frame in one pass, without storing the camera rays: d_hits[width *
 * height] receives the hits of the frame's rays (in `order`), d_out / d_src /
 * *d_count the bounce rays exactly as bvht_bounce_sorted gives them for
 * those rays and `seed`.  (PathTracer::evaluate's primary pass,
 * renderer.rs:345-368, with the frame's first diffuse bounce generated in
 * the same kernel.)  Staging buffers come from the context's scratch. */
int32_t bvht_trace_camera_bounce(bvht_ctx ctx, bvht_blas blas, const bvht_camera* cam, uint32_t width,
                                 uint32_t height, int32_t order, uint64_t seed, bvht_hit* d_hits, bvht_ray* d_out,
                                 uint32_t* d_src, uint32_t* d_count);

/* bvht_trace_camera_bounce without any device allocation.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory in
 * the call (code under HIP graph capture, serving loops, hosts that forbid
 * hipMalloc).  Same arguments and the same results bit for bit; the bounce
 * rays are made again from the stored hits instead of staged.  Its scratch
 * (per-wave direction-bin counts and scan partials) is a fixed array of the
 * library, 6.0 MiB on each GPU, held until the process ends; nothing spills.
 * Limits, else BVHT_EINVAL: width * height <= BVHT_NOALLOC_MAX_PIXELS, and a
 * BLAS depth of at most 33 (or a stack limit of at most 32 entries, see
 * bvht_set_stack_limit).  Every call on one GPU uses that GPU's one array:
 * calls that overlap on one GPU (on several contexts' streams) race and give
 * wrong bounce rays; calls in order on one stream do not. */
#define BVHT_NOALLOC_MAX_PIXELS 4194304u
int32_t bvht_trace_camera_bounce_noalloc(bvht_ctx ctx, bvht_blas blas, const bvht_camera* cam, uint32_t width,
                                         uint32_t height, int32_t order, uint64_t seed, bvht_hit* d_hits,
                                         bvht_ray* d_out, uint32_t* d_src, uint32_t* d_count);

/* Depth accumulator (renderer.rs:176-194): d_depth[i] = hit t or f32::MAX. */
int32_t bvht_depth(bvht_ctx ctx, const bvht_hit* d_hits, uint32_t n, float* d_depth);

/* Accumulators (renderer.rs:134-235) evaluated from the hits of a frame's
 * rays (d_hits[W*H] in ray order `order`, as produced by bvht_raygen + a
 * trace) and stored at each ray's pixel address y*W + x of the accumulation
 * buffer d_accum (float[3*W*H], AccumulationBuffer, renderer.rs:15-33), as
 * PathTracer::evaluate does (renderer.rs:345-368). */
#define BVHT_ACC_DEPTH 0        /* DepthAccumulator: t (f32::MAX on a miss) in all three channels */
#define BVHT_ACC_INTERSECTION 1 /* IntersectionAccumulator: hit_value or miss_value */
#define BVHT_ACC_UV 2           /* UvMappingAccumulator: (u, v, 1 - (u + v)), 0 on a miss */
#define BVHT_ACC_NORMAL 3       /* NormalMappingAccumulator (bvht_accumulate_blas / _tlas) */
#define BVHT_ACC_TEXTURE 4      /* TextureMaterialAccumulator (bvht_accumulate_blas / _tlas) */
typedef struct {
    float hit_value[3];
    float miss_value[3];
} bvht_accum_params;
int32_t bvht_accumulate(bvht_ctx ctx, const bvht_hit* d_hits, uint32_t width, uint32_t height,
                        int32_t order, int32_t kind, const bvht_accum_params* params, float* d_accum);

/* Shading attributes of a model (Mesh::normals / tex_coords, mesh/mesh.rs;
 * Model::texture, model/model.rs:108-110), host arrays copied to the device,
 * triangles in BUILD order (permute them with bvht_blas_permutation; the
 * reference leaves them in input order, misaligned with the triangles):
 * normals float[9 n] (three corner normals per triangle), tex_coords
 * float[6 n] (three corner (u, v)), texture RGB8 tex_width x tex_height,
 * row-major.  NULL removes that attribute. */
int32_t bvht_blas_set_attributes(bvht_ctx ctx, bvht_blas blas, const float* normals,
                                 const float* tex_coords, const uint8_t* texture, uint32_t tex_width,
                                 uint32_t tex_height);
/* NormalMappingAccumulator (renderer.rs:248-290) and
 * TextureMaterialAccumulator (renderer.rs:293-334, with
 * TextureMaterial::evaluate, materials/material.rs:44-52) for the hits of a
 * frame, stored at each ray's pixel like bvht_accumulate.  kind is
 * BVHT_ACC_NORMAL or BVHT_ACC_TEXTURE.  The hit's true primitive and object
 * indices select the attributes (the reference's packed word has object bits
 * 0, so it reads object 0's model for every hit).  bvht_accumulate_blas: hits
 * of bvht_trace_blas, object transform the identity; bvht_accumulate_tlas:
 * hits of bvht_trace_tlas, normals carried to world space by each object's
 * M.  EINVAL if a model lacks the attribute. */
int32_t bvht_accumulate_blas(bvht_ctx ctx, bvht_blas blas, const bvht_hit* d_hits, uint32_t width,
                             uint32_t height, int32_t order, int32_t kind, float* d_accum);
int32_t bvht_accumulate_tlas(bvht_ctx ctx, bvht_tlas tlas, const bvht_hit* d_hits, uint32_t width,
                             uint32_t height, int32_t order, int32_t kind, float* d_accum);

/* Pixel shaders (renderer.rs:113-222) over an accumulation buffer of
 * n_pixels into the RGBA8 frame buffer d_rgba (uint8[4*n_pixels], row-major
 * like FrameBuffer / TextureBuffer2D, texture_buffer.rs:216-240). */
#define BVHT_SHADE_RADIANCE_TO_RGB 0 /* RadianceToRgbShader */
#define BVHT_SHADE_INTERSECTION 1    /* IntersectionShader: hit_rgba iff the radiance is non-zero */
#define BVHT_SHADE_DEPTH_MAPPING 2   /* DepthMappingShader(scale, offset), wrapping u32 arithmetic */
typedef struct {
    float scale, offset;  /* DepthMappingShader::new(scale, offset) */
    uint8_t hit_rgba[4];  /* IntersectionShader::new(hit_value, miss_value) */
    uint8_t miss_rgba[4];
} bvht_shader_params;
int32_t bvht_shade(bvht_ctx ctx, const float* d_accum, uint32_t n_pixels, int32_t kind,
                   const bvht_shader_params* params, uint8_t* d_rgba);

/* ---- TLAS / instances ------------------------------------------------- */

/* TlasBuilder::new().build_for(&[SceneObject]) (tlas.rs:253-279): objects
 * are set with their transforms (SceneObjectBuilder::with_transform, world
 * bounds from the current BLAS bounds, scene_object.rs:109-126) and the TLAS
 * is built.  blases[] are the models the objects reference by index. */
int32_t bvht_tlas_create(bvht_ctx ctx, const bvht_blas* blases, uint32_t blas_count,
                         const bvht_instance* objects, uint32_t n, bvht_tlas* out);
/* SceneObject::set_transform for objects [first, first+count) (host
 * records; scene_object.rs:60-75): stores M, M^-1 and recomputes the world
 * bounds from the referenced BLAS's current bounds.  Bounds are NOT refreshed
 * by a BLAS refit alone, as in the reference. */
int32_t bvht_tlas_set_objects(bvht_ctx ctx, bvht_tlas tlas, const bvht_instance* objects,
                              uint32_t first, uint32_t count);
/* Copy back objects [first, first+count) as the device holds them (after
 * bvht_scene_run_physics, their M and M^-1 are the bodies'). */
int32_t bvht_tlas_get_objects(bvht_ctx ctx, bvht_tlas tlas, bvht_instance* out, uint32_t first, uint32_t count);
/* Device-side variant: refresh every object's world bounds from its current
 * M and the current BLAS bounds (set_transform with the stored transform). */
int32_t bvht_tlas_refresh_bounds(bvht_ctx ctx, bvht_tlas tlas);
/* Tlas::rebuild / Scene::rebuild (tlas.rs:204-250, scene.rs:36-38) on the GPU. */
int32_t bvht_tlas_rebuild(bvht_ctx ctx, bvht_tlas tlas);
/* Same without the host round trip (at most 64 objects): traces size their
 * stack from the bound depth <= object count instead of the measured depth. */
int32_t bvht_tlas_rebuild_async(bvht_ctx ctx, bvht_tlas tlas);
/* Copy back nodes[2n] and nodes_used. */
int32_t bvht_tlas_download(bvht_ctx ctx, bvht_tlas tlas, bvht_tnode* nodes, uint32_t* nodes_used);
int32_t bvht_tlas_destroy(bvht_tlas tlas);
/* Tlas::intersect / Scene::intersect (tlas.rs:123-177, scene.rs:32-34) with
 * SceneObject::intersect at the leaves (scene_object.rs:77-89).  hit.inst is
 * the true object index; hit.packed keeps the reference's instance bits 0. */
int32_t bvht_trace_tlas(bvht_ctx ctx, bvht_tlas tlas, const bvht_ray* d_rays, uint32_t n,
                        bvht_hit* d_hits, bvht_stats* d_stats);
/* Same, with the ray count read from device memory (clamped to max_n). */
int32_t bvht_trace_tlas_count(bvht_ctx ctx, bvht_tlas tlas, const bvht_ray* d_rays,
                              const uint32_t* d_count, uint32_t max_n, bvht_hit* d_hits,
                              bvht_stats* d_stats);
/* Diffuse bounce rays for TLAS hits (see bvht_bounce); the normal is carried
 * to world space by the transpose of the object's M^-1. */
int32_t bvht_bounce_tlas(bvht_ctx ctx, bvht_tlas tlas, const bvht_ray* d_rays,
                         const bvht_hit* d_hits, uint32_t n, uint64_t seed, bvht_ray* d_out,
                         uint32_t* d_count);

/* bvht_bounce_sorted for TLAS hits. */
int32_t bvht_bounce_tlas_sorted(bvht_ctx ctx, bvht_tlas tlas, const bvht_ray* d_rays,
                                const bvht_hit* d_hits, uint32_t n, uint64_t seed, bvht_ray* d_out,
                                uint32_t* d_src, uint32_t* d_count);

/* ---- batched scenes (no reference counterpart) ------------------------ */

/* A batch updates and traces many scenes (TLASes) with one launch per stage.
 * The scenes must belong to ctx and outlive the batch. */
int32_t bvht_batch_create(bvht_ctx ctx, const bvht_tlas* scenes, uint32_t n_scenes, bvht_batch* out);
int32_t bvht_batch_destroy(bvht_batch batch);
/* Bvh::refit of every distinct BLAS of the batch, one launch. */
int32_t bvht_batch_refit(bvht_ctx ctx, bvht_batch batch);
/* set_transform (bounds from the current BLAS boxes) for every object, then
 * Tlas::rebuild of every scene: two launches, no host round trip. */
int32_t bvht_batch_rebuild(bvht_ctx ctx, bvht_batch batch);
/* Ray layout for bvht_batch_trace: scene s owns rays [ray_offsets[s],
 * ray_offsets[s+1]) of one device buffer (host array of n_scenes + 1
 * entries); work items hold rays_per_item rays (64, 128 or 256). */
int32_t bvht_batch_set_rays(bvht_ctx ctx, bvht_batch batch, const uint32_t* ray_offsets,
                            uint32_t rays_per_item);
/* Persistent-kernel configuration: workgroup (64 or 256), LDS stack depth
 * (16 or 24), rays per work item (64, 128, 256).  Chosen by the autotuner. */
int32_t bvht_batch_config(bvht_ctx ctx, bvht_batch batch, int32_t wg, int32_t lds_depth,
                          int32_t rays_per_item);
/* Tlas::intersect for every scene's rays in one persistent launch fed by an
 * atomic work queue; d_hits[i] is the hit of d_rays[i] in its own scene. */
int32_t bvht_batch_trace(bvht_ctx ctx, bvht_batch batch, const bvht_ray* d_rays, bvht_hit* d_hits);

/* Per-frame deformation for the refit workloads (input synthesis, no
 * reference counterpart): BLAS vertices = originals (device float[9n], build
 * order) with y += amp * sin(kx * x + w * frame), frame read from *d_frame. */
int32_t bvht_deform_wave(bvht_ctx ctx, bvht_blas blas, const float* d_orig, float amp, float kx,
                         float w, const uint32_t* d_frame);
/* *d_value += v on the stream (advances a frame counter inside a graph). */
int32_t bvht_counter_add(bvht_ctx ctx, uint32_t* d_value, uint32_t v);

/* ---- differentiable depth/shading: forward/adjoint pair ----------------
 * No reference counterpart.  A mesh binds an indexed triangle mesh (vertices
 * V float[3 n_vertices], faces F uint32[3 n_faces]) to a BLAS built from its
 * face soup, so that BLAS triangle k has corners V[F[face_of[k]][0..2]].  One
 * object with inverse transform M^-1 (host float[16], column-major) is
 * traced: rays go to model space as o' = M^-1 (o,1), d' = M^-1 (d,0) with the
 * arithmetic of SceneObject::intersect (scene_object.rs:77-89).  Per pixel:
 *   depth = hit t (f32::MAX on a miss),
 *   shade = max(0, (m.L)/|m|), m = (1-u-v) N0 + u N1 + v N2 with N the
 *           area-weighted vertex normals and L a model-space light direction
 *           (0 on a miss).
 * Loss = sum over pixels hit in both the image and the target of
 *   w_depth (depth - depth*)^2 + w_shade (shade - shade*)^2.
 * Visibility changes (silhouettes) are not differentiated, and the vertex
 * normals are held fixed within a gradient evaluation. */
typedef struct bvht_mesh_s* bvht_mesh;

/* Bind V and F (host arrays, copied) to blas.  face_of (host uint32[n_tris]
 * of the BLAS, may be NULL) maps build positions to faces; NULL takes the
 * BLAS's build permutation (the BLAS was built by bvht_blas_build from the
 * soup of F in face order).  Fails with EINVAL unless the BLAS triangles are
 * bitwise the gathered faces.  Computes the vertex normals. */
int32_t bvht_mesh_create(bvht_ctx ctx, bvht_blas blas, const float* vertices, uint32_t n_vertices,
                         const uint32_t* faces, uint32_t n_faces, const uint32_t* face_of, bvht_mesh* out);
int32_t bvht_mesh_destroy(bvht_mesh mesh);
/* Device pointers to the mesh's vertices (float[3 n_vertices], writable by
 * the caller; bvht_mesh_update publishes changes) and vertex normals
 * (float[3 n_vertices]). */
int32_t bvht_mesh_buffers(bvht_mesh mesh, float** d_vertices, float** d_normals);
/* After the vertices changed: gather them into the BLAS triangles, Bvh::refit
 * (bvh.rs:469-493), and, if normals != 0, recompute the vertex normals. */
int32_t bvht_mesh_update(bvht_ctx ctx, bvht_mesh mesh, int32_t normals);
/* Gradient step on the vertices: v -= lr * g_v for each vertex (d_grad:
 * device float[3 n_vertices]); with BVHT_SGD_ALONG_NORMAL only the component
 * along the current vertex normal, lr (g_v.n) n.  Each vertex's step is
 * shortened to max_step if longer (max_step <= 0: unclipped): silhouette
 * pixels carry gradients orders of magnitude above the rest. */
#define BVHT_SGD_ALONG_NORMAL 1
int32_t bvht_mesh_sgd(bvht_ctx ctx, bvht_mesh mesh, const float* d_grad, float lr, float max_step,
                      int32_t flags);
/* Forward: trace d_rays[n] through the object and shade.  d_hits[n] as
 * bvht_trace_tlas with one object (hit.inst = 0); d_image[2n] = {depth,
 * shade} per ray.  Each call records how long every 64-ray group took and
 * the next call with the same n starts the groups longest first; only the
 * order of the work changes, never a result. */
int32_t bvht_trace_forward(bvht_ctx ctx, bvht_mesh mesh, const float* minv, const float* light,
                           const bvht_ray* d_rays, uint32_t n, bvht_hit* d_hits, float* d_image);
/* Adjoint of the forward for the loss above against d_target[2n]: for each
 * hit in d_hits (from bvht_trace_forward with the same mesh state and minv)
 * recomputes t,u,v of the hit triangle with the Moeller-Trumbore arithmetic
 * (in binary64, on the binary32 inputs: grazing rays make the binary32 solve
 * ill-conditioned) and propagates dLoss/d(t,u,v) in closed form to the
 * triangle's vertices (d_grad_vertices, float[3 n_vertices]) and through
 * o', d' to M^-1 (d_grad_minv, float[16] column-major; the bottom row is 0).
 * *d_loss (device double) receives the loss.  accumulate = 0 overwrites the
 * three outputs, else adds to them.  Every sum runs in a fixed order (no
 * floating-point atomics): the same inputs give the same bits on every run.
 * dLoss/dM follows as -(M^-1)^T G (M^-1)^T. */
int32_t bvht_trace_adjoint(bvht_ctx ctx, bvht_mesh mesh, const float* minv, const float* light,
                           const bvht_ray* d_rays, uint32_t n, const bvht_hit* d_hits,
                           const float* d_target, float w_depth, float w_shade, int32_t accumulate,
                           float* d_grad_vertices, float* d_grad_minv, double* d_loss);

/* ---- rigid-body physics driving scene objects ---------------------------- */

/* RigidBody (physics/rigid_body.rs:66-86), every field, binary32.  Matrices
 * column-major: inertia tensors m[c*3+r], transform m[c*4+r]; orientation is
 * the quaternion (s, x, y, z).  RigidBody::default(): inverse_mass FLT_MAX,
 * identity inverse inertia tensors, drags 1, orientation (0, 0, 0, 1)
 * (Quaternion::unit_z), asleep, can_sleep, identity transform, the rest 0. */
typedef struct {
    float inverse_mass;
    float inverse_inertia_tensor[9];
    float linear_drag;
    float angular_drag;
    float position[3];
    float orientation[4];
    float velocity[3];
    float rotation[3];
    float inverse_inertia_tensor_world[9];
    float motion;
    uint32_t is_awake;
    uint32_t can_sleep;
    float transform[16];
    float force_accumulator[3];
    float torque_accumulator[3];
    float acceleration[3];
    float last_frame_acceleration[3];
    float pad[3];
} bvht_body;

#define BVHT_FORCE_NULL 0    /* NullForce (force_generator.rs:9-31) */
#define BVHT_FORCE_GRAVITY 1 /* Gravity (force_generator.rs:34-56): v0 = g */
#define BVHT_FORCE_SPRING 2  /* Spring (force_generator.rs:58-121): v0 = connection point on `body`,
                                v1 = on `other` (body coordinates), spring_constant, rest_length */
/* One ForceRegistry entry (force_registry.rs:31-78): a generator acting on `body`. */
typedef struct {
    uint32_t kind;
    uint32_t body;
    uint32_t other;
    uint32_t pad;
    float v0[3];
    float v1[3];
    float spring_constant;
    float rest_length;
} bvht_force;

typedef struct bvht_world_s* bvht_world;

/* World (physics/world.rs:14-61) on the device: bodies and the force
 * registry in registration order.  Forces acting on one body are applied in
 * registry order, as ForceRegistry::apply_forces does. */
int32_t bvht_world_create(bvht_ctx ctx, const bvht_body* bodies, uint32_t n_bodies, const bvht_force* forces,
                          uint32_t n_forces, bvht_world* out);
int32_t bvht_world_destroy(bvht_ctx ctx, bvht_world world);
/* Replace bodies [first, first+count) / copy them back. */
int32_t bvht_world_set_bodies(bvht_ctx ctx, bvht_world world, const bvht_body* bodies, uint32_t first,
                              uint32_t count);
int32_t bvht_world_get_bodies(bvht_ctx ctx, bvht_world world, bvht_body* out, uint32_t first, uint32_t count);
/* World::start_frame (world.rs:44-49): clear accumulators, derived data. */
int32_t bvht_world_start_frame(bvht_ctx ctx, bvht_world world);
/* World::run_physics (world.rs:57-60): every force of the registry, then
 * RigidBody::integrate (rigid_body.rs:138-203) of every body.  powf is
 * evaluated in binary64 (ln by the atanh series, exp by range reduction and
 * Taylor series) and rounded: exact for powf(1, y), powf(x, 0) and powers of
 * two of integer exponent, within one binary32 ulp otherwise. */
int32_t bvht_world_run_physics(bvht_ctx ctx, bvht_world world, float duration);
/* Attach a world to a TLAS: object i follows body object_body[i]
 * (BVHT_MISS: no body) with its builder transform transform_init[16 i]
 * (SceneObjectBuilder::with_transform, column-major). */
int32_t bvht_tlas_bind_world(bvht_ctx ctx, bvht_tlas tlas, bvht_world world, const uint32_t* object_body,
                             const float* transform_init);
/* Scene::run_physics (scene/scene.rs:40-45): World::run_physics, then
 * SceneObject::update_transform (scene_object.rs:52-58) of every object with
 * a body -- M = body transform x transform_init, M^-1 (binary64 adjugate,
 * rounded), world bounds -- all enqueued on the device.  Scene::run is this
 * followed by bvht_tlas_rebuild or bvht_tlas_rebuild_async. */
int32_t bvht_scene_run_physics(bvht_ctx ctx, bvht_tlas tlas, float elapsed);

/* ---- hipGraph capture of enqueued work ---------------------------------- */

/* Capture the calls made on ctx between begin and end into a graph (calls
 * that synchronise or allocate are not capturable: build, upload, download,
 * bvht_tlas_rebuild; use the batch calls, which only enqueue kernels).  A
 * graph holds the addresses of the context's scratch and spill buffers: from
 * bvht_graph_begin until every graph of the context is destroyed, a call that
 * would have to grow one of them returns BVHT_EINVAL instead of moving it, so
 * make one eager call of each captured kind at its largest size first. */
int32_t bvht_graph_begin(bvht_ctx ctx);
int32_t bvht_graph_end(bvht_ctx ctx, bvht_graph* out);
int32_t bvht_graph_launch(bvht_ctx ctx, bvht_graph graph);
int32_t bvht_graph_destroy(bvht_graph graph);

#ifdef __cplusplus
}
#endif

#endif
