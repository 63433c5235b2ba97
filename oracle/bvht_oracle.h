/*
 * bvht_oracle.h -- CPU oracle for the ray/BVH traversal hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * crate's CPU algorithms (lambdaxymox/bvhtracer, Rust), used as the checker
 * for the HIP library and as the CPU baseline that is timed beside it.  It is
 * never linked into, loaded by, or called from the product library.
 *
 * Every function cites the reference file:line it restates.  Arithmetic is
 * single precision in the reference's operation order; the file must be
 * compiled with -ffp-contract=off (no FMA contraction) and without
 * -ffast-math so that results match IEEE binary32 evaluation step by step.
 *
 * Conventions that the reference leaves to the third-party crate cglinalg
 * (unpinned git dependency, bvhtracer/Cargo.toml:20-22):
 *   dot(a,b)   = (a.x*b.x + a.y*b.y) + a.z*b.z
 *   cross(a,b) = (a.y*b.z - a.z*b.y, a.z*b.x - a.x*b.z, a.x*b.y - a.y*b.x)
 *   M*v        = r_i = ((m[0][i]*x + m[1][i]*y) + m[2][i]*z) + m[3][i]*w,
 *                column-major m[col][row], stored as m[col*4+row]
 *   normalize  = v / sqrt(dot(v,v)), component-wise division
 *   component_min/max = IEEE minNum/maxNum (fminf/fmaxf)
 * These are the natural left-to-right forms; the exact-t known-answer tests
 * of the reference (test_triangle_intersection.rs, test_scene_quad.rs) are
 * satisfied by them, which the oracle's own tests check.
 */
#ifndef BVHT_ORACLE_H
#define BVHT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_OK 0
#define ORC_EINVAL (-1)
#define ORC_ENOMEM (-2)

/* 32-byte BLAS node: aabb + primitive_count + left_first union
 * (bvhtracer/src/model/bvh.rs:129-134).  Leaf iff count > 0. */
typedef struct {
    float bmin[3];
    uint32_t left_first;
    float bmax[3];
    uint32_t count;
} orc_node;

/* 32-byte TLAS node: aabb + LeftRightIndex + blas (scene/tlas.rs:41-46). */
typedef struct {
    float bmin[3];
    uint32_t left_right;
    float bmax[3];
    uint32_t blas;
} orc_tnode;

/* Ray record {origin, tmax}{direction, pad}; the reciprocal direction is
 * recomputed by ray_new (query/ray.rs:23-31). */
typedef struct {
    float o[3];
    float tmax;
    float d[3];
    float pad;
} orc_ray;

/* Hit record: t,u,v of the accepted SurfaceInteraction, the true primitive
 * and instance ids, and the reference's packed InstancePrimitiveIndex word
 * (query/intersection.rs:41-47).  Miss: t = FLT_MAX, ids = 0xFFFFFFFF. */
typedef struct {
    float t, u, v;
    uint32_t prim;
    uint32_t inst;
    uint32_t packed;
} orc_hit;

/* Scene object as seen by the TLAS: cached inverse transform (column-major
 * 4x4, transform_component.rs:17-27) and the BLAS it references. */
typedef struct {
    float minv[16];
    uint32_t blas;
    uint32_t pad[3];
} orc_instance;

/* Per-ray work counters (instrumented traversal): interior BLAS nodes
 * expanded (= child-pair fetches), triangle tests, interior TLAS nodes
 * expanded, TLAS leaves visited. */
typedef struct {
    uint32_t pairs, tests, tlas_pairs, inst_visits;
} orc_stats;

/* Pinhole camera: eye-space viewport corners and the eye-to-world matrix
 * (camera/camera.rs:994-1010). */
typedef struct {
    float tl[3];
    float tr[3];
    float bl[3];
    float view_inv[16];
} orc_camera;

/* geometry / query primitives */
void orc_normalize(const float in[3], float out[3]);
void orc_ray_recip(const float d[3], float out[3]);
int orc_aabb_intersect(const float bmin[3], const float bmax[3], const orc_ray* ray, float* tmin);
int orc_triangle_intersect(const float tri[9], const orc_ray* ray, float* t, float* u, float* v);
void orc_triangle_centroid(const float tri[9], float out[3]);
uint32_t orc_pack_index(uint32_t inst, uint32_t prim);

/* BLAS */
int orc_bvh_build(float* tris, uint32_t n, int faithful, uint32_t max_leaf,
                  orc_node* nodes /* 2n */, uint32_t* nodes_used);
int orc_bvh_refit(orc_node* nodes, uint32_t nodes_used, const float* tris);
uint32_t orc_bvh_depth(const orc_node* nodes);
int orc_bvh_intersect(const orc_node* nodes, const float* tris, const orc_ray* ray,
                      orc_hit* hit, orc_stats* st);
int orc_bvh_trace(const orc_node* nodes, const float* tris, const orc_ray* rays, uint64_t n,
                  orc_hit* hits, orc_stats* stats, int nthreads);

/* brute force nearest hit (no BVH), ties[i] = number of triangles at the minimum t */
int orc_brute_trace(const float* tris, uint32_t ntris, const orc_ray* rays, uint64_t n, orc_hit* hits,
                    uint32_t* ties);

/* transforms / instances / TLAS */
void orc_transform_point(const float m[16], const float p[3], float out[3]);
void orc_transform_vector(const float m[16], const float v[3], float out[3]);
void orc_instance_bounds(const float m[16], const float model_bmin[3], const float model_bmax[3],
                         float out_bmin[3], float out_bmax[3]);
int orc_tlas_build(const float* bounds /* n*6 */, uint32_t n, orc_tnode* nodes /* 2n */,
                   uint32_t* nodes_used);
int orc_tlas_intersect(const orc_tnode* tnodes, const orc_instance* inst,
                       const orc_node* const* blas_nodes, const float* const* blas_tris,
                       const orc_ray* ray, orc_hit* hit, orc_stats* st);
int orc_tlas_trace(const orc_tnode* tnodes, const orc_instance* inst,
                   const orc_node* const* blas_nodes, const float* const* blas_tris,
                   const orc_ray* rays, uint64_t n, orc_hit* hits, orc_stats* stats, int nthreads);

/* camera */
void orc_camera_ray(const orc_camera* cam, float u, float v, orc_ray* out);
int orc_raygen(const orc_camera* cam, uint32_t width, uint32_t height, orc_ray* rays);

/* differentiable depth/shading checker (no reference counterpart; restates
 * the library's forward/adjoint pair: binary32 forward, binary64 gradient) */
void orc_mesh_normals(const float* V, uint32_t nv, const uint32_t* F, uint32_t nf, float* N);
int orc_diff_forward(const orc_node* nodes, const float* tris, const uint32_t* vid, const float* normals,
                     const float minv[16], const float light[3], const orc_ray* rays, uint64_t n, orc_hit* hits,
                     float* image, int nthreads);
int orc_diff_grad64(const double* V, uint32_t nv, const uint32_t* vid, uint32_t ntris, const double* N,
                    const double minv[16], const double light[3], const orc_ray* rays, const uint32_t* prims,
                    uint64_t n, const float* target, double w_depth, double w_shade, double* loss, double* grad_v,
                    double* grad_minv, double* image);

/* ---- physics (bvhtracer/src/physics/) ------------------------------------ */
/* RigidBody (rigid_body.rs:66-86): field for field, binary32; matrices
 * column-major, orientation (s, x, y, z).  Byte-identical to bvht_body. */
typedef struct {
    float inverse_mass;
    float inverse_inertia_tensor[9];
    float linear_drag, angular_drag;
    float position[3];
    float orientation[4];
    float velocity[3];
    float rotation[3];
    float inverse_inertia_tensor_world[9];
    float motion;
    uint32_t is_awake, can_sleep;
    float transform[16];
    float force_accumulator[3];
    float torque_accumulator[3];
    float acceleration[3];
    float last_frame_acceleration[3];
    float pad[3];
} orc_body;

/* ForceRegistry entry (force_registry.rs:31-42): kind 0 NullForce, 1 Gravity
 * (v0 = g), 2 Spring (v0 / v1 connection points on body / other, constant,
 * rest length).  Byte-identical to bvht_force. */
typedef struct {
    uint32_t kind, body, other, pad;
    float v0[3], v1[3];
    float spring_constant, rest_length;
} orc_force;

/* f32::powf restated in binary64 (ln by the atanh series, exp by range
 * reduction and Taylor series), rounded to binary32. */
float orc_powf(float x, float y);
/* World::start_frame (world.rs:44-49). */
void orc_world_start_frame(orc_body* bodies, uint32_t n);
/* World::run_physics (world.rs:57-60): the registry in order, then every
 * body's integrate. */
void orc_world_run_physics(orc_body* bodies, uint32_t n, const orc_force* forces, uint32_t nf, float duration);
/* SceneObject::update_transform (scene_object.rs:52-58) for n objects:
 * M = body transform x init, M^-1 by the binary64 adjugate, rounded. */
void orc_objects_from_bodies(const orc_body* bodies, const uint32_t* object_body, const float* init, uint32_t n,
                             float* m_out, float* minv_out);

#ifdef __cplusplus
}
#endif

#endif
