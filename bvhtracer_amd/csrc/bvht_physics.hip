// bvht_physics.hip -- World / ForceRegistry / RigidBody (bvhtracer/src/physics/)
// and Scene::run_physics (scene/scene.rs:40-45) on the GPU.
//
// One thread per body applies that body's registry entries in registry order
// (ForceRegistry::apply_forces, force_registry.rs:72-77, visits entries in
// order and only touches the entry's own body, so per-body order is the whole
// contract); a second launch integrates every body, because a Spring reads the
// other body's transform as the previous integration left it.  A third
// launch gives every object with a body its new M = body transform x builder
// transform, M^-1 and world bounds, ready for the TLAS rebuild.
//
// Binary32 arithmetic in the reference's statement order, compiled without
// FMA contraction.  cglinalg's own operation order for quaternion products,
// quaternion norms, matrix-vector and matrix-matrix products is written
// left to right here (cglinalg is not vendored).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>
#include <cstring>
#include <vector>

#include "bvht_device.hpp"
#include "bvht_internal.hpp"

struct bvht_world_s {
    bvht_ctx ctx = nullptr;
    uint32_t n_bodies = 0;
    uint32_t n_forces = 0;
    bvht_body* d_bodies = nullptr;
    bvht_force* d_forces = nullptr;  // grouped by body, registry order within a body
    uint32_t* d_force_off = nullptr;  // n_bodies + 1 offsets into d_forces
};

using namespace bvht;

namespace {

inline uint32_t nblk(size_t n, uint32_t wg) { return (uint32_t)((n + wg - 1) / wg); }

// ---- powf in binary64 -------------------------------------------------------
// ln x = e ln 2 + 2 atanh((m - 1) / (m + 1)) with x = m 2^e, m in
// [sqrt(1/2), sqrt(2)); exp z = 2^k exp(r), z = k ln 2 + r, |r| <= ln2/2,
// exp(r) by its Taylor series.  Only IEEE +, -, *, / and exact bit
// operations, so every platform gives the same bits.
constexpr double kLn2Hi = 6.93147180369123816490e-01;  // ln 2, upper 32 bits of its significand
constexpr double kLn2Lo = 1.90821492927058770002e-10;  // ln 2 - kLn2Hi
constexpr double kLn2 = kLn2Hi + kLn2Lo;
constexpr double kSqrtHalf = 0.70710678118654752440;

__device__ double ln_pos(double x)
{
    uint64_t b = (uint64_t)__double_as_longlong(x);
    int e = (int)((b >> 52) & 0x7FF);
    if (e == 0) {  // subnormal: scale by 2^54
        x *= 18014398509481984.0;
        b = (uint64_t)__double_as_longlong(x);
        e = (int)((b >> 52) & 0x7FF) - 54;
    }
    e -= 1022;  // x = m 2^e, m in [0.5, 1)
    double m = __longlong_as_double((long long)((b & 0x000FFFFFFFFFFFFFull) | 0x3FE0000000000000ull));
    if (m < kSqrtHalf) {
        m = m * 2.0;
        e -= 1;
    }
    double s = (m - 1.0) / (m + 1.0);
    double s2 = s * s;
    double term = s;
    double sum = s;
    for (int k = 3; k <= 41; k += 2) {
        term = term * s2;
        sum = sum + term / (double)k;
    }
    return (double)e * kLn2 + 2.0 * sum;
}

__device__ double exp_small(double z)  // |z| < 700
{
    double kf = z / kLn2;
    double k = kf < 0.0 ? -(double)(long long)(0.5 - kf) : (double)(long long)(kf + 0.5);
    double r = (z - k * kLn2Hi) - k * kLn2Lo;
    double p = 1.0;
    for (int i = 24; i >= 1; --i)
        p = 1.0 + (r / (double)i) * p;
    long long ki = (long long)k;
    // 2^ki by exponent-field construction, in two steps to stay normal
    long long k1 = ki / 2, k2 = ki - k1;
    double f1 = __longlong_as_double((long long)((uint64_t)(k1 + 1023) << 52));
    double f2 = __longlong_as_double((long long)((uint64_t)(k2 + 1023) << 52));
    return (p * f1) * f2;
}

// f32::powf for the cases the integrator meets: x >= 0.  x < 0 gives NaN.
__device__ float powf_ref(float x, float y)
{
    if (y == 0.0f || x == 1.0f)
        return 1.0f;
    if (x != x || y != y || x < 0.0f)
        return __int_as_float(0x7FC00000);
    if (x == 0.0f)
        return y > 0.0f ? 0.0f : __int_as_float(0x7F800000);
    if (x == __int_as_float(0x7F800000))
        return y > 0.0f ? x : 0.0f;
    if (y == __int_as_float(0x7F800000) || y == -__int_as_float(0x7F800000))
        return ((x > 1.0f) == (y > 0.0f)) ? __int_as_float(0x7F800000) : 0.0f;
    double z = (double)y * ln_pos((double)x);
    if (z > 700.0)
        return __int_as_float(0x7F800000);
    if (z < -800.0)
        return 0.0f;
    return (float)exp_small(z);
}

// ---- RigidBody ---------------------------------------------------------------
constexpr float kSleepThreshold = 0.3f;  // RigidBody::sleep_threshold (rigid_body.rs:115-118)

// Transform3::transform_point (transform.rs:219-223): M (p, 1), contracted.
__device__ void transform_point(const float* m, const float* p, float* out)
{
    for (int r = 0; r < 3; ++r)
        out[r] = m[0 * 4 + r] * p[0] + m[1 * 4 + r] * p[1] + m[2 * 4 + r] * p[2] + m[3 * 4 + r] * 1.0f;
}

// RigidBody::calculate_derived_data (rigid_body.rs:120-135): normalise the
// orientation, Transform3::from_translation_rotation (transform.rs:190-217),
// and _transform_inertia_tensor (rigid_body.rs:28-64) into world space.
__device__ void calculate_derived_data(bvht_body& b)
{
    float* q = b.orientation;
    float mag = sqrtf(q[0] * q[0] + ((q[1] * q[1] + q[2] * q[2]) + q[3] * q[3]));
    q[0] = q[0] / mag;
    q[1] = q[1] / mag;
    q[2] = q[2] / mag;
    q[3] = q[3] / mag;
    const float s = q[0], x = q[1], y = q[2], z = q[3];
    const float one = 1.0f, two = one + one;
    float* m = b.transform;
    m[0] = one - two * y * y - two * z * z;
    m[1] = two * x * y + two * s * z;
    m[2] = two * x * z - two * s * y;
    m[3] = 0.0f;
    m[4] = two * x * y - two * s * z;
    m[5] = one - two * x * x - two * z * z;
    m[6] = two * y * z + two * s * x;
    m[7] = 0.0f;
    m[8] = two * x * z + two * s * y;
    m[9] = two * y * z - two * s * x;
    m[10] = one - two * x * x - two * y * y;
    m[11] = 0.0f;
    m[12] = b.position[0];
    m[13] = b.position[1];
    m[14] = b.position[2];
    m[15] = one;
    // rot_mat[c][r] = m[c*4+r], iit_body[c][r] = it[c*3+r]
    const float* it = b.inverse_inertia_tensor;
#define R(c, r) m[(c) * 4 + (r)]
#define I(c, r) it[(c) * 3 + (r)]
    float m00 = R(0, 0) * I(0, 0) + R(1, 0) * I(0, 1) + R(2, 0) * I(0, 2);
    float m10 = R(0, 0) * I(1, 0) + R(1, 0) * I(1, 1) + R(2, 0) * I(1, 2);
    float m20 = R(0, 0) * I(2, 0) + R(1, 0) * I(2, 1) + R(2, 0) * I(2, 2);
    float m01 = R(0, 1) * I(0, 0) + R(1, 1) * I(0, 1) + R(2, 1) * I(0, 2);
    float m11 = R(0, 1) * I(1, 0) + R(1, 1) * I(1, 1) + R(2, 1) * I(1, 2);
    float m21 = R(0, 1) * I(2, 0) + R(1, 1) * I(2, 1) + R(2, 1) * I(2, 2);
    float m02 = R(0, 2) * I(0, 0) + R(1, 2) * I(0, 1) + R(2, 2) * I(0, 2);
    float m12 = R(0, 2) * I(1, 0) + R(1, 2) * I(1, 1) + R(2, 2) * I(1, 2);
    float m22 = R(0, 2) * I(2, 0) + R(1, 2) * I(2, 1) + R(2, 2) * I(2, 2);
    float* w = b.inverse_inertia_tensor_world;  // w[c*3+r] = iit_world[c][r]
    w[0 * 3 + 0] = m00 * R(0, 0) + m10 * R(1, 0) + m20 * R(2, 0);
    w[1 * 3 + 0] = m00 * R(0, 1) + m10 * R(1, 1) + m20 * R(2, 1);
    w[2 * 3 + 0] = m00 * R(0, 2) + m10 * R(1, 2) + m20 * R(2, 2);
    w[0 * 3 + 1] = m01 * R(0, 0) + m11 * R(1, 0) + m21 * R(2, 0);
    w[1 * 3 + 1] = m01 * R(0, 1) + m11 * R(1, 1) + m21 * R(2, 1);
    w[2 * 3 + 1] = m01 * R(0, 2) + m11 * R(1, 2) + m21 * R(2, 2);
    w[0 * 3 + 2] = m02 * R(0, 0) + m12 * R(1, 0) + m22 * R(2, 0);
    w[1 * 3 + 2] = m02 * R(0, 1) + m12 * R(1, 1) + m22 * R(2, 1);
    w[2 * 3 + 2] = m02 * R(0, 2) + m12 * R(1, 2) + m22 * R(2, 2);
#undef R
#undef I
}

__device__ void clear_accumulators(bvht_body& b)
{
    for (int k = 0; k < 3; ++k) {
        b.force_accumulator[k] = 0.0f;
        b.torque_accumulator[k] = 0.0f;
    }
}

// RigidBody::set_awake(false) (rigid_body.rs:393-403)
__device__ void fall_asleep(bvht_body& b)
{
    b.is_awake = 0u;
    for (int k = 0; k < 3; ++k) {
        b.velocity[k] = 0.0f;
        b.rotation[k] = 0.0f;
    }
}

// RigidBody::integrate (rigid_body.rs:138-203), Newton-Euler.
__device__ void integrate(bvht_body& b, float duration)
{
    if (!b.is_awake)
        return;
    float ang[3];
    for (int k = 0; k < 3; ++k)
        b.last_frame_acceleration[k] = b.acceleration[k] + b.force_accumulator[k] * b.inverse_mass;
    const float* w = b.inverse_inertia_tensor_world;
    const float* t = b.torque_accumulator;
    for (int r = 0; r < 3; ++r)
        ang[r] = w[0 * 3 + r] * t[0] + w[1 * 3 + r] * t[1] + w[2 * 3 + r] * t[2];
    for (int k = 0; k < 3; ++k)
        b.velocity[k] = b.velocity[k] + b.last_frame_acceleration[k] * duration;
    for (int k = 0; k < 3; ++k)
        b.rotation[k] = b.rotation[k] + ang[k] * duration;
    float lin = powf_ref(b.linear_drag, duration);
    float angd = powf_ref(b.angular_drag, duration);
    for (int k = 0; k < 3; ++k)
        b.velocity[k] = b.velocity[k] * lin;
    for (int k = 0; k < 3; ++k)
        b.rotation[k] = b.rotation[k] * angd;
    for (int k = 0; k < 3; ++k)
        b.position[k] = b.position[k] + b.velocity[k] * duration;
    // q = Quaternion::from_pure(rotation * duration) * orientation
    const float ps = 0.0f, px = b.rotation[0] * duration, py = b.rotation[1] * duration, pz = b.rotation[2] * duration;
    const float os = b.orientation[0], ox = b.orientation[1], oy = b.orientation[2], oz = b.orientation[3];
    float qs = ps * os - px * ox - py * oy - pz * oz;
    float qx = ps * ox + px * os + py * oz - pz * oy;
    float qy = ps * oy + py * os + pz * ox - px * oz;
    float qz = ps * oz + pz * os + px * oy - py * ox;
    const float one_half = 0.5f;
    b.orientation[0] = os + qs * one_half;
    b.orientation[1] = ox + qx * one_half;
    b.orientation[2] = oy + qy * one_half;
    b.orientation[3] = oz + qz * one_half;
    calculate_derived_data(b);
    clear_accumulators(b);
    if (b.can_sleep) {
        const float* v = b.velocity;
        const float* o = b.rotation;
        float current_motion = ((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]) + ((o[0] * o[0] + o[1] * o[1]) + o[2] * o[2]);
        float bias = powf_ref(one_half, duration);
        b.motion = bias * b.motion + (1.0f - bias) * current_motion;
        const float ten = 10.0f;
        if (b.motion < kSleepThreshold)
            fall_asleep(b);
        else if (b.motion > ten * kSleepThreshold)
            b.motion = ten * kSleepThreshold;
    }
}

// The part of a body the force generators read and write.
struct ForceState {
    float inverse_mass;
    float position[3];
    float transform[16];
    float force[3];
    float torque[3];
    uint32_t awake;
};

// Gravity::apply_force (force_generator.rs:50-55) and RigidBody::apply_force
// (rigid_body.rs:427-431).
__device__ void apply_gravity(ForceState& b, const bvht_force& f)
{
    if (!(b.inverse_mass >= 0.0f))
        return;
    float mass = b.inverse_mass == 0.0f ? FLT_MAX : 1.0f / b.inverse_mass;
    for (int k = 0; k < 3; ++k)
        b.force[k] = b.force[k] + f.v0[k] * mass;
    b.awake = 1u;
}

// Spring::apply_force (force_generator.rs:101-120) with
// RigidBody::apply_force_at_point (rigid_body.rs:434-439).  The other body's
// transform is read as the previous integration left it (a spring from a
// body to itself makes the reference's RefCell borrow panic; here it reads
// the body's own transform).
__device__ void apply_spring(ForceState& b, const float* ot, const bvht_force& f)
{
    float lws[3], ows[3], d[3];
    transform_point(b.transform, f.v0, lws);
    transform_point(ot, f.v1, ows);
    for (int k = 0; k < 3; ++k)
        d[k] = lws[k] - ows[k];
    float len = sqrtf((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
    float force_magnitude = f.spring_constant * fabsf(len - f.rest_length);
    float force[3];
    for (int k = 0; k < 3; ++k)
        force[k] = (d[k] / len) * (-force_magnitude);
    float pcm[3];
    for (int k = 0; k < 3; ++k)
        pcm[k] = lws[k] - b.position[k];
    for (int k = 0; k < 3; ++k)
        b.force[k] = b.force[k] + force[k];
    b.torque[0] = b.torque[0] + (pcm[1] * force[2] - pcm[2] * force[1]);
    b.torque[1] = b.torque[1] + (pcm[2] * force[0] - pcm[0] * force[2]);
    b.torque[2] = b.torque[2] + (pcm[0] * force[1] - pcm[1] * force[0]);
    b.awake = 1u;
}

__global__ void k_apply_forces(bvht_body* __restrict__ bodies, const bvht_force* __restrict__ forces,
                               const uint32_t* __restrict__ off, uint32_t n_bodies)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_bodies)
        return;
    uint32_t a = off[i], e = off[i + 1];
    if (a == e)
        return;
    bvht_body* bi = bodies + i;
    ForceState b;
    b.inverse_mass = bi->inverse_mass;
    for (int k = 0; k < 3; ++k) {
        b.position[k] = bi->position[k];
        b.force[k] = bi->force_accumulator[k];
        b.torque[k] = bi->torque_accumulator[k];
    }
    for (int k = 0; k < 16; ++k)
        b.transform[k] = bi->transform[k];
    b.awake = bi->is_awake;
    for (uint32_t j = a; j < e; ++j) {
        const bvht_force f = forces[j];
        if (f.kind == BVHT_FORCE_GRAVITY) {
            apply_gravity(b, f);
        } else if (f.kind == BVHT_FORCE_SPRING) {
            float ot[16];
            const float* src = f.other == i ? b.transform : bodies[f.other].transform;
            for (int k = 0; k < 16; ++k)
                ot[k] = src[k];
            apply_spring(b, ot, f);
        }
    }
    // only the accumulators and the awake flag change
    for (int k = 0; k < 3; ++k) {
        bi->force_accumulator[k] = b.force[k];
        bi->torque_accumulator[k] = b.torque[k];
    }
    bi->is_awake = b.awake;
}

__global__ void k_integrate(bvht_body* __restrict__ bodies, uint32_t n_bodies, float duration)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_bodies)
        return;
    bvht_body b = bodies[i];
    integrate(b, duration);
    bodies[i] = b;
}

// World::start_frame (world.rs:44-49)
__global__ void k_start_frame(bvht_body* __restrict__ bodies, uint32_t n_bodies)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_bodies)
        return;
    bvht_body b = bodies[i];
    clear_accumulators(b);
    calculate_derived_data(b);
    bodies[i] = b;
}

// Inverse of a 4x4 matrix by its adjugate, in binary64, rounded to binary32.
__device__ void inverse4(const float* mf, float* out)
{
    double m[16], inv[16];
    for (int k = 0; k < 16; ++k)
        m[k] = (double)mf[k];
    inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] +
             m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] -
             m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] +
             m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] -
              m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] -
             m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] +
             m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] -
             m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] +
              m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] +
             m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] -
             m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] +
              m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] -
              m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] -
             m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] +
             m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] -
              m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] +
              m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
    for (int k = 0; k < 16; ++k)
        out[k] = (float)(inv[k] / det);
}

// SceneObject::update_transform (scene_object.rs:52-58): M = body transform x
// transform_init (Matrix4x4 product, (A B)[c][r] = sum_k A[k][r] B[c][k]),
// then TransformComponent3::set_transform's cached inverse.
__global__ void k_objects_from_bodies(const bvht_body* __restrict__ bodies, const uint32_t* __restrict__ object_body,
                                      const float* __restrict__ init, uint32_t n_obj, bvht_instance* __restrict__ objs)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_obj)
        return;
    uint32_t j = object_body[i];
    if (j == BVHT_MISS)
        return;
    const float* A = bodies[j].transform;
    const float* B = init + 16 * (size_t)i;
    float M[16], Mi[16];
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r)
            M[c * 4 + r] = A[0 * 4 + r] * B[c * 4 + 0] + A[1 * 4 + r] * B[c * 4 + 1] + A[2 * 4 + r] * B[c * 4 + 2] +
                           A[3 * 4 + r] * B[c * 4 + 3];
    inverse4(M, Mi);
    for (int k = 0; k < 16; ++k) {
        objs[i].m[k] = M[k];
        objs[i].minv[k] = Mi[k];
    }
}

void world_free(bvht_world w)
{
    (void)hipFree(w->d_bodies);
    (void)hipFree(w->d_forces);
    (void)hipFree(w->d_force_off);
    delete w;
}

}  // namespace

int32_t bvht::world_run_physics(bvht_world w, float duration)
{
    hipStream_t st = w->ctx->stream;
    if (w->n_forces)
        hipLaunchKernelGGL(k_apply_forces, dim3(nblk(w->n_bodies, 256)), dim3(256), 0, st, w->d_bodies, w->d_forces,
                           w->d_force_off, w->n_bodies);
    hipLaunchKernelGGL(k_integrate, dim3(nblk(w->n_bodies, 256)), dim3(256), 0, st, w->d_bodies, w->n_bodies,
                       duration);
    BVHT_HIP(hipGetLastError());
    return BVHT_OK;
}

extern "C" {

int32_t bvht_world_create(bvht_ctx ctx, const bvht_body* bodies, uint32_t n_bodies, const bvht_force* forces,
                          uint32_t n_forces, bvht_world* out)
{
    BVHT_REQUIRE(ctx && out, "NULL argument");
    BVHT_REQUIRE(n_bodies > 0 && bodies, "a world needs at least one body");
    BVHT_REQUIRE(n_forces == 0 || forces, "NULL force array");
    *out = nullptr;
    // group the registry by body, keeping registry order within a body
    std::vector<uint32_t> off(n_bodies + 1, 0);
    for (uint32_t j = 0; j < n_forces; ++j) {
        const bvht_force& f = forces[j];
        BVHT_REQUIRE(f.kind <= BVHT_FORCE_SPRING, "force %u: unknown kind %u", j, f.kind);
        BVHT_REQUIRE(f.body < n_bodies, "force %u acts on body %u of %u", j, f.body, n_bodies);
        BVHT_REQUIRE(f.kind != BVHT_FORCE_SPRING || f.other < n_bodies, "spring %u names body %u of %u", j, f.other,
                     n_bodies);
        ++off[f.body + 1];
    }
    for (uint32_t i = 0; i < n_bodies; ++i)
        off[i + 1] += off[i];
    std::vector<bvht_force> grouped(n_forces);
    std::vector<uint32_t> cur(off.begin(), off.end() - 1);
    for (uint32_t j = 0; j < n_forces; ++j)
        grouped[cur[forces[j].body]++] = forces[j];
    bvht_world w = new bvht_world_s;
    w->ctx = ctx;
    w->n_bodies = n_bodies;
    w->n_forces = n_forces;
    if (hipMalloc(&w->d_bodies, sizeof(bvht_body) * n_bodies) != hipSuccess ||
        hipMalloc(&w->d_forces, sizeof(bvht_force) * std::max<uint32_t>(n_forces, 1)) != hipSuccess ||
        hipMalloc(&w->d_force_off, sizeof(uint32_t) * (n_bodies + 1)) != hipSuccess) {
        world_free(w);
        return set_error(BVHT_ENOMEM, "device allocation for a world of %u bodies failed", n_bodies);
    }
    hipStream_t st = ctx->stream;
    hipError_t e = hipMemcpyAsync(w->d_bodies, bodies, sizeof(bvht_body) * n_bodies, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && n_forces)
        e = hipMemcpyAsync(w->d_forces, grouped.data(), sizeof(bvht_force) * n_forces, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(w->d_force_off, off.data(), sizeof(uint32_t) * (n_bodies + 1), hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        world_free(w);
        return set_error(BVHT_EDEVICE, "world upload failed: %s", hipGetErrorString(e));
    }
    *out = w;
    return BVHT_OK;
}

int32_t bvht_world_destroy(bvht_ctx ctx, bvht_world world)
{
    if (!world)
        return BVHT_OK;
    (void)hipStreamSynchronize(world->ctx->stream);
    world_free(world);
    (void)ctx;
    return BVHT_OK;
}

int32_t bvht_world_set_bodies(bvht_ctx ctx, bvht_world world, const bvht_body* bodies, uint32_t first,
                              uint32_t count)
{
    BVHT_REQUIRE(ctx && world && (count == 0 || bodies), "NULL argument");
    BVHT_REQUIRE((uint64_t)first + count <= world->n_bodies, "bodies [%u, +%u) out of range (%u)", first, count,
                 world->n_bodies);
    BVHT_HIP(hipMemcpyAsync(world->d_bodies + first, bodies, sizeof(bvht_body) * count, hipMemcpyHostToDevice,
                            ctx->stream));
    BVHT_HIP(hipStreamSynchronize(ctx->stream));
    return BVHT_OK;
}

int32_t bvht_world_get_bodies(bvht_ctx ctx, bvht_world world, bvht_body* out, uint32_t first, uint32_t count)
{
    BVHT_REQUIRE(ctx && world && (count == 0 || out), "NULL argument");
    BVHT_REQUIRE((uint64_t)first + count <= world->n_bodies, "bodies [%u, +%u) out of range (%u)", first, count,
                 world->n_bodies);
    BVHT_HIP(hipMemcpyAsync(out, world->d_bodies + first, sizeof(bvht_body) * count, hipMemcpyDeviceToHost,
                            ctx->stream));
    BVHT_HIP(hipStreamSynchronize(ctx->stream));
    return BVHT_OK;
}

int32_t bvht_world_start_frame(bvht_ctx ctx, bvht_world world)
{
    BVHT_REQUIRE(ctx && world, "NULL argument");
    hipLaunchKernelGGL(k_start_frame, dim3(nblk(world->n_bodies, 256)), dim3(256), 0, ctx->stream, world->d_bodies,
                       world->n_bodies);
    BVHT_HIP(hipGetLastError());
    return BVHT_OK;
}

int32_t bvht_world_run_physics(bvht_ctx ctx, bvht_world world, float duration)
{
    BVHT_REQUIRE(ctx && world, "NULL argument");
    return world_run_physics(world, duration);
}

int32_t bvht_tlas_bind_world(bvht_ctx ctx, bvht_tlas tlas, bvht_world world, const uint32_t* object_body,
                             const float* transform_init)
{
    BVHT_REQUIRE(ctx && tlas && world && object_body && transform_init, "NULL argument");
    for (uint32_t i = 0; i < tlas->n; ++i)
        BVHT_REQUIRE(object_body[i] == BVHT_MISS || object_body[i] < world->n_bodies,
                     "object %u follows body %u of %u", i, object_body[i], world->n_bodies);
    if (!tlas->d_object_body &&
        (hipMalloc(&tlas->d_object_body, sizeof(uint32_t) * tlas->n) != hipSuccess ||
         hipMalloc(&tlas->d_transform_init, sizeof(float) * 16 * (size_t)tlas->n) != hipSuccess))
        return set_error(BVHT_ENOMEM, "device allocation of the object-body binding failed");
    BVHT_HIP(hipMemcpyAsync(tlas->d_object_body, object_body, sizeof(uint32_t) * tlas->n, hipMemcpyHostToDevice,
                            ctx->stream));
    BVHT_HIP(hipMemcpyAsync(tlas->d_transform_init, transform_init, sizeof(float) * 16 * (size_t)tlas->n,
                            hipMemcpyHostToDevice, ctx->stream));
    BVHT_HIP(hipStreamSynchronize(ctx->stream));
    tlas->world = world;
    return BVHT_OK;
}

int32_t bvht_scene_run_physics(bvht_ctx ctx, bvht_tlas tlas, float elapsed)
{
    BVHT_REQUIRE(ctx && tlas, "NULL argument");
    BVHT_REQUIRE(tlas->world, "no world bound to this TLAS (bvht_tlas_bind_world)");
    int32_t rc = world_run_physics(tlas->world, elapsed);
    if (rc)
        return rc;
    hipLaunchKernelGGL(k_objects_from_bodies, dim3(nblk(tlas->n, 64)), dim3(64), 0, ctx->stream,
                       tlas->world->d_bodies, tlas->d_object_body, tlas->d_transform_init, tlas->n, tlas->d_objects);
    BVHT_HIP(hipGetLastError());
    return tlas_refresh(tlas);
}

}  // extern "C"
