// bvht_device.hpp -- device-side primitives shared by the gfx950 kernels.
//
// Arithmetic follows the reference crate's operation order exactly (built
// with -ffp-contract=off and correctly rounded division, so every binary32
// operation is rounded on its own):
//   dot(a,b)   = (a.x*b.x + a.y*b.y) + a.z*b.z
//   cross(a,b) = (a.y*b.z - a.z*b.y, a.z*b.x - a.x*b.z, a.x*b.y - a.y*b.x)
//   M*(x,y,z,w) row i = ((m[0][i]x + m[1][i]y) + m[2][i]z) + m[3][i]w
// min/max are IEEE minNum/maxNum (fminf/fmaxf -> v_min_f32/v_max_f32).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/bvht.h"

namespace bvht {

constexpr float kFltMax = 3.402823466e+38f;

struct V3 {
    float x, y, z;
};

__device__ __forceinline__ V3 mk3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 sub3(V3 a, V3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 add3(V3 a, V3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 scl3(V3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross3(V3 a, V3 b)
{
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ V3 min3(V3 a, V3 b) { return mk3(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
__device__ __forceinline__ V3 max3(V3 a, V3 b) { return mk3(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); }
__device__ __forceinline__ float comp3(V3 a, int axis) { return axis == 0 ? a.x : (axis == 1 ? a.y : a.z); }

// Ray with reciprocal direction (Ray::new, query/ray.rs:23-31).
struct DRay {
    V3 o, d, r;
    float t;
};

__device__ __forceinline__ DRay make_ray(V3 o, V3 d, float t)
{
    DRay r;
    r.o = o;
    r.d = d;
    r.r = mk3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    r.t = t;
    return r;
}

__device__ __forceinline__ DRay load_ray(const bvht_ray* rays, uint32_t i)
{
    const float4* p = reinterpret_cast<const float4*>(rays + i);
    float4 a = p[0];
    float4 b = p[1];
    return make_ray(mk3(a.x, a.y, a.z), mk3(b.x, b.y, b.z), a.w);
}

// Aabb::intersect slab test (geometry/aabb.rs:106-125).
__device__ __forceinline__ bool slab(float4 lo, float4 hi, const DRay& ray, float t_closest, float& tmin_out)
{
    float tx1 = (lo.x - ray.o.x) * ray.r.x;
    float tx2 = (hi.x - ray.o.x) * ray.r.x;
    float tmin = fminf(tx1, tx2);
    float tmax = fmaxf(tx1, tx2);
    float ty1 = (lo.y - ray.o.y) * ray.r.y;
    float ty2 = (hi.y - ray.o.y) * ray.r.y;
    tmin = fmaxf(tmin, fminf(ty1, ty2));
    tmax = fminf(tmax, fmaxf(ty1, ty2));
    float tz1 = (lo.z - ray.o.z) * ray.r.z;
    float tz2 = (hi.z - ray.o.z) * ray.r.z;
    tmin = fmaxf(tmin, fminf(tz1, tz2));
    tmax = fminf(tmax, fmaxf(tz1, tz2));
    tmin_out = tmin;
    return (tmax >= tmin) && (tmin < t_closest) && (tmax > 0.0f);
}

// The slab test's distance key for a ray whose t is not above FLT_MAX: tmin
// when the box is hit regardless of t (tmax >= tmin, tmax > 0), else
// FLT_MAX.  With t_closest <= FLT_MAX (or NaN) the box passes slab() exactly
// when its key is below t_closest, and between two children the key orders
// a hit child before a missed one as slab()'s tmin-or-FLT_MAX does.
__device__ __forceinline__ float slab_key(float4 lo, float4 hi, const DRay& ray)
{
    float tx1 = (lo.x - ray.o.x) * ray.r.x;
    float tx2 = (hi.x - ray.o.x) * ray.r.x;
    float tmin = fminf(tx1, tx2);
    float tmax = fmaxf(tx1, tx2);
    float ty1 = (lo.y - ray.o.y) * ray.r.y;
    float ty2 = (hi.y - ray.o.y) * ray.r.y;
    tmin = fmaxf(tmin, fminf(ty1, ty2));
    tmax = fminf(tmax, fmaxf(ty1, ty2));
    float tz1 = (lo.z - ray.o.z) * ray.r.z;
    float tz2 = (hi.z - ray.o.z) * ray.r.z;
    tmin = fmaxf(tmin, fminf(tz1, tz2));
    tmax = fminf(tmax, fmaxf(tz1, tz2));
    return (tmax >= tmin && tmax > 0.0f) ? tmin : kFltMax;  // a NaN tmin or tmax gives kFltMax: never NaN
}

// Triangle::intersect, Moeller-Trumbore with the absolute 1e-4 thresholds
// (geometry/triangle.rs:41-72).  Returns t = min(ray.t, t).
__device__ __forceinline__ bool tri_test(const float* __restrict__ tri, const DRay& ray, float& t_out,
                                         float& u_out, float& v_out)
{
    V3 v0 = mk3(tri[0], tri[1], tri[2]);
    V3 v1 = mk3(tri[3], tri[4], tri[5]);
    V3 v2 = mk3(tri[6], tri[7], tri[8]);
    V3 edge1 = sub3(v1, v0);
    V3 edge2 = sub3(v2, v0);
    V3 normal = cross3(ray.d, edge2);
    float area = dot3(edge1, normal);
    const float threshold = 0.0001f;
    if (fabsf(area) < threshold)
        return false;
    float f = 1.0f / area;
    V3 s = sub3(ray.o, v0);
    float u = f * dot3(s, normal);
    if (u < 0.0f || u > 1.0f)
        return false;
    V3 q = cross3(s, edge1);
    float v = f * dot3(ray.d, q);
    if (v < 0.0f || u + v > 1.0f)
        return false;
    float t = f * dot3(edge2, q);
    if (t > threshold) {
        t_out = fminf(ray.t, t);
        u_out = u;
        v_out = v;
        return true;
    }
    return false;
}

// Triangle::centroid (geometry/triangle.rs:33-39).
__device__ __forceinline__ V3 tri_centroid(V3 v0, V3 v1, V3 v2)
{
    const float one = 1.0f;
    const float one_third = one / (one + one + one);
    return scl3(add3(add3(v0, v1), v2), one_third);
}

// Matrix4x4 * Vector4 followed by contract() (transform.rs:219-234); m is
// column-major, m[c*4+r].
__device__ __forceinline__ V3 mat_mul4(const float* m, V3 p, float w)
{
    float r0 = m[0] * p.x + m[4] * p.y + m[8] * p.z + m[12] * w;
    float r1 = m[1] * p.x + m[5] * p.y + m[9] * p.z + m[13] * w;
    float r2 = m[2] * p.x + m[6] * p.y + m[10] * p.z + m[14] * w;
    return mk3(r0, r1, r2);
}

// Same with the 3x4 row block stored row-major {m[0][i], m[1][i], m[2][i], m[3][i]}.
__device__ __forceinline__ V3 rows_mul(float4 r0, float4 r1, float4 r2, V3 p, float w)
{
    return mk3(r0.x * p.x + r0.y * p.y + r0.z * p.z + r0.w * w,
               r1.x * p.x + r1.y * p.y + r1.z * p.z + r1.w * w,
               r2.x * p.x + r2.y * p.y + r2.z * p.z + r2.w * w);
}

__device__ __forceinline__ uint32_t pack_index(uint32_t inst, uint32_t prim)
{
    return ((inst & 0x00000FFFu) << 20) | (prim & 0x000FFFFFu);
}

__device__ __forceinline__ void store_hit(bvht_hit* hits, uint32_t i, bool found, float t, float u, float v,
                                          uint32_t prim, uint32_t inst)
{
    bvht_hit h;
    if (found) {
        h.t = t;
        h.u = u;
        h.v = v;
        h.prim = prim;
        h.inst = inst;
        h.packed = pack_index(0, prim);
    } else {
        h.t = kFltMax;
        h.u = 0.0f;
        h.v = 0.0f;
        h.prim = BVHT_MISS;
        h.inst = BVHT_MISS;
        h.packed = BVHT_MISS;
    }
    uint2* p = reinterpret_cast<uint2*>(hits + i);
    p[0] = make_uint2(__float_as_uint(h.t), __float_as_uint(h.u));
    p[1] = make_uint2(__float_as_uint(h.v), h.prim);
    p[2] = make_uint2(h.inst, h.packed);
}

// PCG32 (O'Neill 2014), XSH-RR output; pcg_seed = pcg32_srandom(seed, seq).
struct Pcg {
    uint64_t state, inc;
};
__device__ __forceinline__ uint32_t pcg_next(Pcg& r)
{
    uint64_t old = r.state;
    r.state = old * 6364136223846793005ULL + r.inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}
__device__ __forceinline__ Pcg pcg_seed(uint64_t seed, uint64_t seq)
{
    Pcg r;
    r.state = 0;
    r.inc = (seq << 1u) | 1u;
    pcg_next(r);
    r.state += seed;
    pcg_next(r);
    return r;
}
__device__ __forceinline__ float pcg_f32(Pcg& r) { return (float)(pcg_next(r) >> 8) * (1.0f / 16777216.0f); }

// Diffuse bounce ray: unnormalised geometric normal `nrm` (flipped against
// the incoming direction d), hit point p, origin p + 1e-4 n, direction
// cosine-distributed about n from PCG32 stream (seed, idx), orthonormal basis
// of Duff et al. 2017.  Input synthesis for the bounce pass; not a reference
// function.
__device__ __forceinline__ void bounce_ray(V3 nrm, V3 p, V3 d, uint64_t seed, uint32_t idx, float4& a, float4& b)
{
    float len = sqrtf(dot3(nrm, nrm));
    nrm = mk3(nrm.x / len, nrm.y / len, nrm.z / len);
    if (dot3(nrm, d) > 0.0f)
        nrm = mk3(-nrm.x, -nrm.y, -nrm.z);
    V3 org = add3(p, scl3(nrm, 1e-4f));
    Pcg rng = pcg_seed(seed, idx);
    float r1 = pcg_f32(rng), r2 = pcg_f32(rng);
    float phi = 6.28318530718f * r1;
    float sr = sqrtf(r2);
    float lx = sr * cosf(phi), ly = sr * sinf(phi), lz = sqrtf(fmaxf(0.0f, 1.0f - r2));
    float sgn = copysignf(1.0f, nrm.z);
    float aa = -1.0f / (sgn + nrm.z);
    float bb = nrm.x * nrm.y * aa;
    V3 tx = mk3(1.0f + sgn * nrm.x * nrm.x * aa, sgn * bb, -sgn * nrm.x);
    V3 ty = mk3(bb, sgn + nrm.y * nrm.y * aa, -nrm.y);
    V3 dir = add3(add3(scl3(tx, lx), scl3(ty, ly)), scl3(nrm, lz));
    float dl = sqrtf(dot3(dir, dir));
    dir = mk3(dir.x / dl, dir.y / dl, dir.z / dl);
    a = make_float4(org.x, org.y, org.z, kFltMax);
    b = make_float4(dir.x, dir.y, dir.z, 0.0f);
}

// Order-preserving float <-> uint mapping for atomic min/max (total order,
// -0 < +0).
__device__ __forceinline__ uint32_t f2key(float f)
{
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key2f(uint32_t k)
{
    uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    return __uint_as_float(u);
}

// ------------------------------------------------------ lane exchanges --
// DPP moves within 16-lane rows (quad_perm, row_shr:n = 0x110 + n, row_shl:n
// = 0x100 + n, row_ror:n = 0x120 + n): lane i of row_shr:n reads lane i - n;
// a lane whose source lies outside its row reads 0 (bound_ctrl).  With no
// "old" value to keep, the move needs no register set up first, and the
// compiler folds it into the VALU op that consumes it (v_add_f32_dpp ...).
// Callers use the zero only where 0 is their identity, or not at all.
// gfx950 has no DPP wave shifts or row broadcasts; across rows the permlane
// swaps and readlane are used.
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, true);
}
// The value of lane `lane ^ J` (J a power of two below 64) without LDS:
// quad permutes for 1 and 2, row shifts for 4, a row rotate for 8, and the
// permlane swaps (v_permlane16/32_swap: odd rows of the first operand with
// even rows of the second; the upper half of the first with the lower half
// of the second) for 16 and 32.
template <uint32_t J>
__device__ __forceinline__ uint32_t xor_lane(uint32_t v, uint32_t lane)
{
    if constexpr (J == 1) {
        return dpp_u<0xB1>(v);  // quad_perm [1, 0, 3, 2]
    } else if constexpr (J == 2) {
        return dpp_u<0x4E>(v);  // quad_perm [2, 3, 0, 1]
    } else if constexpr (J == 4) {
        // each lane takes the one of the two whose source is in its row
        uint32_t down = dpp_u<0x114>(v), up = dpp_u<0x104>(v);
        return (lane & 4u) ? down : up;
    } else if constexpr (J == 8) {
        return dpp_u<0x128>(v);  // row_ror:8
    } else if constexpr (J == 16) {
        auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        return ((lane >> 4) & 1u) ? r[0] : r[1];
    } else {
        static_assert(J == 32, "xor_lane: J must be a power of two below 64");
        auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
        return lane >= 32u ? r[0] : r[1];
    }
}
__device__ __forceinline__ float readlane_f(float v, uint32_t lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), (int)lane));
}

}  // namespace bvht
