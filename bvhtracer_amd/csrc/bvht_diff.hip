// bvht_diff.hip -- differentiable depth/shading: indexed mesh bound to a
// BLAS, forward trace + shade, and the closed-form Moeller-Trumbore adjoint.
//
// No reference counterpart: the reference has no differentiable path.  The
// forward's intersection arithmetic is the reference's (SceneObject::intersect
// scene_object.rs:77-89 around Bvh::intersect bvh.rs:242-309), so its hits
// equal a one-object TLAS trace bit for bit.
//
// Adjoint.  A hit solves o' + t d' = v0 + u e1 + v e2, i.e. A x = s with
// A = [-d' | e1 | e2], x = (t, u, v), s = o' - v0.  For g = dLoss/dx,
//   lambda = A^-T g = [(e1 x e2) g_t + (e2 x -d') g_u + (-d' x e1) g_v] / det,
// det = (-d').(e1 x e2) = e1.(d' x e2) (the Moeller-Trumbore determinant), and
//   dLoss/do' = lambda,  dLoss/dd' = t lambda,
//   dLoss/dv0 = -(1-u-v) lambda,  dLoss/dv1 = -u lambda,  dLoss/dv2 = -v lambda,
//   dLoss/dM^-1[r][c] = lambda_r o_c + t lambda_r d_c   (o_3 = 1, d_3 = 0).
// Shading s = (m.L)/|m|, m = (1-u-v) N0 + u N1 + v N2, adds
//   g_u += g_s (dS/dm).(N1 - N0),  g_v += g_s (dS/dm).(N2 - N0),
//   dS/dm = (L - s m/|m|) / |m|    (zero where the shade is clamped to 0).
// The forward is binary32 (bitwise the reference arithmetic); the adjoint
// solves each pixel in binary64 and scatters binary32.
#include <algorithm>
#include <type_traits>
#include <vector>

#include "bvht_internal.hpp"
#include "bvht_trace.hpp"

struct bvht_mesh_s {
    bvht_ctx ctx = nullptr;
    bvht_blas blas = nullptr;
    uint32_t n_vertices = 0, n_faces = 0;
    float* d_vertices = nullptr;   // 3 n_vertices
    float* d_normals = nullptr;    // 3 n_vertices
    uint32_t* d_faces = nullptr;   // 3 n_faces
    uint32_t* d_vid = nullptr;     // 3 per BLAS triangle (build order): vertex ids of its corners
    uint32_t* d_vf_off = nullptr;  // n_vertices + 1: incident faces of vertex i are d_vf_idx[off[i], off[i+1])
    uint32_t* d_vf_idx = nullptr;  // 3 n_faces, ascending face index per vertex
    // Forward wave schedule: each forward records every 64-ray group's
    // duration and orders the groups longest first for the next forward
    // with the same ray count (the mesh changes little between steps).
    uint32_t* d_sched = nullptr;   // groups: group run by wave slot k
    uint32_t* d_cost = nullptr;    // groups: duration of each group's wave (10 ns ticks)
    uint32_t* d_bins = nullptr;    // kSchedBins counts, kSchedBins cursors, the order's unit count
    uint32_t sched_cap = 0;        // groups the buffers hold
    uint32_t sched_n = 0;          // ray count d_sched was made for (0: none)
    // Adjoint gather.  Vertex -> (BLAS triangle, corner) entries 3k + c in
    // ascending order; per-triangle corner gradients (9 floats); per-triangle
    // record counts (scanned in place to bucket offsets) and cursors, both
    // left zero by every adjoint for the next one; scan tile sums.
    uint32_t* d_vt_off = nullptr;  // n_vertices + 1
    uint32_t* d_vt_idx = nullptr;  // 3 n_tris
    float* d_tgrad = nullptr;      // 9 n_tris
    uint32_t* d_poff = nullptr;    // n_tris
    uint32_t* d_pcur = nullptr;    // n_tris
    uint32_t* d_psums = nullptr;   // scan tile sums of d_poff
    uint32_t* d_big = nullptr;     // n_tris: triangles whose records k_adj_big sums
    uint32_t* d_nbig = nullptr;    // their count
};

namespace bvht {
namespace {

inline uint32_t nblk(size_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }
inline size_t align256(size_t bytes) { return (bytes + 255) & ~(size_t)255; }

template <uint32_t J>
__device__ __forceinline__ uint32_t xor_key(uint32_t v, uint32_t lane, std::integral_constant<uint32_t, J>)
{
    return xor_lane<J>(v, lane);
}
template <uint32_t J>
__device__ __forceinline__ uint64_t xor_key(uint64_t v, uint32_t lane, std::integral_constant<uint32_t, J>)
{
    uint32_t lo = xor_lane<J>((uint32_t)v, lane), hi = xor_lane<J>((uint32_t)(v >> 32), lane);
    return ((uint64_t)hi << 32) | lo;
}
template <uint32_t J>
__device__ __forceinline__ double xor_double(double v, uint32_t lane)
{
    uint64_t u = (uint64_t)__double_as_longlong(v);
    uint32_t lo = xor_lane<J>((uint32_t)u, lane), hi = xor_lane<J>((uint32_t)(u >> 32), lane);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
// Butterfly sum over the 64 lanes (partners 32, 16, .., 1), every lane
// getting the same total.
__device__ __forceinline__ double wave_sum_d(double v, uint32_t lane)
{
    v += xor_double<32>(v, lane);
    v += xor_double<16>(v, lane);
    v += xor_double<8>(v, lane);
    v += xor_double<4>(v, lane);
    v += xor_double<2>(v, lane);
    v += xor_double<1>(v, lane);
    return v;
}

// 64-lane bitonic sort of one key per lane, ascending in lane order:
// stages K = 2, 4, .., 64, each merging with partner distances K/2 .. 1.
template <class T, uint32_t K, uint32_t J>
__device__ __forceinline__ void bitonic_merge(T& k, uint32_t lane)
{
    T other = xor_key(k, lane, std::integral_constant<uint32_t, J>{});
    bool keep_min = ((lane & J) == 0) == ((lane & K) == 0);
    k = keep_min ? (other < k ? other : k) : (other > k ? other : k);
    if constexpr (J > 1)
        bitonic_merge<T, K, J / 2>(k, lane);
}
template <class T, uint32_t K>
__device__ __forceinline__ void bitonic_sort(T& k, uint32_t lane)
{
    if constexpr (K > 2)
        bitonic_sort<T, K / 2>(k, lane);
    bitonic_merge<T, K, K / 2>(k, lane);
}

// One Hillis-Steele step of an inclusive segmented sum within 16-lane rows
// (f: a run starts at or before this lane within the row).
template <uint32_t D>
__device__ __forceinline__ void seg_step(float (&sv)[9], uint32_t& f, uint32_t lr)
{
    uint32_t tf = dpp_u<0x110 + D>(f);
    float t[9];
#pragma unroll
    for (int c = 0; c < 9; ++c)
        t[c] = dpp_f<0x110 + D>(sv[c]);
    if (lr >= D) {
        if (!f) {
#pragma unroll
            for (int c = 0; c < 9; ++c)
                sv[c] = sv[c] + t[c];
        }
        f |= tf;
    }
}

__device__ __forceinline__ V3 ld3(const float* p) { return mk3(p[0], p[1], p[2]); }

// Lambert term of the barycentric-interpolated vertex normal:
// max(0, m.L/|m|), m = ((1-u)-v) N0 + u N1 + v N2 (0 where m vanishes).
__device__ __forceinline__ float shade_at(const float* __restrict__ normals, const uint32_t* __restrict__ vid,
                                          uint32_t prim, float u, float v, V3 light)
{
    V3 n0 = ld3(normals + 3 * (size_t)vid[3 * (size_t)prim + 0]);
    V3 n1 = ld3(normals + 3 * (size_t)vid[3 * (size_t)prim + 1]);
    V3 n2 = ld3(normals + 3 * (size_t)vid[3 * (size_t)prim + 2]);
    float w = (1.0f - u) - v;
    V3 m = add3(add3(scl3(n0, w), scl3(n1, u)), scl3(n2, v));
    float ml = dot3(m, light);
    float len = sqrtf(dot3(m, m));
    return (ml > 0.0f && len > 0.0f) ? ml / len : 0.0f;
}

struct Rows {
    float4 r0, r1, r2;  // rows 0..2 of M^-1: {m[0][i], m[1][i], m[2][i], m[3][i]}
};

inline Rows rows_of(const float* minv)
{
    Rows r;
    r.r0 = make_float4(minv[0], minv[4], minv[8], minv[12]);
    r.r1 = make_float4(minv[1], minv[5], minv[9], minv[13]);
    r.r2 = make_float4(minv[2], minv[6], minv[10], minv[14]);
    return r;
}

// ------------------------------------------------------------- mesh kernels

__global__ void k_mesh_gather(const float* __restrict__ verts, const uint32_t* __restrict__ vid, uint32_t n,
                              float* __restrict__ tris)
{
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n)
        return;
    float* out = tris + 9 * (size_t)k;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float* p = verts + 3 * (size_t)vid[3 * (size_t)k + c];
        out[3 * c + 0] = p[0];
        out[3 * c + 1] = p[1];
        out[3 * c + 2] = p[2];
    }
}

__global__ void k_mesh_check(const float* __restrict__ verts, const uint32_t* __restrict__ vid, uint32_t n,
                             const float* __restrict__ tris, uint32_t* __restrict__ err)
{
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n)
        return;
    const float* t = tris + 9 * (size_t)k;
    bool ok = true;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float* p = verts + 3 * (size_t)vid[3 * (size_t)k + c];
        for (int j = 0; j < 3; ++j)
            ok = ok && __float_as_uint(p[j]) == __float_as_uint(t[3 * c + j]);
    }
    if (!ok)
        atomicOr(err, 1u);
}

// Area-weighted vertex normals: the cross products (v1-v0) x (v2-v0) of the
// incident faces summed in ascending face order, then normalised (0 if the
// sum vanishes).  Deterministic.
__global__ void k_vertex_normals(const float* __restrict__ verts, const uint32_t* __restrict__ faces,
                                 const uint32_t* __restrict__ off, const uint32_t* __restrict__ idx, uint32_t nv,
                                 float* __restrict__ normals)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv)
        return;
    V3 acc = mk3(0.0f, 0.0f, 0.0f);
    for (uint32_t k = off[i]; k < off[i + 1]; ++k) {
        const uint32_t* f = faces + 3 * (size_t)idx[k];
        V3 a = ld3(verts + 3 * (size_t)f[0]);
        V3 b = ld3(verts + 3 * (size_t)f[1]);
        V3 c = ld3(verts + 3 * (size_t)f[2]);
        acc = add3(acc, cross3(sub3(b, a), sub3(c, a)));
    }
    float len = sqrtf(dot3(acc, acc));
    V3 nrm = len > 0.0f ? mk3(acc.x / len, acc.y / len, acc.z / len) : mk3(0.0f, 0.0f, 0.0f);
    normals[3 * (size_t)i + 0] = nrm.x;
    normals[3 * (size_t)i + 1] = nrm.y;
    normals[3 * (size_t)i + 2] = nrm.z;
}

// Per-vertex gradient step: v -= lr g_v (with along_normal, the component
// of g_v along the vertex normal n: lr (g_v.n) n), the step shortened to
// max_step if longer (max_step <= 0: no limit).
__global__ void k_sgd(float* __restrict__ x, const float* __restrict__ g, const float* __restrict__ normals,
                      uint32_t nv, float lr, float max_step, int along_normal)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv)
        return;
    V3 gv = ld3(g + 3 * (size_t)i);
    if (along_normal) {
        V3 nrm = ld3(normals + 3 * (size_t)i);
        gv = scl3(nrm, dot3(gv, nrm));
    }
    V3 st = scl3(gv, lr);
    if (max_step > 0.0f) {
        float len = sqrtf(dot3(st, st));
        if (len > max_step)
            st = scl3(st, max_step / len);
    }
    float* p = x + 3 * (size_t)i;
    p[0] = p[0] - st.x;
    p[1] = p[1] - st.y;
    p[2] = p[2] - st.z;
}

// ------------------------------------------------------------------ forward

// Each wave runs one 64-ray group: group sched[slot] (slot = the wave's
// position in the grid), or group slot without a schedule.  Lane 0 records
// the wave's duration in cost[group] (s_memrealtime, 10 ns ticks).  Every
// ray is traced exactly as without a schedule; only the order of the groups
// over the launch changes.
template <int WG, int LDS_DEPTH, bool PACKED>
__global__ __launch_bounds__(WG) void k_trace_forward(const float4* __restrict__ nodes, const float* __restrict__ tris,
                                                      const uint32_t* __restrict__ vid,
                                                      const float* __restrict__ normals, Rows rows, V3 light,
                                                      const bvht_ray* __restrict__ rays, uint32_t n,
                                                      bvht_hit* __restrict__ hits, float2* __restrict__ image,
                                                      uint32_t* __restrict__ spill, uint32_t cap,
                                                      uint32_t* __restrict__ status,
                                                      const uint32_t* __restrict__ sched,
                                                      const uint32_t* __restrict__ sched_n,
                                                      uint32_t* __restrict__ cost)
{
    __shared__ uint32_t lds[LDS_DEPTH * WG];
    uint64_t t0 = (uint64_t)wall_clock64();
    uint32_t groups = (n + 63) / 64;
    uint32_t slot = blockIdx.x * (WG / 64) + (threadIdx.x >> 6);
    if (slot >= groups)
        return;
    uint32_t group = (sched && *sched_n == groups) ? sched[slot] : slot;
    uint32_t i = group * 64 + (threadIdx.x & 63);
    if (i >= n)
        return;
    Stack<LDS_DEPTH, WG> st{(lds_u32*)(lds + threadIdx.x), spill + i, (size_t)n, cap, status};
    const float4* rp = reinterpret_cast<const float4*>(rays + i);
    float4 a = rp[0], b = rp[1];
    V3 o = rows_mul(rows.r0, rows.r1, rows.r2, mk3(a.x, a.y, a.z), 1.0f);
    V3 d = rows_mul(rows.r0, rows.r1, rows.r2, mk3(b.x, b.y, b.z), 0.0f);
    DRay ray = make_ray(o, d, a.w);
    uint32_t pairs = 0, tests = 0;
    BlasResult r = __builtin_amdgcn_ballot_w64(ray.t > kFltMax) == 0
                       ? traverse_blas<false, PACKED, true>(nodes, tris, ray, st, 0u, pairs, tests)
                       : traverse_blas<false, PACKED, false>(nodes, tris, ray, st, 0u, pairs, tests);
    store_hit(hits, i, r.found, r.t, r.u, r.v, r.prim, 0);
    float depth = kFltMax, shade = 0.0f;
    if (r.found) {
        depth = r.t;
        shade = shade_at(normals, vid, r.prim, r.u, r.v, light);
    }
    image[i] = make_float2(depth, shade);
    // lane 0 is always a ray of the group (group * 64 < n)
    if ((threadIdx.x & 63) == 0) {
        uint64_t dt = (uint64_t)wall_clock64() - t0;
        cost[group] = dt > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)dt;
    }
}

typedef void (*forward_fn)(const float4*, const float*, const uint32_t*, const float*, Rows, V3, const bvht_ray*,
                           uint32_t, bvht_hit*, float2*, uint32_t*, uint32_t, uint32_t*, const uint32_t*,
                           const uint32_t*, uint32_t*);

forward_fn pick_forward(int wg, int lds, bool packed)
{
#define X(W, D)                                                           \
    if (wg == W && lds == D)                                              \
        return packed ? k_trace_forward<W, D, true> : k_trace_forward<W, D, false>;
    BVHT_TRACE_CONFIGS(X)
#undef X
    return nullptr;
}

// ------------------------------------------------------------------ adjoint

constexpr int kAdjWG = 512;
constexpr int kRed = 13;  // 12 entries of rows 0..2 of dLoss/dM^-1 (row-major r*4+c) + loss

// Binary64 vector helpers for the adjoint's per-pixel solve.
struct D3 {
    double x, y, z;
};
__device__ __forceinline__ D3 dk(double x, double y, double z) { return D3{x, y, z}; }
__device__ __forceinline__ D3 dof(V3 a) { return D3{(double)a.x, (double)a.y, (double)a.z}; }
__device__ __forceinline__ D3 dsub(D3 a, D3 b) { return dk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ D3 dadd(D3 a, D3 b) { return dk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ D3 dscl(D3 a, double s) { return dk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ double ddot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 dcross(D3 a, D3 b)
{
    return dk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ D3 drows(const Rows& m, D3 p, double w)
{
    return dk((double)m.r0.x * p.x + (double)m.r0.y * p.y + (double)m.r0.z * p.z + (double)m.r0.w * w,
              (double)m.r1.x * p.x + (double)m.r1.y * p.y + (double)m.r1.z * p.z + (double)m.r1.w * w,
              (double)m.r2.x * p.x + (double)m.r2.y * p.y + (double)m.r2.z * p.z + (double)m.r2.w * w);
}

// Per pixel the solve runs in binary64 on the binary32 inputs: a grazing
// ray (|cos| ~ 1e-3, determinant near the 1e-4 threshold) amplifies
// binary32 rounding of t, u, v and lambda to percent-level errors in the
// gradients of the vertices it touches, and such pixels carry the largest
// gradients of all.  The scatter and the outputs stay binary32.
__global__ __launch_bounds__(kAdjWG) void k_trace_adjoint(
    const float* __restrict__ tris, uint32_t ntris, const uint32_t* __restrict__ vid,
    const float* __restrict__ normals, Rows rows, V3 light, const bvht_ray* __restrict__ rays, uint32_t n,
    const bvht_hit* __restrict__ hits, const float2* __restrict__ target, float w_depth, float w_shade,
    uint32_t* __restrict__ rec_key, float* __restrict__ rec_val, uint32_t* __restrict__ tri_count,
    double* __restrict__ partials)
{
    __shared__ double red[kAdjWG / 64][kRed];
    uint32_t i = blockIdx.x * kAdjWG + threadIdx.x;
    double acc[kRed];
#pragma unroll
    for (int c = 0; c < kRed; ++c)
        acc[c] = 0.0;
    uint32_t prim = i < n ? hits[i].prim : BVHT_MISS;
    float2 tg = i < n ? target[i] : make_float2(kFltMax, 0.0f);
    uint32_t key = BVHT_MISS;  // the hit triangle, for contributing rays
    float cv[9];
#pragma unroll
    for (int c = 0; c < 9; ++c)
        cv[c] = 0.0f;
    if (prim < ntris && tg.x < kFltMax) {
        const float4* rp = reinterpret_cast<const float4*>(rays + i);
        float4 ra = rp[0], rb = rp[1];
        D3 ow = dk(ra.x, ra.y, ra.z), dw = dk(rb.x, rb.y, rb.z);
        D3 o = drows(rows, ow, 1.0);
        D3 d = drows(rows, dw, 0.0);
        const float* tp = tris + 9 * (size_t)prim;
        D3 v0 = dof(ld3(tp)), v1 = dof(ld3(tp + 3)), v2 = dof(ld3(tp + 6));
        // Moeller-Trumbore (geometry/triangle.rs:41-72), binary64
        D3 e1 = dsub(v1, v0), e2 = dsub(v2, v0);
        D3 pn = dcross(d, e2);
        double det = ddot(e1, pn);
        double f = 1.0 / det;
        D3 s = dsub(o, v0);
        double u = f * ddot(s, pn);
        D3 q = dcross(s, e1);
        double v = f * ddot(d, q);
        double t = f * ddot(e2, q);
        const uint32_t* vv = vid + 3 * (size_t)prim;
        D3 n0 = dof(ld3(normals + 3 * (size_t)vv[0]));
        D3 n1 = dof(ld3(normals + 3 * (size_t)vv[1]));
        D3 n2 = dof(ld3(normals + 3 * (size_t)vv[2]));
        D3 lt = dof(light);
        double w = 1.0 - u - v;
        D3 m = dadd(dadd(dscl(n0, w), dscl(n1, u)), dscl(n2, v));
        double ml = ddot(m, lt), len = sqrt(ddot(m, m));
        bool lit = ml > 0.0 && len > 0.0;
        double sh = lit ? ml / len : 0.0;
        double rt = t - (double)tg.x;
        double rs = sh - (double)tg.y;
        double gt = 2.0 * (double)w_depth * rt;
        double gs = 2.0 * (double)w_shade * rs;
        double gu = 0.0, gv = 0.0;
        if (lit) {
            double il = 1.0 / len;
            D3 dsdm = dscl(dsub(lt, dscl(m, sh * il)), il);
            gu = gs * ddot(dsdm, dsub(n1, n0));
            gv = gs * ddot(dsdm, dsub(n2, n0));
        }
        D3 a0 = dk(-d.x, -d.y, -d.z);
        D3 lam = dscl(dadd(dadd(dscl(dcross(e1, e2), gt), dscl(dcross(e2, a0), gu)), dscl(dcross(a0, e1), gv)), f);
        double bw[3] = {-w, -u, -v};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            cv[3 * k + 0] = (float)(lam.x * bw[k]);
            cv[3 * k + 1] = (float)(lam.y * bw[k]);
            cv[3 * k + 2] = (float)(lam.z * bw[k]);
        }
        key = prim;
        double lr[3] = {lam.x, lam.y, lam.z};
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            double tl = t * lr[r];
            acc[4 * r + 0] = lr[r] * ow.x + tl * dw.x;
            acc[4 * r + 1] = lr[r] * ow.y + tl * dw.y;
            acc[4 * r + 2] = lr[r] * ow.z + tl * dw.z;
            acc[4 * r + 3] = lr[r];
        }
        acc[12] = (double)w_depth * rt * rt + (double)w_shade * rs * rs;
    }
    // Per-triangle sums within the wave, in a fixed order, without floating-
    // point atomics.  The 64 lanes are sorted by (hit triangle, lane) with a
    // bitonic network of shuffles (lanes without a contribution sort last);
    // sorted position p takes the contributions of the lane it came from;
    // a segmented inclusive scan over the sorted positions sums each
    // triangle's run, and the run's last position writes the record
    // (triangle, 9 sums) in its own slot and counts it for the triangle.
    // Every other slot is marked empty.
    uint32_t lane64 = threadIdx.x & 63u;
    uint32_t rkey = BVHT_MISS;
    if (__ballot(key != BVHT_MISS)) {
        uint32_t src, skey;
        if (ntris < (1u << 26)) {
            // 32-bit keys; 2^26 - 1 marks "no contribution" and sorts last
            constexpr uint32_t kNoKey = (1u << 26) - 1u;
            uint32_t sk = ((key == BVHT_MISS ? kNoKey : key) << 6) | lane64;
            bitonic_sort<uint32_t, 64>(sk, lane64);
            src = sk & 63u;
            skey = sk >> 6;
            if (skey == kNoKey)
                skey = BVHT_MISS;
        } else {
            uint64_t sk = ((uint64_t)key << 6) | lane64;
            bitonic_sort<uint64_t, 64>(sk, lane64);
            src = (uint32_t)(sk & 63u);
            skey = (uint32_t)(sk >> 6);
        }
        float sv[9];
#pragma unroll
        for (int c = 0; c < 9; ++c)
            sv[c] = __shfl(cv[c], (int)src, 64);
        uint32_t prev = __shfl_up(skey, 1, 64), next = __shfl_down(skey, 1, 64);
        uint32_t f = (lane64 == 0 || prev != skey) ? 1u : 0u;
        const uint32_t lr = lane64 & 15u;
        seg_step<1>(sv, f, lr);
        seg_step<2>(sv, f, lr);
        seg_step<4>(sv, f, lr);
        seg_step<8>(sv, f, lr);
        // carry into row r: the segmented sum through lane 16 r - 1
        const uint32_t row = lane64 >> 4;
        const bool f31 = __builtin_amdgcn_readlane((int)f, 31) != 0;
        const bool f47 = __builtin_amdgcn_readlane((int)f, 47) != 0;
#pragma unroll
        for (int c = 0; c < 9; ++c) {
            float c1 = readlane_f(sv[c], 15);
            float v31 = readlane_f(sv[c], 31), v47 = readlane_f(sv[c], 47);
            float c2 = f31 ? v31 : c1 + v31;
            float c3 = f47 ? v47 : c2 + v47;
            float carry = row == 1 ? c1 : (row == 2 ? c2 : c3);
            if (row > 0 && !f)
                sv[c] = carry + sv[c];
        }
        if (skey != BVHT_MISS && (lane64 == 63u || next != skey)) {
            rkey = skey;
            float* out = rec_val + 9 * (size_t)i;
#pragma unroll
            for (int c = 0; c < 9; ++c)
                out[c] = sv[c];
            atomicAdd(tri_count + skey, 1u);
        }
    }
    rec_key[i] = rkey;
    // fixed-order reduction: butterfly within the wave, then the waves in order
#pragma unroll
    for (int c = 0; c < kRed; ++c)
        acc[c] = wave_sum_d(acc[c], threadIdx.x & 63u);
    uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    if (lane == 0) {
#pragma unroll
        for (int c = 0; c < kRed; ++c)
            red[wave][c] = acc[c];
    }
    __syncthreads();
    if (threadIdx.x < (uint32_t)kRed) {
        double sum = 0.0;
        for (int k = 0; k < kAdjWG / 64; ++k)
            sum += red[k][threadIdx.x];
        partials[(size_t)blockIdx.x * 16 + threadIdx.x] = sum;
    }
}

// Sum the per-block partials in a fixed order; write (or add) dLoss/dM^-1
// (column-major, bottom row 0) and the loss.  Thread t sums blocks t,
// t + kFinWG, ... for all components at once (each block's partials are one
// 128-byte row), then a butterfly within each wave and the waves in order.
constexpr int kFinWG = 1024;

__global__ __launch_bounds__(kFinWG) void k_adjoint_finish(const double* __restrict__ partials, uint32_t nblocks,
                                                           int accumulate, float* __restrict__ grad_minv,
                                                           double* __restrict__ loss)
{
    __shared__ double red[kFinWG / 64][kRed];
    double acc[kRed];
#pragma unroll
    for (int c = 0; c < kRed; ++c)
        acc[c] = 0.0;
    for (uint32_t b = threadIdx.x; b < nblocks; b += kFinWG) {
        const double2* row = reinterpret_cast<const double2*>(partials + (size_t)b * 16);
#pragma unroll
        for (int c = 0; c + 1 < kRed; c += 2) {
            double2 v = row[c / 2];
            acc[c] += v.x;
            acc[c + 1] += v.y;
        }
        acc[kRed - 1] += partials[(size_t)b * 16 + kRed - 1];
    }
#pragma unroll
    for (int c = 0; c < kRed; ++c)
        acc[c] = wave_sum_d(acc[c], threadIdx.x & 63u);
    uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    if (lane == 0) {
#pragma unroll
        for (int c = 0; c < kRed; ++c)
            red[wave][c] = acc[c];
    }
    __syncthreads();
    if (threadIdx.x < (uint32_t)kRed) {
        int c = (int)threadIdx.x;
        double sum = 0.0;
        for (int w = 0; w < kFinWG / 64; ++w)
            sum += red[w][c];
        if (c < 12) {
            int r = c / 4, col = c % 4;
            float g = (float)sum;
            float* p = grad_minv + col * 4 + r;
            *p = accumulate ? *p + g : g;
        } else {
            *loss = accumulate ? *loss + sum : sum;
        }
    } else if (threadIdx.x >= 64 && threadIdx.x < 68 && !accumulate) {
        grad_minv[(threadIdx.x - 64) * 4 + 3] = 0.0f;
    }
}

// Bucket the wave records by triangle: tri_off[t] is the exclusive prefix of
// the per-triangle record counts, tri_cur[t] starts at 0.  The order inside
// a bucket comes from the atomics; k_adj_triangles sorts it.
__global__ __launch_bounds__(256) void k_adj_place(const uint32_t* __restrict__ rec_key, uint32_t slots,
                                                    const uint32_t* __restrict__ tri_off,
                                                    uint32_t* __restrict__ tri_cur, uint32_t* __restrict__ bucket)
{
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= slots)
        return;
    uint32_t t = rec_key[i];
    if (t != BVHT_MISS)
        bucket[tri_off[t] + atomicAdd(tri_cur + t, 1u)] = i;
}

// Corner gradients of triangle t: its records summed in ascending slot order,
// so the sum is the same on every run.  A bucket of at most kAdjSmall
// records is copied to this thread's LDS column and insertion-sorted there;
// a larger one is listed for k_adj_big.  Leaves tri_off[t] and tri_cur[t]
// zero for the next adjoint pass (k_adj_big does so for its triangles).
constexpr uint32_t kAdjSmall = 32;
constexpr uint32_t kAdjBigMax = 32768;  // records of one triangle in one pass: one per wave of 2^21 rays
constexpr int kAdjBigWG = 1024;

__device__ __forceinline__ void store9(float* __restrict__ out, const float (&g)[9])
{
#pragma unroll
    for (int c = 0; c < 9; ++c)
        out[c] = g[c];
}

__global__ __launch_bounds__(256) void k_adj_triangles(uint32_t nt, uint32_t* __restrict__ tri_off,
                                                        uint32_t* __restrict__ tri_cur,
                                                        const uint32_t* __restrict__ bucket,
                                                        const float* __restrict__ rec_val, float* __restrict__ tgrad,
                                                        uint32_t* __restrict__ big, uint32_t* __restrict__ n_big)
{
    __shared__ uint32_t col[kAdjSmall][256];
    uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= nt)
        return;
    uint32_t cnt = tri_cur[t];
    if (cnt > kAdjSmall) {
        big[atomicAdd(n_big, 1u)] = t;
        return;
    }
    const uint32_t* b = bucket + tri_off[t];
    for (uint32_t x = 0; x < cnt; ++x) {
        uint32_t v = b[x], y = x;
        for (; y > 0 && col[y - 1][threadIdx.x] > v; --y)
            col[y][threadIdx.x] = col[y - 1][threadIdx.x];
        col[y][threadIdx.x] = v;
    }
    float g[9];
#pragma unroll
    for (int c = 0; c < 9; ++c)
        g[c] = 0.0f;
    for (uint32_t x = 0; x < cnt; ++x) {
        const float* r = rec_val + 9 * (size_t)col[x][threadIdx.x];
#pragma unroll
        for (int c = 0; c < 9; ++c)
            g[c] += r[c];
    }
    store9(tgrad + 9 * (size_t)t, g);
    tri_off[t] = 0u;
    tri_cur[t] = 0u;
}

// Triangles with more than kAdjSmall records (at most kAdjBigMax): one
// workgroup per triangle, the gridded workgroups taking the listed triangles
// in turn.  The bucket is bitonic-sorted in LDS, thread k sums sorted
// records [k c, (k+1) c) in order (c = ceil(cnt / kAdjBigWG)), the partial
// sums are added by a butterfly within each wave and the waves in order.
__global__ __launch_bounds__(kAdjBigWG) void k_adj_big(const uint32_t* __restrict__ big,
                                                      const uint32_t* __restrict__ n_big,
                                                      uint32_t* __restrict__ tri_off, uint32_t* __restrict__ tri_cur,
                                                      const uint32_t* __restrict__ bucket,
                                                      const float* __restrict__ rec_val, float* __restrict__ tgrad)
{
    __shared__ uint32_t keys[kAdjBigMax];
    __shared__ float wsum[kAdjBigWG / 64][9];
    uint32_t nb = *n_big;
    for (uint32_t q = blockIdx.x; q < nb; q += gridDim.x) {
        uint32_t t = big[q];
        uint32_t cnt = tri_cur[t];
        const uint32_t* b = bucket + tri_off[t];
        uint32_t p2 = 64;
        while (p2 < cnt)
            p2 <<= 1;
        for (uint32_t x = threadIdx.x; x < p2; x += kAdjBigWG)
            keys[x] = x < cnt ? b[x] : 0xFFFFFFFFu;
        __syncthreads();
        for (uint32_t k = 2; k <= p2; k <<= 1)
            for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                for (uint32_t x = threadIdx.x; x < p2; x += kAdjBigWG) {
                    uint32_t y = x ^ j;
                    if (y > x) {
                        uint32_t a = keys[x], c = keys[y];
                        bool up = (x & k) == 0;
                        if (up ? a > c : a < c) {
                            keys[x] = c;
                            keys[y] = a;
                        }
                    }
                }
                __syncthreads();
            }
        uint32_t per = (cnt + kAdjBigWG - 1) / kAdjBigWG;
        float g[9];
#pragma unroll
        for (int c = 0; c < 9; ++c)
            g[c] = 0.0f;
        for (uint32_t x = threadIdx.x * per; x < min(cnt, (threadIdx.x + 1) * per); ++x) {
            const float* r = rec_val + 9 * (size_t)keys[x];
#pragma unroll
            for (int c = 0; c < 9; ++c)
                g[c] += r[c];
        }
#pragma unroll
        for (int c = 0; c < 9; ++c)
            for (int m = 32; m >= 1; m >>= 1)
                g[c] += __shfl_xor(g[c], m, 64);
        if ((threadIdx.x & 63u) == 0) {
#pragma unroll
            for (int c = 0; c < 9; ++c)
                wsum[threadIdx.x >> 6][c] = g[c];
        }
        __syncthreads();
        if (threadIdx.x < 9) {
            float s = 0.0f;
            for (int w = 0; w < kAdjBigWG / 64; ++w)
                s += wsum[w][threadIdx.x];
            tgrad[9 * (size_t)t + threadIdx.x] = s;
        }
        if (threadIdx.x == 0) {
            tri_off[t] = 0u;
            tri_cur[t] = 0u;
        }
        __syncthreads();
    }
}

// dLoss/dvertex: the corner gradients of the vertex's incident (triangle,
// corner) entries, in their fixed ascending order; written, or added with
// accumulate.
__global__ __launch_bounds__(256) void k_adj_vertices(uint32_t nv, const uint32_t* __restrict__ vt_off,
                                                       const uint32_t* __restrict__ vt_idx,
                                                       const float* __restrict__ tgrad, int accumulate,
                                                       float* __restrict__ grad_v)
{
    uint32_t v = blockIdx.x * 256 + threadIdx.x;
    if (v >= nv)
        return;
    float gx = 0.0f, gy = 0.0f, gz = 0.0f;
    for (uint32_t e = vt_off[v]; e < vt_off[v + 1]; ++e) {
        const float* g = tgrad + 3 * (size_t)vt_idx[e];  // 9 t + 3 c = 3 (3 t + c)
        gx += g[0];
        gy += g[1];
        gz += g[2];
    }
    float* out = grad_v + 3 * (size_t)v;
    if (accumulate) {
        out[0] = out[0] + gx;
        out[1] = out[1] + gy;
        out[2] = out[2] + gz;
    } else {
        out[0] = gx;
        out[1] = gy;
        out[2] = gz;
    }
}

}  // namespace
}  // namespace bvht

using namespace bvht;

extern "C" int32_t bvht_blas_permutation(bvht_ctx ctx, bvht_blas blas, uint32_t* perm)
{
    BVHT_REQUIRE(ctx && blas && perm, "NULL argument");
    BVHT_REQUIRE(blas->d_perm != nullptr, "this BLAS was adopted, not built: it has no build permutation");
    BVHT_HIP(hipMemcpyAsync(perm, blas->d_perm, sizeof(uint32_t) * (size_t)blas->n_tris, hipMemcpyDeviceToHost,
                            ctx->stream));
    BVHT_HIP(hipStreamSynchronize(ctx->stream));
    return BVHT_OK;
}

extern "C" int32_t bvht_mesh_destroy(bvht_mesh m)
{
    if (!m)
        return BVHT_OK;
    if (m->ctx)
        (void)hipStreamSynchronize(m->ctx->stream);
    (void)hipFree(m->d_vertices);
    (void)hipFree(m->d_normals);
    (void)hipFree(m->d_faces);
    (void)hipFree(m->d_vid);
    (void)hipFree(m->d_vf_off);
    (void)hipFree(m->d_vf_idx);
    (void)hipFree(m->d_sched);
    (void)hipFree(m->d_cost);
    (void)hipFree(m->d_bins);
    (void)hipFree(m->d_vt_off);
    (void)hipFree(m->d_vt_idx);
    (void)hipFree(m->d_tgrad);
    (void)hipFree(m->d_poff);
    (void)hipFree(m->d_pcur);
    (void)hipFree(m->d_psums);
    (void)hipFree(m->d_big);
    (void)hipFree(m->d_nbig);
    delete m;
    return BVHT_OK;
}

static int32_t launch_normals(bvht_ctx ctx, bvht_mesh m)
{
    hipLaunchKernelGGL(k_vertex_normals, dim3(nblk(m->n_vertices, 256)), dim3(256), 0, ctx->stream, m->d_vertices,
                       m->d_faces, m->d_vf_off, m->d_vf_idx, m->n_vertices, m->d_normals);
    BVHT_HIP(hipGetLastError());
    return BVHT_OK;
}

extern "C" int32_t bvht_mesh_create(bvht_ctx ctx, bvht_blas blas, const float* vertices, uint32_t n_vertices,
                                    const uint32_t* faces, uint32_t n_faces, const uint32_t* face_of, bvht_mesh* out)
{
    BVHT_REQUIRE(ctx && blas && vertices && faces && out, "NULL argument");
    BVHT_REQUIRE(blas->ctx == ctx, "the BLAS belongs to another context");
    BVHT_REQUIRE(n_vertices > 0 && n_faces > 0, "empty mesh");
    BVHT_REQUIRE(n_faces <= (1u << 30) && n_vertices <= (1u << 30), "mesh too large");
    uint32_t nt = blas->n_tris;
    for (size_t k = 0; k < 3 * (size_t)n_faces; ++k)
        BVHT_REQUIRE(faces[k] < n_vertices, "face %zu references vertex %u of %u", k / 3, faces[k], n_vertices);
    std::vector<uint32_t> fo(nt);
    if (face_of) {
        std::copy(face_of, face_of + nt, fo.begin());
    } else {
        BVHT_REQUIRE(blas->d_perm != nullptr, "face_of is NULL and the BLAS has no build permutation");
        BVHT_REQUIRE(nt == n_faces, "the BLAS has %u triangles, the mesh %u faces", nt, n_faces);
        BVHT_HIP(hipMemcpyAsync(fo.data(), blas->d_perm, sizeof(uint32_t) * nt, hipMemcpyDeviceToHost, ctx->stream));
        BVHT_HIP(hipStreamSynchronize(ctx->stream));
    }
    std::vector<uint32_t> vid(3 * (size_t)nt);
    for (uint32_t k = 0; k < nt; ++k) {
        BVHT_REQUIRE(fo[k] < n_faces, "face_of[%u] = %u out of range (%u faces)", k, fo[k], n_faces);
        for (int c = 0; c < 3; ++c)
            vid[3 * (size_t)k + c] = faces[3 * (size_t)fo[k] + c];
    }
    // vertex -> incident faces (ascending), by counting sort
    std::vector<uint32_t> off(n_vertices + 1, 0), idx(3 * (size_t)n_faces);
    for (size_t k = 0; k < 3 * (size_t)n_faces; ++k)
        ++off[faces[k] + 1];
    for (uint32_t i = 0; i < n_vertices; ++i)
        off[i + 1] += off[i];
    {
        std::vector<uint32_t> cur(off.begin(), off.end() - 1);
        for (uint32_t f = 0; f < n_faces; ++f)
            for (int c = 0; c < 3; ++c)
                idx[cur[faces[3 * (size_t)f + c]]++] = f;
    }
    // vertex -> (BLAS triangle k, corner c) entries 3k + c, ascending
    std::vector<uint32_t> vt_off(n_vertices + 1, 0), vt_idx(3 * (size_t)nt);
    for (size_t e = 0; e < vid.size(); ++e)
        ++vt_off[vid[e] + 1];
    for (uint32_t i = 0; i < n_vertices; ++i)
        vt_off[i + 1] += vt_off[i];
    {
        std::vector<uint32_t> cur(vt_off.begin(), vt_off.end() - 1);
        for (size_t e = 0; e < vid.size(); ++e)
            vt_idx[cur[vid[e]]++] = (uint32_t)e;
    }
    bvht_mesh m = new bvht_mesh_s;
    m->ctx = ctx;
    m->blas = blas;
    m->n_vertices = n_vertices;
    m->n_faces = n_faces;
    size_t vb = sizeof(float) * 3 * (size_t)n_vertices;
    bool ok = hipMalloc(&m->d_vertices, vb) == hipSuccess && hipMalloc(&m->d_normals, vb) == hipSuccess &&
              hipMalloc(&m->d_faces, sizeof(uint32_t) * 3 * (size_t)n_faces) == hipSuccess &&
              hipMalloc(&m->d_vid, sizeof(uint32_t) * vid.size()) == hipSuccess &&
              hipMalloc(&m->d_vf_off, sizeof(uint32_t) * off.size()) == hipSuccess &&
              hipMalloc(&m->d_vf_idx, sizeof(uint32_t) * idx.size()) == hipSuccess &&
              hipMalloc(&m->d_vt_off, sizeof(uint32_t) * vt_off.size()) == hipSuccess &&
              hipMalloc(&m->d_vt_idx, sizeof(uint32_t) * vt_idx.size()) == hipSuccess &&
              hipMalloc(&m->d_tgrad, sizeof(float) * 9 * (size_t)nt) == hipSuccess &&
              hipMalloc(&m->d_poff, sizeof(uint32_t) * std::max<size_t>(nt, 4)) == hipSuccess &&
              hipMalloc(&m->d_pcur, sizeof(uint32_t) * std::max<size_t>(nt, 4)) == hipSuccess &&
              hipMalloc(&m->d_psums, sizeof(uint32_t) * scan_sums_entries(nt)) == hipSuccess &&
              hipMalloc(&m->d_big, sizeof(uint32_t) * std::max<size_t>(nt, 4)) == hipSuccess &&
              hipMalloc(&m->d_nbig, sizeof(uint32_t) * 4) == hipSuccess;
    if (!ok) {
        bvht_mesh_destroy(m);
        return set_error(BVHT_ENOMEM, "device allocation for a mesh of %u vertices failed", n_vertices);
    }
    hipStream_t st = ctx->stream;
    int32_t rc = BVHT_OK;
    uint32_t* d_err = nullptr;
    uint32_t err = 0;
    if (hipMemcpyAsync(m->d_vertices, vertices, vb, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(m->d_faces, faces, sizeof(uint32_t) * 3 * (size_t)n_faces, hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        hipMemcpyAsync(m->d_vid, vid.data(), sizeof(uint32_t) * vid.size(), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(m->d_vf_off, off.data(), sizeof(uint32_t) * off.size(), hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        hipMemcpyAsync(m->d_vf_idx, idx.data(), sizeof(uint32_t) * idx.size(), hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        hipMemcpyAsync(m->d_vt_off, vt_off.data(), sizeof(uint32_t) * vt_off.size(), hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        hipMemcpyAsync(m->d_vt_idx, vt_idx.data(), sizeof(uint32_t) * vt_idx.size(), hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        hipMalloc(&d_err, sizeof(uint32_t)) != hipSuccess) {
        rc = set_error(BVHT_EDEVICE, "mesh upload failed");
    }
    if (!rc)
        rc = fill_u32(ctx, d_err, 0u, 1);
    if (!rc)
        rc = fill_u32(ctx, m->d_poff, 0u, nt);
    if (!rc)
        rc = fill_u32(ctx, m->d_pcur, 0u, nt);
    if (!rc) {
        hipLaunchKernelGGL(k_mesh_check, dim3(nblk(nt, 256)), dim3(256), 0, st, m->d_vertices, m->d_vid, nt,
                           blas->d_tris, d_err);
        rc = launch_normals(ctx, m);
    }
    if (!rc && (hipMemcpyAsync(&err, d_err, sizeof err, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess))
        rc = set_error(BVHT_EDEVICE, "mesh check failed to run");
    (void)hipFree(d_err);
    if (!rc && err)
        rc = set_error(BVHT_EINVAL, "the BLAS triangles are not the mesh faces in face_of order");
    if (rc) {
        bvht_mesh_destroy(m);
        return rc;
    }
    *out = m;
    return BVHT_OK;
}

extern "C" int32_t bvht_mesh_buffers(bvht_mesh mesh, float** d_vertices, float** d_normals)
{
    BVHT_REQUIRE(mesh, "NULL mesh");
    if (d_vertices)
        *d_vertices = mesh->d_vertices;
    if (d_normals)
        *d_normals = mesh->d_normals;
    return BVHT_OK;
}

extern "C" int32_t bvht_mesh_update(bvht_ctx ctx, bvht_mesh mesh, int32_t normals)
{
    BVHT_REQUIRE(ctx && mesh && mesh->ctx == ctx, "NULL argument or mesh of another context");
    bvht_blas b = mesh->blas;
    hipLaunchKernelGGL(k_mesh_gather, dim3(nblk(b->n_tris, 256)), dim3(256), 0, ctx->stream, mesh->d_vertices,
                       mesh->d_vid, b->n_tris, b->d_tris);
    BVHT_HIP(hipGetLastError());
    int32_t rc = bvht_blas_refit(ctx, b);
    if (rc)
        return rc;
    return normals ? launch_normals(ctx, mesh) : BVHT_OK;
}

extern "C" int32_t bvht_mesh_sgd(bvht_ctx ctx, bvht_mesh mesh, const float* d_grad, float lr, float max_step,
                                 int32_t flags)
{
    BVHT_REQUIRE(ctx && mesh && d_grad && mesh->ctx == ctx, "NULL argument or mesh of another context");
    BVHT_REQUIRE((flags & ~BVHT_SGD_ALONG_NORMAL) == 0, "unknown SGD flags 0x%x", flags);
    uint32_t n = mesh->n_vertices;
    hipLaunchKernelGGL(k_sgd, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, mesh->d_vertices, d_grad,
                       mesh->d_normals, n, lr, max_step, (int)(flags & BVHT_SGD_ALONG_NORMAL));
    BVHT_HIP(hipGetLastError());
    return BVHT_OK;
}

extern "C" int32_t bvht_trace_forward(bvht_ctx ctx, bvht_mesh mesh, const float* minv, const float* light,
                                      const bvht_ray* d_rays, uint32_t n, bvht_hit* d_hits, float* d_image)
{
    BVHT_REQUIRE(ctx && mesh && minv && light && mesh->ctx == ctx, "NULL argument or mesh of another context");
    if (n == 0)
        return BVHT_OK;
    BVHT_REQUIRE(d_rays && d_hits && d_image, "NULL device buffer");
    BVHT_REQUIRE(n <= 0xFFFFFF00u, "too many rays (%u)", n);
    bvht_blas b = mesh->blas;
    uint32_t need = stack_cap(ctx, b->depth > 0 ? b->depth - 1 : 0);
    int wg, lds;
    trace_config(ctx, need, &wg, &lds);
    uint32_t* spill = nullptr;
    if (need > (uint32_t)lds) {
        int32_t rc = ensure_spill(ctx, &ctx->spill, &ctx->spill_words, (size_t)(need - lds) * n);
        if (rc)
            return rc;
        spill = ctx->spill;
    }
    forward_fn fn = pick_forward(wg, lds, mesh->blas->packed_stack);
    if (!fn)
        return set_error(BVHT_EINVAL, "no forward kernel for workgroup %d, LDS depth %d", wg, lds);
    uint32_t groups = nblk(n, 64);
    if (groups > mesh->sched_cap) {
        BVHT_REQUIRE(ctx->live_graphs == 0 && !ctx->capturing,
                     "the forward schedule must grow to %u groups, but a graph captured on this context holds its "
                     "address: make one eager call of the largest ray count first", groups);
        BVHT_HIP(hipStreamSynchronize(ctx->stream));
        (void)hipFree(mesh->d_sched);
        (void)hipFree(mesh->d_cost);
        mesh->d_sched = mesh->d_cost = nullptr;
        mesh->sched_cap = mesh->sched_n = 0;
        if (hipMalloc(&mesh->d_sched, sizeof(uint32_t) * groups) != hipSuccess ||
            hipMalloc(&mesh->d_cost, sizeof(uint32_t) * groups) != hipSuccess ||
            (!mesh->d_bins && hipMalloc(&mesh->d_bins, sizeof(uint32_t) * kSchedBinsWords) != hipSuccess))
            return set_error(BVHT_ENOMEM, "forward schedule for %u groups", groups);
        mesh->sched_cap = groups;
    }
    // the schedule made by the last forward over the same number of rays
    const uint32_t* sched = mesh->sched_n == n ? mesh->d_sched : nullptr;
    hipLaunchKernelGGL(fn, dim3(nblk(groups, wg / 64)), dim3(wg), 0, ctx->stream,
                       reinterpret_cast<const float4*>(b->d_nodes), b->d_tris, mesh->d_vid, mesh->d_normals,
                       rows_of(minv), V3{light[0], light[1], light[2]}, d_rays, n, d_hits,
                       reinterpret_cast<float2*>(d_image), spill, need, ctx->d_status, sched,
                       sched_units(mesh->d_bins), mesh->d_cost);
    // the schedule for the next forward: longest groups first
    int32_t rc = launch_longest_first(ctx, mesh->d_cost, groups, mesh->d_bins, mesh->d_sched);
    if (rc)
        return rc;
    mesh->sched_n = n;
    BVHT_HIP(hipGetLastError());
    return BVHT_OK;
}

extern "C" int32_t bvht_trace_adjoint(bvht_ctx ctx, bvht_mesh mesh, const float* minv, const float* light,
                                      const bvht_ray* d_rays, uint32_t n, const bvht_hit* d_hits,
                                      const float* d_target, float w_depth, float w_shade, int32_t accumulate,
                                      float* d_grad_vertices, float* d_grad_minv, double* d_loss)
{
    BVHT_REQUIRE(ctx && mesh && minv && light && mesh->ctx == ctx, "NULL argument or mesh of another context");
    BVHT_REQUIRE(d_grad_vertices && d_grad_minv && d_loss, "NULL output buffer");
    BVHT_REQUIRE(n == 0 || (d_rays && d_hits && d_target), "NULL device buffer");
    BVHT_REQUIRE(n <= 0xFFFFFF00u, "too many rays (%u)", n);
    // Passes of at most kAdjPass rays: a triangle then gets at most one record
    // per wave of the pass (kAdjBigMax), which k_adj_big's LDS holds.  The
    // vertex gradients of later passes add to the earlier ones in pass order.
    constexpr uint32_t kAdjPass = kAdjBigMax * 64;
    uint32_t nb = std::max<uint32_t>(nblk(n, kAdjWG), 1);
    uint32_t pass_slots = std::min<uint32_t>(nb, kAdjPass / kAdjWG) * kAdjWG;
    // context scratch: per-block partials of the whole call, then per-slot
    // record keys, record values (9 floats) and the triangle buckets of a pass
    size_t keys_at = align256(sizeof(double) * 16 * (size_t)nb);
    size_t vals_at = align256(keys_at + sizeof(uint32_t) * (size_t)pass_slots);
    size_t bucket_at = align256(vals_at + sizeof(float) * 9 * (size_t)pass_slots);
    int32_t rc = ensure_scratch(ctx, bucket_at + sizeof(uint32_t) * (size_t)pass_slots);
    if (rc)
        return rc;
    char* base = static_cast<char*>(ctx->scratch);
    double* partials = reinterpret_cast<double*>(base);
    uint32_t* rec_key = reinterpret_cast<uint32_t*>(base + keys_at);
    float* rec_val = reinterpret_cast<float*>(base + vals_at);
    uint32_t* bucket = reinterpret_cast<uint32_t*>(base + bucket_at);
    bvht_blas b = mesh->blas;
    uint32_t nt = b->n_tris;
    for (uint32_t first = 0; first == 0 || first < n; first += kAdjPass) {
        uint32_t m = std::min(n - first, kAdjPass);
        uint32_t blocks = std::max<uint32_t>(nblk(m, kAdjWG), 1);
        uint32_t slots = blocks * kAdjWG;
        hipLaunchKernelGGL(k_trace_adjoint, dim3(blocks), dim3(kAdjWG), 0, ctx->stream, b->d_tris, nt, mesh->d_vid,
                           mesh->d_normals, rows_of(minv), V3{light[0], light[1], light[2]}, d_rays + first, m,
                           d_hits + first, reinterpret_cast<const float2*>(d_target) + first, w_depth, w_shade,
                           rec_key, rec_val, mesh->d_poff, partials + 16 * (size_t)(first / kAdjWG));
        launch_scan(ctx, mesh->d_poff, nt, mesh->d_psums, nullptr);
        hipLaunchKernelGGL(k_adj_place, dim3(nblk(slots, 256)), dim3(256), 0, ctx->stream, rec_key, slots,
                           mesh->d_poff, mesh->d_pcur, bucket);
        rc = fill_u32(ctx, mesh->d_nbig, 0u, 1);
        if (rc)
            return rc;
        hipLaunchKernelGGL(k_adj_triangles, dim3(nblk(nt, 256)), dim3(256), 0, ctx->stream, nt, mesh->d_poff,
                           mesh->d_pcur, bucket, rec_val, mesh->d_tgrad, mesh->d_big, mesh->d_nbig);
        hipLaunchKernelGGL(k_adj_big, dim3(std::min<uint32_t>(nt, 64)), dim3(kAdjBigWG), 0, ctx->stream,
                           mesh->d_big, mesh->d_nbig, mesh->d_poff, mesh->d_pcur, bucket, rec_val, mesh->d_tgrad);
        hipLaunchKernelGGL(k_adj_vertices, dim3(nblk(mesh->n_vertices, 256)), dim3(256), 0, ctx->stream,
                           mesh->n_vertices, mesh->d_vt_off, mesh->d_vt_idx, mesh->d_tgrad,
                           (int)(accumulate || first > 0), d_grad_vertices);
        BVHT_HIP(hipGetLastError());
        if (n - first <= kAdjPass)
            break;
    }
    hipLaunchKernelGGL(k_adjoint_finish, dim3(1), dim3(kFinWG), 0, ctx->stream, partials, nb, (int)accumulate,
                       d_grad_minv, d_loss);
    BVHT_HIP(hipGetLastError());
    return BVHT_OK;
}
