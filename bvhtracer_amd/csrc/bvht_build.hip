// bvht_build.hip -- GPU BLAS build (BvhBuilder::build_for, bvh.rs:509-541).
//
// The reference builds top-down, depth first, with a binned SAH of 8 bins
// and an in-place two-pointer partition.  The result (node array and triangle
// permutation) depends only on each node's own primitive range, so it can be
// produced breadth first:
//
//   * Large nodes (more than kSmall primitives) are processed a level at a
//     time.  Their primitive ranges are cut into chunks of kChunk elements,
//     one workgroup per chunk; per-node reductions (centroid range, bin
//     counts and boxes, child boxes) combine chunk partials with atomics on
//     order-preserving float keys (min/max are exact, so the order of
//     combination does not matter).  The split decision is made by one
//     thread per node with the reference's sequential arithmetic.  The
//     two-pointer partition permutation is reproduced exactly from prefix
//     counts (derivation at k_part_scatter).
//   * Small nodes are finished by one thread each, running the reference's
//     sequential subdivide on their range.
//   * Node numbering: the reference allocates child pairs in depth-first
//     pre-order (bvh.rs:440-466), so the children of the internal node with
//     pre-order rank r are at 2+2r and 3+2r.  Ranks follow from the internal
//     counts of the subtrees once the tree is known.
//
// Per element the build keeps its centroid (Triangle::centroid, exact), the
// min/max of its three vertices (growing a box by the three vertices equals
// growing it by their min and max) and its original index.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "bvht_device.hpp"
#include "bvht_internal.hpp"

using namespace bvht;

namespace {

constexpr uint32_t kSmall = 32;     // nodes with <= kSmall primitives are finished by one thread
constexpr uint32_t kChunk = 4096;   // elements per workgroup in the level kernels
constexpr uint32_t kWG = 256;
constexpr int kBins = 8;

// Element record, held in three arrays for aligned loads:
//   ec  float4  centroid xyz, original triangle index (bits) in w
//   elo float4  vertex min xyz, vertex max x in w
//   ehi float2  vertex max yz
struct Task {       // a large node of the current level
    uint32_t first, count;
    float box[6];   // node box (update_node_bounds)
};

struct Chunk {
    uint32_t slot, begin, end, pad;
};

struct Decision {
    int32_t kind;   // 0 leaf, 1 plane split, 2 median split
    int32_t axis;
    float pos;
    uint32_t nl;    // left count after the partition
};

constexpr uint32_t kEmptyMin = 0xFFFFFFFFu;  // key sentinels never produced by a non-NaN value
constexpr uint32_t kEmptyMax = 0u;

__device__ __forceinline__ bool not_nan(float f) { return f == f; }

__device__ __forceinline__ void key_min(uint32_t& k, float f)
{
    if (not_nan(f))
        k = min(k, f2key(f));
}
__device__ __forceinline__ void key_max(uint32_t& k, float f)
{
    if (not_nan(f))
        k = max(k, f2key(f));
}

__device__ __forceinline__ uint32_t wave_min(uint32_t v)
{
    for (int o = 32; o > 0; o >>= 1)
        v = min(v, (uint32_t)__shfl_xor((int)v, o));
    return v;
}
__device__ __forceinline__ uint32_t wave_max(uint32_t v)
{
    for (int o = 32; o > 0; o >>= 1)
        v = max(v, (uint32_t)__shfl_xor((int)v, o));
    return v;
}

// Block-wide reduction of `nv` key values per thread into global atomics:
// first half of vals[] reduced with min, second half with max.
template <int NV>
__device__ __forceinline__ void block_reduce_keys(uint32_t (&vals)[NV], uint32_t* __restrict__ gdst,
                                                  uint32_t* lds /* NV * 4 */)
{
    uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        uint32_t v = (k < NV / 2) ? wave_min(vals[k]) : wave_max(vals[k]);
        if (lane == 0)
            lds[k * 4 + wave] = v;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        int k = threadIdx.x;
        uint32_t a = lds[k * 4 + 0], b = lds[k * 4 + 1], c = lds[k * 4 + 2], d = lds[k * 4 + 3];
        if (k < NV / 2) {
            uint32_t v = min(min(a, b), min(c, d));
            if (v != kEmptyMin)
                atomicMin(gdst + k, v);
        } else {
            uint32_t v = max(max(a, b), max(c, d));
            if (v != kEmptyMax)
                atomicMax(gdst + k, v);
        }
    }
    __syncthreads();
}

// ------------------------------------------------------------- kernels --

// Per triangle: centroid, vertex min/max, original index.
__global__ void k_init_elems(const float* __restrict__ tris, uint32_t n, float4* __restrict__ ec,
                             float4* __restrict__ elo, float2* __restrict__ ehi)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float* t = tris + 9 * (size_t)i;
    V3 v0 = mk3(t[0], t[1], t[2]), v1 = mk3(t[3], t[4], t[5]), v2 = mk3(t[6], t[7], t[8]);
    V3 c = tri_centroid(v0, v1, v2);
    V3 lo = min3(min3(v0, v1), v2);
    V3 hi = max3(max3(v0, v1), v2);
    ec[i] = make_float4(c.x, c.y, c.z, __uint_as_float(i));
    elo[i] = make_float4(lo.x, lo.y, lo.z, hi.x);
    ehi[i] = make_float2(hi.y, hi.z);
}

__global__ void k_init_acc(uint32_t* __restrict__ cacc, uint32_t* __restrict__ bacc, uint32_t* __restrict__ child,
                           uint32_t L)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 6 * L)
        cacc[i] = (i % 6) < 3 ? kEmptyMin : kEmptyMax;
    if (i < 12 * L)
        child[i] = (i % 12) < 6 ? kEmptyMin : kEmptyMax;
    if (i < 3 * kBins * 7 * L) {
        uint32_t f = i % 7;
        bacc[i] = f == 0 ? 0u : (f <= 3 ? kEmptyMin : kEmptyMax);
    }
}

__global__ void k_fill(uint32_t* p, uint32_t v, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

// Box of a range of elements (min over vertex mins, max over vertex maxes),
// accumulated into keys[6] = {minx,miny,minz,maxx,maxy,maxz}.
__global__ __launch_bounds__(kWG) void k_range_box(const float4* __restrict__ elo, const float2* __restrict__ ehi,
                                                   uint32_t begin, uint32_t end, uint32_t* __restrict__ keys)
{
    __shared__ uint32_t lds[6 * 4];
    uint32_t v[6] = {kEmptyMin, kEmptyMin, kEmptyMin, kEmptyMax, kEmptyMax, kEmptyMax};
    for (uint32_t e = begin + blockIdx.x * kWG + threadIdx.x; e < end; e += gridDim.x * kWG) {
        float4 lo = elo[e];
        float2 hi = ehi[e];
        key_min(v[0], lo.x); key_min(v[1], lo.y); key_min(v[2], lo.z);
        key_max(v[3], lo.w); key_max(v[4], hi.x); key_max(v[5], hi.y);
    }
    block_reduce_keys<6>(v, keys, lds);
}

// Centroid range per node and axis (find_best_split_plane, bvh.rs:338-344).
__global__ __launch_bounds__(kWG) void k_centroid_range(const float4* __restrict__ ec, const Chunk* __restrict__ chunks,
                                                        uint32_t* __restrict__ cacc /* slot*6 */)
{
    __shared__ uint32_t lds[6 * 4];
    Chunk ch = chunks[blockIdx.x];
    uint32_t v[6] = {kEmptyMin, kEmptyMin, kEmptyMin, kEmptyMax, kEmptyMax, kEmptyMax};
    for (uint32_t e = ch.begin + threadIdx.x; e < ch.end; e += kWG) {
        float4 c = ec[e];
        key_min(v[0], c.x); key_min(v[1], c.y); key_min(v[2], c.z);
        key_max(v[3], c.x); key_max(v[4], c.y); key_max(v[5], c.z);
    }
    block_reduce_keys<6>(v, cacc + 6 * ch.slot, lds);
}

struct AxisRange {
    float bmin, bmax, scale;
    bool valid;
};

// bounds_min/bounds_max start at (1e30, 1e-30) in the reference (Q2 of the
// reference's quirks; bvh.rs:339-340) and at (MAX, -MAX) in SAH mode.
__device__ __forceinline__ AxisRange axis_range(const uint32_t* cacc, int axis, int faithful)
{
    float init_min = faithful ? 1e30f : kFltMax;
    float init_max = faithful ? 1e-30f : -kFltMax;
    uint32_t kmin = cacc[axis], kmax = cacc[3 + axis];
    float bmin = kmin == kEmptyMin ? init_min : fminf(init_min, key2f(kmin));
    float bmax = kmax == kEmptyMax ? init_max : fmaxf(init_max, key2f(kmax));
    AxisRange r;
    r.bmin = bmin;
    r.bmax = bmax;
    r.valid = !(bmin == bmax);
    r.scale = (float)kBins / (bmax - bmin);
    return r;
}

// `f32 as usize` then min(BIN_COUNT - 1, _) (bvh.rs:352-353).
__device__ __forceinline__ uint32_t bin_index(float p)
{
    if (!(p > 0.0f))
        return 0;
    if (p >= (float)(kBins - 1))
        return kBins - 1;
    return (uint32_t)p;
}

// Bin counts and boxes per node, axis and bin (bvh.rs:349-358).
// bacc layout per slot: [axis][bin][7] = count, minx,miny,minz, maxx,maxy,maxz keys.
__global__ __launch_bounds__(kWG) void k_bins(const float4* __restrict__ ec, const float4* __restrict__ elo,
                                              const float2* __restrict__ ehi, const Chunk* __restrict__ chunks,
                                              const uint32_t* __restrict__ cacc, uint32_t* __restrict__ bacc,
                                              int faithful)
{
    __shared__ uint32_t sb[3 * kBins * 7];
    Chunk ch = chunks[blockIdx.x];
    for (uint32_t k = threadIdx.x; k < 3 * kBins * 7; k += kWG) {
        uint32_t f = k % 7;
        sb[k] = f == 0 ? 0u : (f <= 3 ? kEmptyMin : kEmptyMax);
    }
    AxisRange ar[3];
    for (int a = 0; a < 3; ++a)
        ar[a] = axis_range(cacc + 6 * ch.slot, a, faithful);
    __syncthreads();
    for (uint32_t e = ch.begin + threadIdx.x; e < ch.end; e += kWG) {
        float4 c = ec[e];
        float4 lo = elo[e];
        float2 hi = ehi[e];
        float cc[3] = {c.x, c.y, c.z};
        for (int a = 0; a < 3; ++a) {
            if (!ar[a].valid)
                continue;
            uint32_t b = bin_index((cc[a] - ar[a].bmin) * ar[a].scale);
            uint32_t* s = sb + (a * kBins + b) * 7;
            atomicAdd(s, 1u);
            if (not_nan(lo.x)) atomicMin(s + 1, f2key(lo.x));
            if (not_nan(lo.y)) atomicMin(s + 2, f2key(lo.y));
            if (not_nan(lo.z)) atomicMin(s + 3, f2key(lo.z));
            if (not_nan(lo.w)) atomicMax(s + 4, f2key(lo.w));
            if (not_nan(hi.x)) atomicMax(s + 5, f2key(hi.x));
            if (not_nan(hi.y)) atomicMax(s + 6, f2key(hi.y));
        }
    }
    __syncthreads();
    uint32_t* g = bacc + (size_t)ch.slot * (3 * kBins * 7);
    for (uint32_t k = threadIdx.x; k < 3 * kBins * 7; k += kWG) {
        uint32_t f = k % 7, v = sb[k];
        if (f == 0) {
            if (v)
                atomicAdd(g + k, v);
        } else if (f <= 3) {
            if (v != kEmptyMin)
                atomicMin(g + k, v);
        } else {
            if (v != kEmptyMax)
                atomicMax(g + k, v);
        }
    }
}

struct FBox {
    V3 lo, hi;
};

__device__ __forceinline__ float box_area(const FBox& b)
{
    V3 e = sub3(b.hi, b.lo);
    return e.x * e.y + e.y * e.z + e.z * e.x;
}
__device__ __forceinline__ void box_grow_box(FBox& b, const FBox& o)
{
    if (o.lo.x != kFltMax) {  // Aabb::grow_aabb (aabb.rs:41-46)
        b.lo = min3(b.lo, o.lo);
        b.hi = max3(b.hi, o.lo);
        b.lo = min3(b.lo, o.hi);
        b.hi = max3(b.hi, o.hi);
    }
}
__device__ __forceinline__ FBox box_init(int faithful)
{
    FBox b;
    if (faithful) {
        b.lo = mk3(0.0f, 0.0f, 0.0f);
        b.hi = mk3(0.0f, 0.0f, 0.0f);
    } else {
        b.lo = mk3(kFltMax, kFltMax, kFltMax);
        b.hi = mk3(-kFltMax, -kFltMax, -kFltMax);
    }
    return b;
}

// SAH sweep and plane choice for one axis (bvh.rs:360-390), given the bins.
__device__ __forceinline__ void sah_axis(const uint32_t counts[kBins], const FBox boxes[kBins], AxisRange ar, int axis,
                                         int faithful, int& best_axis, float& best_pos, float& best_cost)
{
    float left_area[kBins - 1], right_area[kBins - 1];
    uint32_t left_count[kBins - 1], right_count[kBins - 1];
    FBox lb = box_init(faithful), rb = box_init(faithful);
    uint32_t ls = 0, rs = 0;
    for (int i = 0; i < kBins - 1; ++i) {
        ls += counts[i];
        left_count[i] = ls;
        box_grow_box(lb, boxes[i]);
        left_area[i] = box_area(lb);
        rs += counts[kBins - 1 - i];
        right_count[kBins - 2 - i] = rs;
        box_grow_box(rb, boxes[kBins - 1 - i]);
        right_area[kBins - 2 - i] = box_area(rb);
    }
    float scale = (ar.bmax - ar.bmin) / (float)kBins;
    for (int i = 0; i < kBins - 1; ++i) {
        if (!faithful && (left_count[i] == 0 || right_count[i] == 0))
            continue;
        float cost = (float)left_count[i] * left_area[i] + (float)right_count[i] * right_area[i];
        if (cost < best_cost) {
            best_axis = axis;
            best_pos = ar.bmin + scale * (float)(i + 1);
            best_cost = cost;
        }
    }
}

// Decide each large node of the level (bvh.rs:405-417).
__global__ void k_decide(const Task* __restrict__ tasks, uint32_t ntasks, const uint32_t* __restrict__ cacc,
                         const uint32_t* __restrict__ bacc, Decision* __restrict__ dec, int faithful, uint32_t max_leaf)
{
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ntasks)
        return;
    Task t = tasks[s];
    int best_axis = -1;
    float best_pos = 0.0f, best_cost = kFltMax;
    for (int a = 0; a < 3; ++a) {
        AxisRange ar = axis_range(cacc + 6 * s, a, faithful);
        if (!ar.valid)
            continue;
        uint32_t counts[kBins];
        FBox boxes[kBins];
        const uint32_t* g = bacc + (size_t)s * (3 * kBins * 7) + a * kBins * 7;
        for (int b = 0; b < kBins; ++b) {
            counts[b] = g[b * 7];
            if (counts[b] == 0) {
                boxes[b] = box_init(faithful);
            } else {
                V3 lo = mk3(key2f(g[b * 7 + 1]), key2f(g[b * 7 + 2]), key2f(g[b * 7 + 3]));
                V3 hi = mk3(key2f(g[b * 7 + 4]), key2f(g[b * 7 + 5]), key2f(g[b * 7 + 6]));
                if (faithful) {  // the bin box starts as the zero box and grows
                    lo = min3(mk3(0.0f, 0.0f, 0.0f), lo);
                    hi = max3(mk3(0.0f, 0.0f, 0.0f), hi);
                }
                boxes[b].lo = lo;
                boxes[b].hi = hi;
            }
        }
        sah_axis(counts, boxes, ar, a, faithful, best_axis, best_pos, best_cost);
    }
    FBox nb;
    nb.lo = mk3(t.box[0], t.box[1], t.box[2]);
    nb.hi = mk3(t.box[3], t.box[4], t.box[5]);
    float no_split_cost = (float)t.count * box_area(nb);
    bool force = !faithful && max_leaf > 0 && t.count > max_leaf;
    Decision d;
    d.axis = best_axis;
    d.pos = best_pos;
    d.nl = 0;
    if (best_axis >= 0 && (!(best_cost >= no_split_cost) || force))
        d.kind = 1;
    else if (force)
        d.kind = 2;
    else
        d.kind = 0;
    dec[s] = d;
}

// Left counts per chunk for the plane splits.
__global__ __launch_bounds__(kWG) void k_part_count(const float4* __restrict__ ec, const Chunk* __restrict__ chunks,
                                                    const Decision* __restrict__ dec, uint32_t* __restrict__ chunk_left)
{
    __shared__ uint32_t wc[4];
    Chunk ch = chunks[blockIdx.x];
    Decision d = dec[ch.slot];
    if (d.kind != 1) {
        if (threadIdx.x == 0)
            chunk_left[blockIdx.x] = 0;
        return;
    }
    uint32_t cnt = 0;
    for (uint32_t e = ch.begin + threadIdx.x; e < ch.end; e += kWG) {
        float4 c = ec[e];
        float v = d.axis == 0 ? c.x : (d.axis == 1 ? c.y : c.z);
        cnt += (v < d.pos) ? 1u : 0u;
    }
    for (int o = 32; o > 0; o >>= 1)
        cnt += (uint32_t)__shfl_xor((int)cnt, o);
    if ((threadIdx.x & 63) == 0)
        wc[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0)
        chunk_left[blockIdx.x] = wc[0] + wc[1] + wc[2] + wc[3];
}

// Exclusive prefix of chunk left counts within each node; node totals.
__global__ void k_part_prefix(const uint32_t* __restrict__ node_chunk_begin, uint32_t ntasks,
                              uint32_t* __restrict__ chunk_left, Decision* __restrict__ dec)
{
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ntasks)
        return;
    uint32_t run = 0;
    for (uint32_t c = node_chunk_begin[s]; c < node_chunk_begin[s + 1]; ++c) {
        uint32_t v = chunk_left[c];
        chunk_left[c] = run;
        run += v;
    }
    dec[s].nl = run;
}

// Two-pointer partition (bvh.rs:420-429) as a permutation.  With n elements
// at relative positions k, nl of them "left" (centroid < split):
//   the examined sequence takes the next left-stream element after a left-goer
//   and the next right-stream element (from the top) after a right-goer, so
//   the left stream is [0, a) with a = nl if position nl is a left-goer (or
//   nl == n), else nl + 1, and the right stream is [a, n) read downwards.
//   Left-goers of the left stream stay put.  The m-th right-goer of the left
//   stream goes to top slot 0 (m = 0) or just below the (m-1)-th left-goer of
//   the right stream.  The m-th left-goer of the right stream (counted from
//   the top) takes the position of the m-th right-goer of the left stream.
//   Right-goers of the right stream move down by one.
// Phase A records, per node, kpos[m] (position of the m-th left-stream
// right-goer) and sel[m] (top-index of the m-th right-stream left-goer).
template <bool PHASE_B>
__global__ __launch_bounds__(kWG) void k_part_scatter(const float4* __restrict__ ec, const float4* __restrict__ elo,
                                                      const float2* __restrict__ ehi, const Chunk* __restrict__ chunks,
                                                      const Task* __restrict__ tasks, const Decision* __restrict__ dec,
                                                      const uint32_t* __restrict__ chunk_left,
                                                      uint32_t* __restrict__ kpos, uint32_t* __restrict__ sel,
                                                      float4* __restrict__ oc, float4* __restrict__ olo,
                                                      float2* __restrict__ ohi, uint32_t* __restrict__ err)
{
    __shared__ uint32_t wc[4];
    Chunk ch = chunks[blockIdx.x];
    Decision d = dec[ch.slot];
    if (d.kind == 0)
        return;
    Task t = tasks[ch.slot];
    uint32_t n = t.count, first = t.first, nl = d.nl;
    uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t a = n;
    if (d.kind == 1 && nl < n) {
        float4 c = ec[first + nl];
        float v = d.axis == 0 ? c.x : (d.axis == 1 ? c.y : c.z);
        a = (v < d.pos) ? nl : nl + 1;
    }
    uint32_t base = chunk_left[blockIdx.x];
    for (uint32_t e0 = ch.begin; e0 < ch.end; e0 += kWG) {
        uint32_t e = e0 + threadIdx.x;
        bool in = e < ch.end;
        bool f = false;
        float4 c, lo;
        float2 hi;
        if (in) {
            c = ec[e];
            if (d.kind == 1) {
                float v = d.axis == 0 ? c.x : (d.axis == 1 ? c.y : c.z);
                f = v < d.pos;
            }
        }
        uint64_t m = __ballot(f);
        if (lane == 0)
            wc[wave] = __popcll(m);
        __syncthreads();
        uint32_t lb = base + __popcll(m & ((1ull << lane) - 1ull));
        for (uint32_t w = 0; w < wave; ++w)
            lb += wc[w];
        uint32_t tot = wc[0] + wc[1] + wc[2] + wc[3];
        __syncthreads();
        base += tot;
        if (!in)
            continue;
        uint32_t k = e - first;
        if (d.kind == 2) {
            if (PHASE_B) {
                oc[e] = c;
                olo[e] = elo[e];
                ohi[e] = ehi[e];
            }
            continue;
        }
        // every index below is range-checked: a violation flags the build as
        // failed instead of touching memory outside the node's range
        if (!PHASE_B) {
            if (k < a && !f) {
                uint32_t mm = k - lb;
                if (mm < n)
                    kpos[first + mm] = k;
                else
                    atomicOr(err, 1u);
            } else if (k >= a && f) {
                uint32_t mm = nl - lb - 1;
                if (mm < n)
                    sel[first + mm] = n - 1 - k;
                else
                    atomicOr(err, 1u);
            }
            continue;
        }
        uint32_t dest = n;
        if (k < a) {
            if (f) {
                dest = k;
            } else {
                uint32_t mm = k - lb;
                if (mm == 0) {
                    dest = n - 1;
                } else if (mm - 1 < n) {
                    uint32_t sv = sel[first + mm - 1];
                    dest = sv + 2 <= n ? n - 2 - sv : n;
                }
            }
        } else {
            if (f) {
                uint32_t mm = nl - lb - 1;
                dest = mm < n ? kpos[first + mm] : n;
            } else {
                dest = k - 1;
            }
        }
        if (dest >= n) {
            atomicOr(err, 2u);
            continue;
        }
        lo = elo[e];
        hi = ehi[e];
        oc[first + dest] = c;
        olo[first + dest] = lo;
        ohi[first + dest] = hi;
    }
}

// Copy the partitioned ranges back and reduce the two child boxes
// (update_node_bounds of the children, bvh.rs:462-463).
__global__ __launch_bounds__(kWG) void k_part_commit(const float4* __restrict__ oc, const float4* __restrict__ olo,
                                                     const float2* __restrict__ ohi, const Chunk* __restrict__ chunks,
                                                     const Task* __restrict__ tasks, const Decision* __restrict__ dec,
                                                     float4* __restrict__ ec, float4* __restrict__ elo,
                                                     float2* __restrict__ ehi, uint32_t* __restrict__ child_keys)
{
    __shared__ uint32_t lds[12 * 4];
    Chunk ch = chunks[blockIdx.x];
    Decision d = dec[ch.slot];
    if (d.kind == 0)
        return;
    Task t = tasks[ch.slot];
    uint32_t nl = d.kind == 1 ? d.nl : t.count / 2;
    // a plane split that leaves one side empty is abandoned (bvh.rs:431-435),
    // or replaced by the median split when the node must be split
    if (d.kind == 1 && (nl == 0 || nl == t.count))
        nl = t.count / 2;
    uint32_t v[12];
    for (int k = 0; k < 12; ++k)
        v[k] = kEmptyMin;
    // layout: l.min xyz, r.min xyz (min) | l.max xyz, r.max xyz (max)
    for (int k = 6; k < 12; ++k)
        v[k] = kEmptyMax;
    for (uint32_t e = ch.begin + threadIdx.x; e < ch.end; e += kWG) {
        float4 c = oc[e];
        float4 lo = olo[e];
        float2 hi = ohi[e];
        ec[e] = c;
        elo[e] = lo;
        ehi[e] = hi;
        int side = (e - t.first) < nl ? 0 : 3;
        key_min(v[side + 0], lo.x);
        key_min(v[side + 1], lo.y);
        key_min(v[side + 2], lo.z);
        key_max(v[6 + side + 0], lo.w);
        key_max(v[6 + side + 1], hi.x);
        key_max(v[6 + side + 2], hi.y);
    }
    block_reduce_keys<12>(v, child_keys + 12 * ch.slot, lds);
}

struct SNode {  // same layout as bvht_node; internal nodes hold a local pair index
    float bmin[3];
    uint32_t lf;
    float bmax[3];
    uint32_t count;
};

__device__ __forceinline__ float comp4(float4 c, int axis) { return axis == 0 ? c.x : (axis == 1 ? c.y : c.z); }

// Sequential subdivide (bvh.rs:396-467) of one small subtree by one thread, on
// the element range [first, first+count); nodes go to out[0 .. 1+2*pairs),
// child pairs allocated in depth-first pre-order (pair p -> out[1+2p], out[2+2p]).
__global__ void k_small_build(float4* __restrict__ ec, float4* __restrict__ elo, float2* __restrict__ ehi,
                              const Task* __restrict__ roots, const uint32_t* __restrict__ out_base, uint32_t nroots,
                              SNode* __restrict__ snodes, uint32_t* __restrict__ pairs_out, int faithful,
                              uint32_t max_leaf, uint32_t* __restrict__ err)
{
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nroots)
        return;
    Task rt = roots[r];
    SNode* out = snodes + out_base[r];
    out[0] = SNode{{rt.box[0], rt.box[1], rt.box[2]}, rt.first, {rt.box[3], rt.box[4], rt.box[5]}, rt.count};
    uint32_t npairs = 0;
    uint32_t stack[kSmall + 2];
    uint32_t sp = 0;
    stack[sp++] = 0;
    while (sp) {
        uint32_t j = stack[--sp];
        SNode nd = out[j];
        uint32_t first = nd.lf, count = nd.count;
        // find_best_split_plane (bvh.rs:333-394)
        int best_axis = -1;
        float best_pos = 0.0f, best_cost = kFltMax;
        for (int a = 0; a < 3; ++a) {
            float bmin = faithful ? 1e30f : kFltMax, bmax = faithful ? 1e-30f : -kFltMax;
            for (uint32_t k = 0; k < count; ++k) {
                float c = comp4(ec[first + k], a);
                bmin = fminf(bmin, c);
                bmax = fmaxf(bmax, c);
            }
            if (bmin == bmax)
                continue;
            AxisRange ar;
            ar.bmin = bmin;
            ar.bmax = bmax;
            ar.scale = (float)kBins / (bmax - bmin);
            ar.valid = true;
            uint32_t counts[kBins];
            FBox boxes[kBins];
            for (int b = 0; b < kBins; ++b) {
                counts[b] = 0;
                boxes[b] = box_init(faithful);
            }
            for (uint32_t k = 0; k < count; ++k) {
                uint32_t b = bin_index((comp4(ec[first + k], a) - bmin) * ar.scale);
                float4 lo = elo[first + k];
                float2 hi = ehi[first + k];
                counts[b] += 1;
                boxes[b].lo = min3(boxes[b].lo, mk3(lo.x, lo.y, lo.z));
                boxes[b].hi = max3(boxes[b].hi, mk3(lo.w, hi.x, hi.y));
            }
            sah_axis(counts, boxes, ar, a, faithful, best_axis, best_pos, best_cost);
        }
        FBox nb;
        nb.lo = mk3(nd.bmin[0], nd.bmin[1], nd.bmin[2]);
        nb.hi = mk3(nd.bmax[0], nd.bmax[1], nd.bmax[2]);
        float no_split_cost = (float)count * box_area(nb);
        bool force = !faithful && max_leaf > 0 && count > max_leaf;
        uint32_t split;
        if (best_axis >= 0 && (!(best_cost >= no_split_cost) || force)) {
            int64_t i = first, jj = (int64_t)first + count - 1;
            while (i <= jj) {
                if (comp4(ec[i], best_axis) < best_pos) {
                    ++i;
                } else {
                    float4 c0 = ec[i], l0 = elo[i];
                    float2 h0 = ehi[i];
                    ec[i] = ec[jj];
                    elo[i] = elo[jj];
                    ehi[i] = ehi[jj];
                    ec[jj] = c0;
                    elo[jj] = l0;
                    ehi[jj] = h0;
                    --jj;
                }
            }
            split = (uint32_t)i;
        } else if (force) {
            split = first + count / 2;
        } else {
            continue;
        }
        uint32_t lc = split - first;
        if (lc == 0 || lc == count) {
            if (!force)
                continue;
            split = first + count / 2;
            lc = count / 2;
        }
        uint32_t p = npairs++;
        uint32_t l = 1 + 2 * p, rr = 2 + 2 * p;
        uint32_t ranges[2][2] = {{first, lc}, {split, count - lc}};
        for (int side = 0; side < 2; ++side) {
            V3 lo = mk3(kFltMax, kFltMax, kFltMax), hi = mk3(-kFltMax, -kFltMax, -kFltMax);
            for (uint32_t k = 0; k < ranges[side][1]; ++k) {
                float4 a = elo[ranges[side][0] + k];
                float2 b = ehi[ranges[side][0] + k];
                lo = min3(lo, mk3(a.x, a.y, a.z));
                hi = max3(hi, mk3(a.w, b.x, b.y));
            }
            out[side ? rr : l] = SNode{{lo.x, lo.y, lo.z}, ranges[side][0], {hi.x, hi.y, hi.z}, ranges[side][1]};
        }
        if (sp + 2 > kSmall + 2 || rr >= 2 * rt.count) {
            atomicOr(err, 4u);
            break;
        }
        out[j].lf = p;
        out[j].count = 0;
        stack[sp++] = rr;
        stack[sp++] = l;
    }
    pairs_out[r] = npairs;
}

// Place the small subtrees: local node 0 at root_final, local pair p at
// 2 + 2*(rank + p).  Internal nodes point at their pair's final index.
__global__ void k_small_emit(const SNode* __restrict__ snodes, const uint32_t* __restrict__ out_base,
                             const uint32_t* __restrict__ pairs, const uint32_t* __restrict__ root_final,
                             const uint32_t* __restrict__ rank, uint32_t nroots, bvht_node* __restrict__ nodes,
                             uint32_t node_count, uint32_t* __restrict__ err)
{
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nroots)
        return;
    const SNode* s = snodes + out_base[r];
    uint32_t np = pairs[r], rk = rank[r];
    for (uint32_t j = 0; j < 1 + 2 * np; ++j) {
        SNode sn = s[j];
        uint32_t fj = j == 0 ? root_final[r] : 2 + 2 * (rk + (j - 1) / 2) + (j - 1) % 2;
        bvht_node o;
        o.bmin[0] = sn.bmin[0]; o.bmin[1] = sn.bmin[1]; o.bmin[2] = sn.bmin[2];
        o.bmax[0] = sn.bmax[0]; o.bmax[1] = sn.bmax[1]; o.bmax[2] = sn.bmax[2];
        o.count = sn.count;
        o.left_first = sn.count > 0 ? sn.lf : 2 + 2 * (rk + sn.lf);
        if (fj >= node_count) {
            atomicOr(err, 8u);
            continue;
        }
        nodes[fj] = o;
    }
}

// Structural check of the finished node array (the same rules the upload
// path checks on the host): children after their parent and inside
// nodes_used, leaf ranges inside the triangle array.
__global__ void k_validate_tree(const bvht_node* __restrict__ nodes, uint32_t nodes_used, uint32_t n,
                                uint32_t* __restrict__ err)
{
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nodes_used || k == 1)
        return;
    bvht_node nd = nodes[k];
    if (nd.count > 0) {
        if ((uint64_t)nd.left_first + nd.count > n)
            atomicOr(err, 32u);
    } else if (nd.left_first <= k || nd.left_first == 1 || nd.left_first + 1 >= nodes_used) {
        atomicOr(err, 64u);
    }
}

__global__ void k_scatter_nodes(const bvht_node* __restrict__ src, const uint32_t* __restrict__ idx, uint32_t m,
                                bvht_node* __restrict__ nodes)
{
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < m)
        nodes[idx[k]] = src[k];
}

// Build order triangle array: tris_out[k] = tris_in[index of element k];
// perm[k] records that index.
__global__ void k_gather_tris(const float* __restrict__ tin, const float4* __restrict__ ec, uint32_t n,
                              float* __restrict__ tout, uint32_t* __restrict__ perm, uint32_t* __restrict__ err)
{
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n)
        return;
    uint32_t src = __float_as_uint(ec[k].w);
    if (src >= n) {
        atomicOr(err, 16u);
        return;
    }
    perm[k] = src;
    const float* a = tin + 9 * (size_t)src;
    float* b = tout + 9 * (size_t)k;
#pragma unroll
    for (int q = 0; q < 9; ++q)
        b[q] = a[q];
}

// ------------------------------------------------------------ host side --

struct HostNode {
    uint32_t first, count;
    float box[6];
    int32_t left = -1;        // host index of the left child (right = left + 1) for large internal nodes
    int32_t small = -1;       // small-root slot
};

inline uint32_t nblk(size_t n, uint32_t wg) { return (uint32_t)((n + wg - 1) / wg); }

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
    int32_t alloc(size_t bytes)
    {
        if (p) {
            (void)hipFree(p);
            p = nullptr;
        }
        if (hipMalloc(&p, bytes ? bytes : 4) != hipSuccess)
            return set_error(BVHT_ENOMEM, "hipMalloc(%zu) in the BLAS build failed", bytes);
        return BVHT_OK;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

int32_t fill(hipStream_t st, uint32_t* p, uint32_t v, size_t n)
{
    if (n == 0)
        return BVHT_OK;
    hipLaunchKernelGGL(k_fill, dim3(nblk(n, 256)), dim3(256), 0, st, p, v, n);
    BVHT_HIP(hipGetLastError());
    return BVHT_OK;
}

float key2f_host(uint32_t k)
{
    uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// keys {min xyz, max xyz} -> box floats; empty ranges give (+MAX, -MAX)
void keys_to_box(const uint32_t* k, float* box)
{
    for (int a = 0; a < 3; ++a) {
        box[a] = k[a] == kEmptyMin ? 3.402823466e+38f : key2f_host(k[a]);
        box[3 + a] = k[3 + a] == kEmptyMax ? -3.402823466e+38f : key2f_host(k[3 + a]);
    }
}

#define RC(x)                    \
    do {                         \
        int32_t _rc = (x);       \
        if (_rc)                 \
            return _rc;          \
    } while (0)

int32_t build_impl(bvht_ctx ctx, float* h_tris, uint32_t n, int faithful, uint32_t max_leaf, bvht_blas* out)
{
    hipStream_t st = ctx->stream;
    bvht_blas b = nullptr;
    RC(bvht::blas_alloc(ctx, n, 2 * n, &b));
    struct Guard {
        bvht_blas* b;
        ~Guard() { if (*b) bvht_blas_destroy(*b); }
    } guard{&b};

    DevBuf tin, ec, elo, ehi, oc, olo, ohi, kpos, sel, d_err;
    RC(d_err.alloc(sizeof(uint32_t)));
    RC(fill(st, d_err.as<uint32_t>(), 0u, 1));
    RC(tin.alloc(sizeof(float) * 9 * (size_t)n));
    RC(ec.alloc(sizeof(float4) * (size_t)n));
    RC(elo.alloc(sizeof(float4) * (size_t)n));
    RC(ehi.alloc(sizeof(float2) * (size_t)n));
    RC(oc.alloc(sizeof(float4) * (size_t)n));
    RC(olo.alloc(sizeof(float4) * (size_t)n));
    RC(ohi.alloc(sizeof(float2) * (size_t)n));
    RC(kpos.alloc(sizeof(uint32_t) * (size_t)n));
    RC(sel.alloc(sizeof(uint32_t) * (size_t)n));
    BVHT_HIP(hipMemcpyAsync(tin.p, h_tris, sizeof(float) * 9 * (size_t)n, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_init_elems, dim3(nblk(n, 256)), dim3(256), 0, st, tin.as<float>(), n, ec.as<float4>(),
                       elo.as<float4>(), ehi.as<float2>());
    BVHT_HIP(hipGetLastError());

    // root box: update_node_bounds over everything (bvh.rs:537)
    DevBuf rootk;
    RC(rootk.alloc(6 * sizeof(uint32_t)));
    RC(fill(st, rootk.as<uint32_t>(), kEmptyMin, 3));
    RC(fill(st, rootk.as<uint32_t>() + 3, kEmptyMax, 3));
    hipLaunchKernelGGL(k_range_box, dim3(std::min<uint32_t>(nblk(n, kWG), 1024)), dim3(kWG), 0, st, elo.as<float4>(),
                       ehi.as<float2>(), 0u, n, rootk.as<uint32_t>());
    BVHT_HIP(hipGetLastError());
    uint32_t rk[6];
    BVHT_HIP(hipMemcpyAsync(rk, rootk.p, sizeof rk, hipMemcpyDeviceToHost, st));
    BVHT_HIP(hipStreamSynchronize(st));

    std::vector<HostNode> tree;
    tree.reserve(1024);
    HostNode root;
    root.first = 0;
    root.count = n;
    keys_to_box(rk, root.box);
    tree.push_back(root);
    std::vector<uint32_t> level, small_roots;
    if (n > kSmall)
        level.push_back(0);
    else
        small_roots.push_back(0);

    DevBuf d_tasks, d_chunks, d_cacc, d_bacc, d_dec, d_chunk_left, d_node_chunk, d_child;
    size_t cap_tasks = 0, cap_chunks = 0;
    std::vector<Task> tasks;
    std::vector<Chunk> chunks;
    std::vector<uint32_t> node_chunk;
    std::vector<Decision> dec;
    std::vector<uint32_t> childk;
    while (!level.empty()) {
        uint32_t L = (uint32_t)level.size();
        tasks.resize(L);
        chunks.clear();
        node_chunk.assign(L + 1, 0);
        for (uint32_t s = 0; s < L; ++s) {
            const HostNode& hn = tree[level[s]];
            tasks[s].first = hn.first;
            tasks[s].count = hn.count;
            std::memcpy(tasks[s].box, hn.box, sizeof hn.box);
            node_chunk[s] = (uint32_t)chunks.size();
            for (uint32_t bgn = hn.first; bgn < hn.first + hn.count; bgn += kChunk)
                chunks.push_back(Chunk{s, bgn, std::min(bgn + kChunk, hn.first + hn.count), 0});
        }
        node_chunk[L] = (uint32_t)chunks.size();
        uint32_t C = (uint32_t)chunks.size();
        if (L > cap_tasks) {
            cap_tasks = std::max<size_t>(L, 2 * cap_tasks);
            RC(d_tasks.alloc(sizeof(Task) * cap_tasks));
            RC(d_cacc.alloc(sizeof(uint32_t) * 6 * cap_tasks));
            RC(d_bacc.alloc(sizeof(uint32_t) * 3 * kBins * 7 * cap_tasks));
            RC(d_dec.alloc(sizeof(Decision) * cap_tasks));
            RC(d_node_chunk.alloc(sizeof(uint32_t) * (cap_tasks + 1)));
            RC(d_child.alloc(sizeof(uint32_t) * 12 * cap_tasks));
        }
        if (C > cap_chunks) {
            cap_chunks = std::max<size_t>(C, 2 * cap_chunks);
            RC(d_chunks.alloc(sizeof(Chunk) * cap_chunks));
            RC(d_chunk_left.alloc(sizeof(uint32_t) * cap_chunks));
        }
        BVHT_HIP(hipMemcpyAsync(d_tasks.p, tasks.data(), sizeof(Task) * L, hipMemcpyHostToDevice, st));
        BVHT_HIP(hipMemcpyAsync(d_chunks.p, chunks.data(), sizeof(Chunk) * C, hipMemcpyHostToDevice, st));
        BVHT_HIP(hipMemcpyAsync(d_node_chunk.p, node_chunk.data(), sizeof(uint32_t) * (L + 1), hipMemcpyHostToDevice,
                                st));
        // accumulators: centroid ranges {min xyz | max xyz}, bins {count, min xyz, max xyz},
        // child boxes {l.min, r.min | l.max, r.max}
        hipLaunchKernelGGL(k_init_acc, dim3(nblk((size_t)3 * kBins * 7 * L, 256)), dim3(256), 0, st,
                           d_cacc.as<uint32_t>(), d_bacc.as<uint32_t>(), d_child.as<uint32_t>(), L);
        hipLaunchKernelGGL(k_centroid_range, dim3(C), dim3(kWG), 0, st, ec.as<float4>(), d_chunks.as<Chunk>(),
                           d_cacc.as<uint32_t>());
        hipLaunchKernelGGL(k_bins, dim3(C), dim3(kWG), 0, st, ec.as<float4>(), elo.as<float4>(), ehi.as<float2>(),
                           d_chunks.as<Chunk>(), d_cacc.as<uint32_t>(), d_bacc.as<uint32_t>(), faithful);
        hipLaunchKernelGGL(k_decide, dim3(nblk(L, 64)), dim3(64), 0, st, d_tasks.as<Task>(), L, d_cacc.as<uint32_t>(),
                           d_bacc.as<uint32_t>(), d_dec.as<Decision>(), faithful, max_leaf);
        hipLaunchKernelGGL(k_part_count, dim3(C), dim3(kWG), 0, st, ec.as<float4>(), d_chunks.as<Chunk>(),
                           d_dec.as<Decision>(), d_chunk_left.as<uint32_t>());
        hipLaunchKernelGGL(k_part_prefix, dim3(nblk(L, 64)), dim3(64), 0, st, d_node_chunk.as<uint32_t>(), L,
                           d_chunk_left.as<uint32_t>(), d_dec.as<Decision>());
        hipLaunchKernelGGL((k_part_scatter<false>), dim3(C), dim3(kWG), 0, st, ec.as<float4>(), elo.as<float4>(),
                           ehi.as<float2>(), d_chunks.as<Chunk>(), d_tasks.as<Task>(), d_dec.as<Decision>(),
                           d_chunk_left.as<uint32_t>(), kpos.as<uint32_t>(), sel.as<uint32_t>(), oc.as<float4>(),
                           olo.as<float4>(), ohi.as<float2>(), d_err.as<uint32_t>());
        hipLaunchKernelGGL((k_part_scatter<true>), dim3(C), dim3(kWG), 0, st, ec.as<float4>(), elo.as<float4>(),
                           ehi.as<float2>(), d_chunks.as<Chunk>(), d_tasks.as<Task>(), d_dec.as<Decision>(),
                           d_chunk_left.as<uint32_t>(), kpos.as<uint32_t>(), sel.as<uint32_t>(), oc.as<float4>(),
                           olo.as<float4>(), ohi.as<float2>(), d_err.as<uint32_t>());
        hipLaunchKernelGGL(k_part_commit, dim3(C), dim3(kWG), 0, st, oc.as<float4>(), olo.as<float4>(),
                           ohi.as<float2>(), d_chunks.as<Chunk>(), d_tasks.as<Task>(), d_dec.as<Decision>(),
                           ec.as<float4>(), elo.as<float4>(), ehi.as<float2>(), d_child.as<uint32_t>());
        BVHT_HIP(hipGetLastError());
        dec.resize(L);
        childk.resize((size_t)12 * L);
        BVHT_HIP(hipMemcpyAsync(dec.data(), d_dec.p, sizeof(Decision) * L, hipMemcpyDeviceToHost, st));
        BVHT_HIP(hipMemcpyAsync(childk.data(), d_child.p, sizeof(uint32_t) * 12 * L, hipMemcpyDeviceToHost, st));
        BVHT_HIP(hipStreamSynchronize(st));

        std::vector<uint32_t> next;
        for (uint32_t s = 0; s < L; ++s) {
            const Decision& d = dec[s];
            uint32_t hid = level[s];
            uint32_t cnt = tree[hid].count, first = tree[hid].first;
            if (d.kind == 0)
                continue;  // leaf
            uint32_t nl = d.kind == 1 ? d.nl : cnt / 2;
            if (d.kind == 1 && (nl == 0 || nl == cnt)) {
                bool force = !faithful && max_leaf > 0 && cnt > max_leaf;
                if (!force)
                    continue;  // abandoned split: stays a leaf, partition already applied
                nl = cnt / 2;
            }
            HostNode l, r;
            l.first = first;
            l.count = nl;
            r.first = first + nl;
            r.count = cnt - nl;
            uint32_t kk[6];
            const uint32_t* ck = &childk[(size_t)12 * s];
            kk[0] = ck[0]; kk[1] = ck[1]; kk[2] = ck[2]; kk[3] = ck[6]; kk[4] = ck[7]; kk[5] = ck[8];
            keys_to_box(kk, l.box);
            kk[0] = ck[3]; kk[1] = ck[4]; kk[2] = ck[5]; kk[3] = ck[9]; kk[4] = ck[10]; kk[5] = ck[11];
            keys_to_box(kk, r.box);
            int32_t li = (int32_t)tree.size();
            tree[hid].left = li;
            tree.push_back(l);
            tree.push_back(r);
            for (int32_t c = li; c <= li + 1; ++c) {
                if (tree[c].count > kSmall)
                    next.push_back((uint32_t)c);
                else
                    small_roots.push_back((uint32_t)c);
            }
        }
        level.swap(next);
    }

    // small subtrees
    uint32_t R = (uint32_t)small_roots.size();
    std::vector<Task> rtasks(R);
    std::vector<uint32_t> rbase(R);
    size_t total_local = 0;
    for (uint32_t r = 0; r < R; ++r) {
        const HostNode& hn = tree[small_roots[r]];
        tree[small_roots[r]].small = (int32_t)r;
        rtasks[r].first = hn.first;
        rtasks[r].count = hn.count;
        std::memcpy(rtasks[r].box, hn.box, sizeof hn.box);
        rbase[r] = (uint32_t)total_local;
        total_local += 2 * (size_t)hn.count;
    }
    DevBuf d_rt, d_rbase, d_snodes, d_pairs, d_rfinal, d_rrank;
    std::vector<uint32_t> pairs(R), rfinal(R), rrank(R);
    if (R) {
        RC(d_rt.alloc(sizeof(Task) * R));
        RC(d_rbase.alloc(sizeof(uint32_t) * R));
        RC(d_snodes.alloc(sizeof(SNode) * total_local));
        RC(d_pairs.alloc(sizeof(uint32_t) * R));
        RC(d_rfinal.alloc(sizeof(uint32_t) * R));
        RC(d_rrank.alloc(sizeof(uint32_t) * R));
        BVHT_HIP(hipMemcpyAsync(d_rt.p, rtasks.data(), sizeof(Task) * R, hipMemcpyHostToDevice, st));
        BVHT_HIP(hipMemcpyAsync(d_rbase.p, rbase.data(), sizeof(uint32_t) * R, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_small_build, dim3(nblk(R, 64)), dim3(64), 0, st, ec.as<float4>(), elo.as<float4>(),
                           ehi.as<float2>(), d_rt.as<Task>(), d_rbase.as<uint32_t>(), R, d_snodes.as<SNode>(),
                           d_pairs.as<uint32_t>(), faithful, max_leaf, d_err.as<uint32_t>());
        BVHT_HIP(hipGetLastError());
        BVHT_HIP(hipMemcpyAsync(pairs.data(), d_pairs.p, sizeof(uint32_t) * R, hipMemcpyDeviceToHost, st));
        BVHT_HIP(hipStreamSynchronize(st));
    }

    // internal counts (post-order) and pre-order ranks over the large tree
    size_t H = tree.size();
    std::vector<uint64_t> ic(H, 0);
    for (size_t h = H; h-- > 0;) {  // children always have larger host indices
        const HostNode& hn = tree[h];
        if (hn.small >= 0)
            ic[h] = pairs[hn.small];
        else if (hn.left >= 0)
            ic[h] = 1 + ic[hn.left] + ic[hn.left + 1];
    }
    uint64_t total_internal = ic[0];
    BVHT_REQUIRE(2 + 2 * total_internal <= 2 * (uint64_t)n, "internal node count overflow");
    std::vector<uint32_t> final_idx(H, 0);
    std::vector<bvht_node> big_nodes;
    std::vector<uint32_t> big_idx;
    {
        std::vector<uint32_t> stk;
        stk.push_back(0);
        final_idx[0] = 0;
        uint64_t counter = 0;
        while (!stk.empty()) {
            uint32_t h = stk.back();
            stk.pop_back();
            const HostNode& hn = tree[h];
            if (hn.small >= 0) {
                rfinal[hn.small] = final_idx[h];
                rrank[hn.small] = (uint32_t)counter;
                counter += pairs[hn.small];
                continue;
            }
            bvht_node o;
            std::memcpy(o.bmin, hn.box, 12);
            std::memcpy(o.bmax, hn.box + 3, 12);
            if (hn.left >= 0) {
                uint32_t r = (uint32_t)counter++;
                o.left_first = 2 + 2 * r;
                o.count = 0;
                final_idx[hn.left] = 2 + 2 * r;
                final_idx[hn.left + 1] = 3 + 2 * r;
                stk.push_back((uint32_t)hn.left + 1);
                stk.push_back((uint32_t)hn.left);
            } else {
                o.left_first = hn.first;
                o.count = hn.count;
            }
            if (final_idx[h] >= 2 * n)
                return set_error(BVHT_EDEVICE, "BLAS build: node index %u out of range", final_idx[h]);
            big_nodes.push_back(o);
            big_idx.push_back(final_idx[h]);
        }
    }
    b->nodes_used = (uint32_t)(2 + 2 * total_internal);
    BVHT_HIP(hipMemsetAsync(b->d_nodes, 0, sizeof(bvht_node) * 2 * (size_t)n, st));
    if (!big_nodes.empty()) {
        DevBuf d_big, d_bigi;
        RC(d_big.alloc(sizeof(bvht_node) * big_nodes.size()));
        RC(d_bigi.alloc(sizeof(uint32_t) * big_idx.size()));
        BVHT_HIP(hipMemcpyAsync(d_big.p, big_nodes.data(), sizeof(bvht_node) * big_nodes.size(), hipMemcpyHostToDevice,
                                st));
        BVHT_HIP(hipMemcpyAsync(d_bigi.p, big_idx.data(), sizeof(uint32_t) * big_idx.size(), hipMemcpyHostToDevice,
                                st));
        hipLaunchKernelGGL(k_scatter_nodes, dim3(nblk(big_nodes.size(), 256)), dim3(256), 0, st,
                           d_big.as<bvht_node>(), d_bigi.as<uint32_t>(), (uint32_t)big_nodes.size(), b->d_nodes);
        BVHT_HIP(hipGetLastError());
        BVHT_HIP(hipStreamSynchronize(st));
    }
    if (R) {
        BVHT_HIP(hipMemcpyAsync(d_rfinal.p, rfinal.data(), sizeof(uint32_t) * R, hipMemcpyHostToDevice, st));
        BVHT_HIP(hipMemcpyAsync(d_rrank.p, rrank.data(), sizeof(uint32_t) * R, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_small_emit, dim3(nblk(R, 64)), dim3(64), 0, st, d_snodes.as<SNode>(),
                           d_rbase.as<uint32_t>(), d_pairs.as<uint32_t>(), d_rfinal.as<uint32_t>(),
                           d_rrank.as<uint32_t>(), R, b->d_nodes, 2 * n, d_err.as<uint32_t>());
        BVHT_HIP(hipGetLastError());
    }
    if (hipMalloc(&b->d_perm, sizeof(uint32_t) * (size_t)n) != hipSuccess)
        return set_error(BVHT_ENOMEM, "device allocation of the %u-entry build permutation failed", n);
    hipLaunchKernelGGL(k_gather_tris, dim3(nblk(n, 256)), dim3(256), 0, st, tin.as<float>(), ec.as<float4>(), n,
                       b->d_tris, b->d_perm, d_err.as<uint32_t>());
    hipLaunchKernelGGL(k_validate_tree, dim3(nblk(b->nodes_used, 256)), dim3(256), 0, st, b->d_nodes, b->nodes_used,
                       n, d_err.as<uint32_t>());
    BVHT_HIP(hipGetLastError());
    uint32_t errv = 0;
    BVHT_HIP(hipMemcpyAsync(&errv, d_err.p, sizeof errv, hipMemcpyDeviceToHost, st));
    BVHT_HIP(hipStreamSynchronize(st));
    if (errv)
        return set_error(BVHT_EDEVICE, "BLAS build self-check failed (flags 0x%x): inconsistent partition or tree",
                         errv);
    BVHT_HIP(hipMemcpyAsync(h_tris, b->d_tris, sizeof(float) * 9 * (size_t)n, hipMemcpyDeviceToHost, st));
    BVHT_HIP(hipStreamSynchronize(st));
    RC(compute_blas_topology(b, nullptr));
    *out = b;
    b = nullptr;
    return BVHT_OK;
}

}  // namespace

extern "C" int32_t bvht_blas_build(bvht_ctx ctx, float* tris, uint32_t n, int32_t mode, uint32_t max_leaf,
                                   bvht_blas* out)
{
    BVHT_REQUIRE(ctx && tris && out, "NULL argument");
    BVHT_REQUIRE(n > 0, "empty triangle array (the reference panics on an empty mesh)");
    BVHT_REQUIRE(n <= (1u << 30), "too many triangles (%u)", n);
    BVHT_REQUIRE(mode == BVHT_BUILD_REFERENCE || mode == BVHT_BUILD_SAH, "unknown build mode %d", mode);
    return build_impl(ctx, tris, n, mode == BVHT_BUILD_REFERENCE ? 1 : 0, max_leaf, out);
}
