// bvht_build.hip -- GPU BLAS build (BvhBuilder::build_for, bvh.rs:509-541).
//
// The reference builds top-down, depth first, with a binned SAH of 8 bins
// and an in-place two-pointer partition.  The result (node array and triangle
// permutation) depends only on each node's own primitive range, so it can be
// produced breadth first:
//
//   * Large nodes (more than kSmall primitives) are processed a level at a
//     time.  Their primitive ranges are cut into chunks of 256 to 4096
//     elements, one workgroup per chunk; per-node reductions (centroid range, bin
//     counts and boxes, child boxes) combine chunk partials with atomics on
//     order-preserving float keys (min/max are exact, so the order of
//     combination does not matter).  The split decision is made by four
//     lanes per node (one per axis, combined in axis order) with the
//     reference's sequential arithmetic.  The
//     two-pointer partition permutation is reproduced exactly from prefix
//     counts (derivation at k_part_scatter).  The level loop runs on the
//     device: the node list of the next level, the chunks and every count
//     stay in device memory (Ctl), and the host reads the counts back only
//     every few levels to see whether any large node is left.
//   * Small nodes are finished by one wave each, running the reference's
//     sequential subdivide on their range.
//   * Node numbering: the reference allocates child pairs in depth-first
//     pre-order (bvh.rs:440-466), so the children of the internal node with
//     pre-order rank r are at 2+2r and 3+2r.  Ranks follow from the internal
//     counts of the subtrees once the tree is known (k_number, on the
//     device).
//   * Every buffer but the BLAS itself lives in a workspace the context keeps
//     between builds.
//
// Per element the build keeps its centroid (Triangle::centroid, exact), the
// min/max of its three vertices (growing a box by the three vertices equals
// growing it by their min and max) and its original index.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "bvht_device.hpp"
#include "bvht_internal.hpp"

using namespace bvht;

namespace {

constexpr uint32_t kSmall = 32;     // nodes with <= kSmall primitives are finished by one thread
// Elements per chunk (a workgroup's share) in the level kernels: a power of
// two from kChunkMin to kChunkMax, picked per build by chunk_elems.
constexpr uint32_t kChunkMin = 256;
constexpr uint32_t kChunkMax = 4096;
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

constexpr uint32_t kNone = 0xFFFFFFFFu;
// Per split node, the keys part_commit_chunk reduces for its two children:
// mins (box lo left, right; centroid min left, right) then maxes (box hi
// left, right; centroid max left, right), three each.
constexpr uint32_t kChildKeys = 24;

// A node of the large-node tree: every node split level by level, their
// children, and the small subtrees' roots.  Children are allocated in pairs
// after their parent, so a child's index is always larger.
struct BNode {
    uint32_t first, count;
    float box[6];
    uint32_t left;   // index of the left child (right = left + 1); kNone for a leaf
    uint32_t small;  // slot in the small-root list; kNone for a large node
};

// Build state the level kernels share on the device; the host reads it back
// only every few levels.
struct Ctl {
    uint32_t L;           // large nodes of the current level
    uint32_t C;           // chunks of the current level
    uint32_t H;           // BNodes allocated
    uint32_t R;           // small roots listed
    uint32_t lvl_off;     // first entry of the current level in the level list
    uint32_t nlevels;     // levels processed
    uint32_t err;         // self-check flags
    uint32_t nodes_used;  // 2 + 2 * internal nodes, after numbering
    uint32_t depth;       // levels of the finished tree, after placement
    uint32_t done;        // workgroups of k_level_end finished (the last resets it)
};

constexpr int kCtlWG = 1024;  // the single-workgroup level kernels

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

// Accumulators of the level's L large nodes (L read on the device).
__global__ void k_init_acc(const Ctl* __restrict__ ctl, uint32_t* __restrict__ bacc, uint32_t* __restrict__ child)
{
    size_t L = ctl->L;
    size_t total = (size_t)3 * kBins * 7 * L;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        if (i < kChildKeys * L)
            child[i] = (i % kChildKeys) < kChildKeys / 2 ? kEmptyMin : kEmptyMax;
        uint32_t f = (uint32_t)(i % 7);
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

// Centroid range of elements [0, n) into keys[6] = {min xyz, max xyz}: the
// root's (find_best_split_plane, bvh.rs:338-344).  The children's come out of
// part_commit_chunk, which reads every element of a split node anyway.
__global__ __launch_bounds__(kWG) void k_centroid_root(const float4* __restrict__ ec, uint32_t n,
                                                       uint32_t* __restrict__ keys)
{
    __shared__ uint32_t lds[6 * 4];
    uint32_t v[6] = {kEmptyMin, kEmptyMin, kEmptyMin, kEmptyMax, kEmptyMax, kEmptyMax};
    for (uint32_t e = blockIdx.x * kWG + threadIdx.x; e < n; e += gridDim.x * kWG) {
        float4 c = ec[e];
        key_min(v[0], c.x); key_min(v[1], c.y); key_min(v[2], c.z);
        key_max(v[3], c.x); key_max(v[4], c.y); key_max(v[5], c.z);
    }
    block_reduce_keys<6>(v, keys, lds);
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
// The current value of an LDS word that other lanes may be changing with
// atomics (a relaxed atomic load: no ordering, no tearing).
__device__ __forceinline__ uint32_t lds_peek(uint32_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__global__ __launch_bounds__(kWG) void k_bins(const float4* __restrict__ ec, const float4* __restrict__ elo,
                                              const float2* __restrict__ ehi, const Chunk* __restrict__ chunks,
                                              const Ctl* __restrict__ ctl, const uint32_t* __restrict__ cacc,
                                              uint32_t* __restrict__ bacc, int faithful)
{
    __shared__ uint32_t sb[3 * kBins * 7];
    uint32_t C = ctl->C;
    for (uint32_t ci = blockIdx.x; ci < C; ci += gridDim.x) {
        Chunk ch = chunks[ci];
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
                // a bound is only lowered (raised); a key that does not pass
                // the value read now cannot pass the final one, so its atomic
                // is skipped (a stale read only costs an extra atomic)
                uint32_t k;
                if (not_nan(lo.x) && (k = f2key(lo.x)) < lds_peek(s + 1)) atomicMin(s + 1, k);
                if (not_nan(lo.y) && (k = f2key(lo.y)) < lds_peek(s + 2)) atomicMin(s + 2, k);
                if (not_nan(lo.z) && (k = f2key(lo.z)) < lds_peek(s + 3)) atomicMin(s + 3, k);
                if (not_nan(lo.w) && (k = f2key(lo.w)) > lds_peek(s + 4)) atomicMax(s + 4, k);
                if (not_nan(hi.x) && (k = f2key(hi.x)) > lds_peek(s + 5)) atomicMax(s + 5, k);
                if (not_nan(hi.y) && (k = f2key(hi.y)) > lds_peek(s + 6)) atomicMax(s + 6, k);
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
        __syncthreads();  // sb is read above before the next chunk resets it
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

// The bins of node s on one axis, as sah_axis takes them.
__device__ __forceinline__ void load_axis_bins(const uint32_t* __restrict__ bacc, uint32_t s, int a, int faithful,
                                               uint32_t counts[kBins], FBox boxes[kBins])
{
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
}

// Decide each large node of the level (bvh.rs:405-417), four lanes per node:
// lane a < 3 runs the SAH sweep of axis a on its own (the first minimal
// plane of that axis), and the node's first lane takes the axes in order,
// each only when strictly cheaper than the ones before -- the sequential
// sweep's choice, since within sah_axis a plane replaces the best so far
// only when strictly cheaper.  One lane per node left most of the chip idle
// on the wide levels (20,000 nodes: 313 waves).
__global__ void k_decide(const Task* __restrict__ tasks, const Ctl* __restrict__ ctl, const uint32_t* __restrict__ cacc,
                         const uint32_t* __restrict__ bacc, Decision* __restrict__ dec, int faithful, uint32_t max_leaf)
{
    const uint32_t L = ctl->L;
    const uint32_t stride = gridDim.x * blockDim.x;
    // every lane of a wave runs the same number of rounds (the shuffles
    // below read lanes of the same quad)
    for (uint32_t q0 = blockIdx.x * blockDim.x; q0 < 4 * L; q0 += stride) {
        const uint32_t q = q0 + threadIdx.x;
        const uint32_t s = q >> 2;
        const int a = (int)(q & 3u);
        int axis = -1;
        float pos = 0.0f, cost = kFltMax;
        if (s < L && a < 3) {
            AxisRange ar = axis_range(cacc + 6 * s, a, faithful);
            if (ar.valid) {
                uint32_t counts[kBins];
                FBox boxes[kBins];
                load_axis_bins(bacc, s, a, faithful, counts, boxes);
                sah_axis(counts, boxes, ar, a, faithful, axis, pos, cost);
            }
        }
        const int ax1 = __shfl(axis, (int)threadIdx.x + 1, 64), ax2 = __shfl(axis, (int)threadIdx.x + 2, 64);
        const float c1 = __shfl(cost, (int)threadIdx.x + 1, 64), c2 = __shfl(cost, (int)threadIdx.x + 2, 64);
        const float p1 = __shfl(pos, (int)threadIdx.x + 1, 64), p2 = __shfl(pos, (int)threadIdx.x + 2, 64);
        if (s >= L || a != 0)
            continue;
        int best_axis = axis;
        float best_pos = pos, best_cost = cost;
        if (ax1 >= 0 && c1 < best_cost) {
            best_axis = ax1;
            best_pos = p1;
            best_cost = c1;
        }
        if (ax2 >= 0 && c2 < best_cost) {
            best_axis = ax2;
            best_pos = p2;
            best_cost = c2;
        }
        Task t = tasks[s];
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
}

// Left counts per chunk for the plane splits.
__global__ __launch_bounds__(kWG) void k_part_count(const float4* __restrict__ ec, const Chunk* __restrict__ chunks,
                                                    const Ctl* __restrict__ ctl, const Decision* __restrict__ dec,
                                                    uint32_t* __restrict__ chunk_left)
{
    __shared__ uint32_t wc[4];
    uint32_t C = ctl->C;
    for (uint32_t ci = blockIdx.x; ci < C; ci += gridDim.x) {
        Chunk ch = chunks[ci];
        Decision d = dec[ch.slot];
        if (d.kind != 1) {
            if (threadIdx.x == 0)
                chunk_left[ci] = 0;
            continue;
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
            chunk_left[ci] = wc[0] + wc[1] + wc[2] + wc[3];
        __syncthreads();  // wc is read above before the next chunk writes it
    }
}

// Exclusive prefix of chunk left counts within each node; node totals.
// Whether node s of the level splits, and its left count (the host-side rule
// of the reference's subdivide, bvh.rs:431-437, with the SAH mode's forced
// median split).
__device__ __forceinline__ bool node_splits(const Decision& d, uint32_t cnt, int faithful, uint32_t max_leaf,
                                            uint32_t& nl)
{
    if (d.kind == 0)
        return false;
    nl = d.kind == 1 ? d.nl : cnt / 2;
    if (d.kind == 1 && (nl == 0 || nl == cnt)) {
        bool force = !faithful && max_leaf > 0 && cnt > max_leaf;
        if (!force)
            return false;  // abandoned split: a leaf (its range stays partitioned)
        nl = cnt / 2;
    }
    return true;
}

// The level's nodes in kLvlWGs contiguous ranges, a multiple of kCtlWG
// nodes each (k_part_prefix counts and k_level_end places them the same way).
constexpr uint32_t kLvlWGs = 64;
__device__ __forceinline__ uint32_t level_range_len(uint32_t L)
{
    return ((L + kLvlWGs - 1) / kLvlWGs + kCtlWG - 1) / kCtlWG * kCtlWG;
}
__device__ __forceinline__ void level_range(uint32_t L, uint32_t& s0, uint32_t& s1)
{
    uint32_t per = level_range_len(L);
    s0 = min(L, blockIdx.x * per);
    s1 = min(L, s0 + per);
}

// Per node: the chunks' left counts to offsets and the node's left count;
// then whether it splits, and how many of its children are large, summed
// into the counts of its range of the level (wgsum, cleared by
// k_level_begin).  A wave's 64 consecutive nodes lie in one range (ranges
// are multiples of kCtlWG nodes), so one lane adds the wave's counts.
__global__ void k_part_prefix(const uint32_t* __restrict__ node_chunk_begin, const Ctl* __restrict__ ctl,
                              uint32_t* __restrict__ chunk_left, Decision* __restrict__ dec,
                              const Task* __restrict__ tasks, int faithful, uint32_t max_leaf,
                              uint32_t* __restrict__ wgsum)
{
    uint32_t L = ctl->L;
    uint32_t per = level_range_len(L);
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < L; s += gridDim.x * blockDim.x) {
        uint32_t run = 0;
        for (uint32_t c = node_chunk_begin[s]; c < node_chunk_begin[s + 1]; ++c) {
            uint32_t v = chunk_left[c];
            chunk_left[c] = run;
            run += v;
        }
        Decision d = dec[s];
        d.nl = run;
        dec[s].nl = run;
        uint32_t cnt = tasks[s].count, nl = 0;
        bool split = node_splits(d, cnt, faithful, max_leaf, nl);
        uint32_t big = split ? (nl > kSmall ? 1u : 0u) + (cnt - nl > kSmall ? 1u : 0u) : 0u;
        uint64_t ms = __builtin_amdgcn_ballot_w64(split);
        uint64_t m1 = __builtin_amdgcn_ballot_w64(big >= 1u), m2 = __builtin_amdgcn_ballot_w64(big == 2u);
        uint64_t act = __builtin_amdgcn_ballot_w64(true);
        if ((threadIdx.x & 63u) == (uint32_t)__builtin_ctzll(act)) {
            uint32_t r = s / per;
            if (ms)
                atomicAdd(&wgsum[2 * r], (uint32_t)__popcll(ms));
            if (m1)
                atomicAdd(&wgsum[2 * r + 1], (uint32_t)(__popcll(m1) + __popcll(m2)));
        }
    }
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
__device__ void part_scatter_chunk(const float4* __restrict__ ec, const float4* __restrict__ elo,
                                   const float2* __restrict__ ehi, Chunk ch, uint32_t base,
                                   const Task* __restrict__ tasks, const Decision* __restrict__ dec,
                                   uint32_t* __restrict__ kpos, uint32_t* __restrict__ sel, float4* __restrict__ oc,
                                   float4* __restrict__ olo, float2* __restrict__ ohi, uint32_t* __restrict__ err,
                                   uint32_t* wc)
{
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

template <bool PHASE_B>
__global__ __launch_bounds__(kWG) void k_part_scatter(const float4* __restrict__ ec, const float4* __restrict__ elo,
                                                      const float2* __restrict__ ehi, const Chunk* __restrict__ chunks,
                                                      const Ctl* __restrict__ ctl, const Task* __restrict__ tasks,
                                                      const Decision* __restrict__ dec,
                                                      const uint32_t* __restrict__ chunk_left,
                                                      uint32_t* __restrict__ kpos, uint32_t* __restrict__ sel,
                                                      float4* __restrict__ oc, float4* __restrict__ olo,
                                                      float2* __restrict__ ohi, uint32_t* __restrict__ err)
{
    __shared__ uint32_t wc[4];
    uint32_t C = ctl->C;
    for (uint32_t ci = blockIdx.x; ci < C; ci += gridDim.x)
        part_scatter_chunk<PHASE_B>(ec, elo, ehi, chunks[ci], chunk_left[ci], tasks, dec, kpos, sel, oc, olo, ohi,
                                    err, wc);
}

__device__ void part_commit_chunk(const float4* __restrict__ oc, const float4* __restrict__ olo,
                                  const float2* __restrict__ ohi, Chunk ch, Task t, Decision d,
                                  float4* __restrict__ ec, float4* __restrict__ elo, float2* __restrict__ ehi,
                                  uint32_t* __restrict__ child_keys, uint32_t* lds)
{
    uint32_t nl = d.kind == 1 ? d.nl : t.count / 2;
    // a plane split that leaves one side empty is abandoned (bvh.rs:431-435),
    // or replaced by the median split when the node must be split
    if (d.kind == 1 && (nl == 0 || nl == t.count))
        nl = t.count / 2;
    uint32_t v[kChildKeys];
    for (uint32_t k = 0; k < kChildKeys; ++k)
        v[k] = k < kChildKeys / 2 ? kEmptyMin : kEmptyMax;
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
        key_min(v[6 + side + 0], c.x);
        key_min(v[6 + side + 1], c.y);
        key_min(v[6 + side + 2], c.z);
        key_max(v[12 + side + 0], lo.w);
        key_max(v[12 + side + 1], hi.x);
        key_max(v[12 + side + 2], hi.y);
        key_max(v[18 + side + 0], c.x);
        key_max(v[18 + side + 1], c.y);
        key_max(v[18 + side + 2], c.z);
    }
    block_reduce_keys<kChildKeys>(v, child_keys + kChildKeys * ch.slot, lds);
}

// Copy the partitioned ranges back and reduce the two child boxes
// (update_node_bounds of the children, bvh.rs:462-463) and the children's
// centroid ranges (the next level's find_best_split_plane, bvh.rs:338-344).
__global__ __launch_bounds__(kWG) void k_part_commit(const float4* __restrict__ oc, const float4* __restrict__ olo,
                                                     const float2* __restrict__ ohi, const Chunk* __restrict__ chunks,
                                                     const Ctl* __restrict__ ctl, const Task* __restrict__ tasks,
                                                     const Decision* __restrict__ dec, float4* __restrict__ ec,
                                                     float4* __restrict__ elo, float2* __restrict__ ehi,
                                                     uint32_t* __restrict__ child_keys)
{
    __shared__ uint32_t lds[kChildKeys * 4];
    uint32_t C = ctl->C;
    for (uint32_t ci = blockIdx.x; ci < C; ci += gridDim.x) {
        Chunk ch = chunks[ci];
        Decision d = dec[ch.slot];
        if (d.kind != 0)
            part_commit_chunk(oc, olo, ohi, ch, tasks[ch.slot], d, ec, elo, ehi, child_keys, lds);
    }
}

struct SNode {  // same layout as bvht_node; internal nodes hold a local pair index
    float bmin[3];
    uint32_t lf;
    float bmax[3];
    uint32_t count;
};

// ------------------------------------------------ small subtrees, a wave each
//
// The reference's sequential subdivide (bvh.rs:396-467) of a subtree of at
// most kSmall elements, run by one wave: lane p holds the element at
// position first + p, lane j holds local node j, lane s holds stack entry s,
// and lane a*8+b gathers bin b of axis a.  Every value equals the sequential
// one: min/max are exact (the reduction order changes at most the sign of a
// zero), counts are integers, the SAH picks the first minimal cost in the
// sequential (axis, plane) order, and the two-pointer partition is applied
// as the permutation it amounts to (the closed form at k_part_scatter,
// looked up with ballots instead of the kpos/sel arrays).

__device__ __forceinline__ float rl(float v, uint32_t lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), (int)lane));
}
__device__ __forceinline__ uint32_t rl(uint32_t v, uint32_t lane)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)lane);
}
constexpr float kQNaN = __builtin_nanf("");
// fminf / fmaxf as one v_min_f32 / v_max_f32, for operands that are never a
// signalling NaN.  The compiler's fminf adds a canonicalize (v_max_f32 x, x)
// of each operand it cannot prove canonical; none of these is needed here:
// every value the small-subtree pass compares is a centroid or box bound made
// by arithmetic or by fminf / fmaxf (canonical), moved by DPP, readlane or
// ds_permute, or a constant (finite, or the quiet NaN that marks "nothing"),
// and on such operands v_min_f32 / v_max_f32 give fminf / fmaxf bit for bit.
// The _s forms take the first operand from an SGPR (a readlane result).
__device__ __forceinline__ float cmin(float a, float b)
{
    float r;
    asm("v_min_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ float cmax(float a, float b)
{
    float r;
    asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ float cmin_s(float s, float v)
{
    float r;
    asm("v_min_f32 %0, %1, %2" : "=v"(r) : "s"(s), "v"(v));
    return r;
}
__device__ __forceinline__ float cmax_s(float s, float v)
{
    float r;
    asm("v_max_f32 %0, %1, %2" : "=v"(r) : "s"(s), "v"(v));
    return r;
}
// cmin(a, dpp_f<CTRL>(b)) / cmax in one DPP instruction (row_shr:n is
// 0x110 + n, row_shl:n 0x100 + n; a source lane outside the row reads 0, as
// dpp_f gives).  The s_nop provides the two wait states between a VALU write
// of b and its DPP read.
#define BVHT_DPP_MINMAX(OP, CTRL)                                                                     \
    do {                                                                                              \
        if constexpr (CTRL == 0x111)                                                                  \
            asm("s_nop 1\n\t" OP "_dpp %0, %2, %1 row_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1"  \
                : "=v"(r) : "v"(a), "v"(b));                                                          \
        else if constexpr (CTRL == 0x112)                                                             \
            asm("s_nop 1\n\t" OP "_dpp %0, %2, %1 row_shr:2 row_mask:0xf bank_mask:0xf bound_ctrl:1"  \
                : "=v"(r) : "v"(a), "v"(b));                                                          \
        else if constexpr (CTRL == 0x114)                                                             \
            asm("s_nop 1\n\t" OP "_dpp %0, %2, %1 row_shr:4 row_mask:0xf bank_mask:0xf bound_ctrl:1"  \
                : "=v"(r) : "v"(a), "v"(b));                                                          \
        else if constexpr (CTRL == 0x118)                                                             \
            asm("s_nop 1\n\t" OP "_dpp %0, %2, %1 row_shr:8 row_mask:0xf bank_mask:0xf bound_ctrl:1"  \
                : "=v"(r) : "v"(a), "v"(b));                                                          \
        else if constexpr (CTRL == 0x101)                                                             \
            asm("s_nop 1\n\t" OP "_dpp %0, %2, %1 row_shl:1 row_mask:0xf bank_mask:0xf bound_ctrl:1"  \
                : "=v"(r) : "v"(a), "v"(b));                                                          \
        else if constexpr (CTRL == 0x102)                                                             \
            asm("s_nop 1\n\t" OP "_dpp %0, %2, %1 row_shl:2 row_mask:0xf bank_mask:0xf bound_ctrl:1"  \
                : "=v"(r) : "v"(a), "v"(b));                                                          \
        else                                                                                          \
        {                                                                                             \
            static_assert(CTRL == 0x104, "unsupported DPP control");                                  \
            asm("s_nop 1\n\t" OP "_dpp %0, %2, %1 row_shl:4 row_mask:0xf bank_mask:0xf bound_ctrl:1"  \
                : "=v"(r) : "v"(a), "v"(b));                                                          \
        }                                                                                             \
    } while (0)
template <int CTRL>
__device__ __forceinline__ float cmin_dpp(float a, float b)
{
    float r;
    BVHT_DPP_MINMAX("v_min_f32", CTRL);
    return r;
}
template <int CTRL>
__device__ __forceinline__ float cmax_dpp(float a, float b)
{
    float r;
    BVHT_DPP_MINMAX("v_max_f32", CTRL);
    return r;
}
#undef BVHT_DPP_MINMAX
// min / max over lanes 0..31, returned to every lane (lanes out of the
// node pass NaN, which fminf / fmaxf ignore): an inclusive scan within each
// 16-lane row, whose lane 15 then holds the row's result, and the two rows
// combined.  Lane 15's chain (15 <- 14, 13, 11, 7 <- ...) reads only lanes of
// its own row, so the zeros DPP gives lanes whose source is outside the row
// reach only lanes that are not read.
__device__ __forceinline__ float wmin32(float v)
{
    v = cmin_dpp<0x111>(v, v);
    v = cmin_dpp<0x112>(v, v);
    v = cmin_dpp<0x114>(v, v);
    v = cmin_dpp<0x118>(v, v);
    return fminf(rl(v, 15), rl(v, 31));
}
__device__ __forceinline__ float wmax32(float v)
{
    v = cmax_dpp<0x111>(v, v);
    v = cmax_dpp<0x112>(v, v);
    v = cmax_dpp<0x114>(v, v);
    v = cmax_dpp<0x118>(v, v);
    return fmaxf(rl(v, 15), rl(v, 31));
}
// One bin's (or a run of bins') count and its contribution to a box grown
// by it: Aabb::grow_aabb (aabb.rs:41-46) grows by both corners, so a box
// {lo, hi} adds min(lo, hi) to the minimum and max(lo, hi) to the maximum
// per component; a bin whose lo.x is still f32::MAX is not grown by (the
// sweep's check, as in sah_axis); NaN marks "nothing".
struct BinSum {
    uint32_t n;
    float lo[3], hi[3];
};
// One scan step over the bins of an axis: where `take`, v combines with the
// value DPP control CTRL brings (the callers take only sources in the same
// 8-lane axis group, so the zeros of out-of-row sources are never used).
template <int CTRL>
__device__ __forceinline__ void bin_scan_step(BinSum& v, bool take)
{
    const uint32_t n = v.n + dpp_u<CTRL>(v.n);
    float lo[3], hi[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        lo[q] = cmin_dpp<CTRL>(v.lo[q], v.lo[q]);
        hi[q] = cmax_dpp<CTRL>(v.hi[q], v.hi[q]);
    }
    if (take) {
        v.n = n;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            v.lo[q] = lo[q];
            v.hi[q] = hi[q];
        }
    }
}
template <int CTRL>
__device__ __forceinline__ BinSum dpp_bin(const BinSum& v)
{
    BinSum t;
    t.n = dpp_u<CTRL>(v.n);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        t.lo[q] = dpp_f<CTRL>(v.lo[q]);
        t.hi[q] = dpp_f<CTRL>(v.hi[q]);
    }
    return t;
}
// position of the m-th set bit (m < popcount) of a 32-bit mask
__device__ __forceinline__ uint32_t nth_set(uint32_t mask, uint32_t m)
{
    uint32_t pos = 0;
    for (uint32_t w = 16; w > 0; w >>= 1) {
        uint32_t c = __popc((mask >> pos) & ((1u << w) - 1u));
        if (m >= c) {
            m -= c;
            pos += w;
        }
    }
    return pos;
}
__device__ __forceinline__ float perm_f(uint32_t to, float v)
{
    return __int_as_float(__builtin_amdgcn_ds_permute((int)(to * 4), __float_as_int(v)));
}

constexpr int kSmallWavesPerWG = 4;

// 8 waves per SIMD (64 VGPRs, two dwords spilled) rather than the 7 its 67
// VGPRs allow: 1.3M build 6.81 -> 6.71 ms, reference mode 3.45 -> 3.35 ms.
__global__ __launch_bounds__(64 * kSmallWavesPerWG) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_small_wave(
    float4* __restrict__ ec, const float4* __restrict__ elo, const float2* __restrict__ ehi,
    const BNode* __restrict__ bn, const uint32_t* __restrict__ small_roots, uint32_t nroots,
    SNode* __restrict__ snodes, uint32_t* __restrict__ pairs_out, int faithful, uint32_t max_leaf,
    uint32_t* __restrict__ err)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t r = blockIdx.x * kSmallWavesPerWG + (threadIdx.x >> 6);
    if (r >= nroots)
        return;  // wave-uniform; no workgroup barrier below
    BNode rn = bn[small_roots[r]];
    const uint32_t first = rn.first, count = rn.count;
    if (count == 0 || count > kSmall) {
        if (lane == 0)
            atomicOr(err, 4u);
        return;
    }
    float4 c = make_float4(0.0f, 0.0f, 0.0f, 0.0f), lo = c;
    float2 hi = make_float2(0.0f, 0.0f);
    if (lane < count) {
        c = ec[first + lane];
        lo = elo[first + lane];
        hi = ehi[first + lane];
    }
    // node j in lane j: box, first (local; the pair index once internal), count
    float nb0 = 0.0f, nb1 = 0.0f, nb2 = 0.0f, nb3 = 0.0f, nb4 = 0.0f, nb5 = 0.0f;
    uint32_t nlf = 0, ncnt = 0;
    if (lane == 0) {
        nb0 = rn.box[0]; nb1 = rn.box[1]; nb2 = rn.box[2];
        nb3 = rn.box[3]; nb4 = rn.box[4]; nb5 = rn.box[5];
        ncnt = count;
    }
    uint32_t stk = 0;  // stack entry `lane`; entry 0 = node 0
    uint32_t sp = 1, npairs = 0;
    const float init_min = faithful ? 1e30f : kFltMax, init_max = faithful ? 1e-30f : -kFltMax;
    const float qnan = __int_as_float(0x7FC00000);
    const uint32_t ba = lane >> 3, bb = lane & 7u;  // bin lanes 0..23
    while (sp > 0) {
        const uint32_t j = rl(stk, sp - 1);
        --sp;
        const uint32_t fl = rl(nlf, j), cn = rl(ncnt, j);
        // one element: either no axis has a centroid range, or (the
        // reference mode's 1e30 / 1e-30 range) any split leaves a side empty,
        // and the partition of one element moves nothing: a leaf as it is
        if (cn <= 1)
            continue;
        const bool act = lane >= fl && lane < fl + cn;
        // find_best_split_plane (bvh.rs:333-394): centroid range per axis
        float bmin[3], bmax[3], ascale[3];
        bool valid[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float v = a == 0 ? c.x : (a == 1 ? c.y : c.z);
            bmin[a] = fminf(init_min, wmin32(act ? v : qnan));
            bmax[a] = fmaxf(init_max, wmax32(act ? v : qnan));
            valid[a] = !(bmin[a] == bmax[a]);
            ascale[a] = (float)kBins / (bmax[a] - bmin[a]);
        }
        // each element's bin on the valid axes, 4 bits per axis
        uint32_t pb = 0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float v = a == 0 ? c.x : (a == 1 ? c.y : c.z);
            if (valid[a])
                pb |= bin_index((v - bmin[a]) * ascale[a]) << (4 * a);
        }
        // bin (ba, bb): count and box over the node's elements
        uint32_t bcnt = 0;
        FBox bx = box_init(faithful);
        for (uint32_t k = fl; k < fl + cn; ++k) {
            uint32_t pk = rl(pb, k);
            V3 el = mk3(rl(lo.x, k), rl(lo.y, k), rl(lo.z, k));
            V3 eh = mk3(rl(lo.w, k), rl(hi.x, k), rl(hi.y, k));
            if (lane < 24 && ((pk >> (4 * ba)) & 15u) == bb) {
                bcnt += 1;
                bx.lo = mk3(cmin_s(el.x, bx.lo.x), cmin_s(el.y, bx.lo.y), cmin_s(el.z, bx.lo.z));
                bx.hi = mk3(cmax_s(eh.x, bx.hi.x), cmax_s(eh.y, bx.hi.y), cmax_s(eh.z, bx.hi.z));
            }
        }
        // SAH sweep (sah_axis): plane (ba, bb), bb < 7, has bins 0..bb on the
        // left and bb+1..7 on the right.  Bins of one axis are 8 adjacent
        // lanes inside a 16-lane row: an inclusive prefix and suffix scan over
        // each 8-lane group, then each plane takes the suffix of the next lane.
        BinSum own;
        own.n = bcnt;
        {
            bool grow = bx.lo.x != kFltMax;
            own.lo[0] = grow ? cmin(bx.lo.x, bx.hi.x) : kQNaN;
            own.lo[1] = grow ? cmin(bx.lo.y, bx.hi.y) : kQNaN;
            own.lo[2] = grow ? cmin(bx.lo.z, bx.hi.z) : kQNaN;
            own.hi[0] = grow ? cmax(bx.lo.x, bx.hi.x) : kQNaN;
            own.hi[1] = grow ? cmax(bx.lo.y, bx.hi.y) : kQNaN;
            own.hi[2] = grow ? cmax(bx.lo.z, bx.hi.z) : kQNaN;
        }
        BinSum pre = own, suf = own;
        bin_scan_step<0x111>(pre, bb >= 1);
        bin_scan_step<0x112>(pre, bb >= 2);
        bin_scan_step<0x114>(pre, bb >= 4);
        bin_scan_step<0x101>(suf, bb + 1 <= 7);
        bin_scan_step<0x102>(suf, bb + 2 <= 7);
        bin_scan_step<0x104>(suf, bb + 4 <= 7);
        BinSum rsum = dpp_bin<0x101>(suf);  // bins bb+1..7 (lanes with bb == 7 are not planes)
        FBox lbx = box_init(faithful), rbx = box_init(faithful);
        lbx.lo = mk3(fminf(lbx.lo.x, pre.lo[0]), fminf(lbx.lo.y, pre.lo[1]), fminf(lbx.lo.z, pre.lo[2]));
        lbx.hi = mk3(fmaxf(lbx.hi.x, pre.hi[0]), fmaxf(lbx.hi.y, pre.hi[1]), fmaxf(lbx.hi.z, pre.hi[2]));
        rbx.lo = mk3(fminf(rbx.lo.x, rsum.lo[0]), fminf(rbx.lo.y, rsum.lo[1]), fminf(rbx.lo.z, rsum.lo[2]));
        rbx.hi = mk3(fmaxf(rbx.hi.x, rsum.hi[0]), fmaxf(rbx.hi.y, rsum.hi[1]), fmaxf(rbx.hi.z, rsum.hi[2]));
        uint32_t lcnt = pre.n, rcnt = rsum.n;
        // this lane's axis, by value (no dynamically indexed arrays)
        bool avalid = ba == 0 ? valid[0] : (ba == 1 ? valid[1] : valid[2]);
        float amin = ba == 0 ? bmin[0] : (ba == 1 ? bmin[1] : bmin[2]);
        float amax = ba == 0 ? bmax[0] : (ba == 1 ? bmax[1] : bmax[2]);
        bool cand = lane < 24 && bb < kBins - 1 && avalid;
        if (!faithful && (lcnt == 0 || rcnt == 0))
            cand = false;
        float cost = (float)lcnt * box_area(lbx) + (float)rcnt * box_area(rbx);
        bool sel = cand && cost < kFltMax;
        float mcost = wmin32(sel ? cost : qnan);
        uint64_t win = __ballot(sel && cost == mcost);
        int best_axis = -1;
        float best_pos = 0.0f, best_cost = kFltMax;
        if (win) {
            uint32_t wl = (uint32_t)__builtin_ctzll(win);
            float scale = (amax - amin) / (float)kBins;
            float pa = amin + scale * (float)(bb + 1);
            best_axis = (int)(wl >> 3);
            best_pos = rl(pa, wl);
            best_cost = mcost;
        }
        FBox nbx;
        nbx.lo = mk3(rl(nb0, j), rl(nb1, j), rl(nb2, j));
        nbx.hi = mk3(rl(nb3, j), rl(nb4, j), rl(nb5, j));
        float no_split_cost = (float)cn * box_area(nbx);
        bool force = !faithful && max_leaf > 0 && cn > max_leaf;
        uint32_t lc;
        if (best_axis >= 0 && (!(best_cost >= no_split_cost) || force)) {
            // the two-pointer partition (bvh.rs:420-429) of positions [fl, fl + cn)
            float v = best_axis == 0 ? c.x : (best_axis == 1 ? c.y : c.z);
            bool f = act && v < best_pos;
            uint32_t k = lane - fl;
            uint32_t F = (uint32_t)(__ballot(f) >> fl);
            uint32_t nl = __popc(F);
            uint32_t a0 = nl == cn ? cn : (((F >> nl) & 1u) ? nl : nl + 1);
            uint32_t RL = (uint32_t)(__ballot(act && k < a0 && !f) >> fl);
            uint32_t LR = (uint32_t)(__ballot(act && k >= a0 && f) >> fl);
            uint32_t to = lane;
            if (act) {
                uint32_t lb = __popc(F & ((1u << k) - 1u));
                uint32_t dest;
                if (k < a0) {
                    uint32_t m = k - lb;
                    dest = f ? k : (m == 0 ? cn - 1 : nth_set(LR, __popc(LR) - m) - 1);
                } else {
                    dest = f ? nth_set(RL, nl - lb - 1) : k - 1;
                }
                to = fl + dest;
            }
            c = make_float4(perm_f(to, c.x), perm_f(to, c.y), perm_f(to, c.z), perm_f(to, c.w));
            lo = make_float4(perm_f(to, lo.x), perm_f(to, lo.y), perm_f(to, lo.z), perm_f(to, lo.w));
            hi = make_float2(perm_f(to, hi.x), perm_f(to, hi.y));
            lc = nl;
        } else if (force) {
            lc = cn / 2;
        } else {
            continue;
        }
        if (lc == 0 || lc == cn) {
            if (!force)
                continue;
            lc = cn / 2;
        }
        // children: boxes over each side (update_node_bounds, bvh.rs:462-463)
        bool left = act && lane - fl < lc, right = act && lane - fl >= lc;
        float cl[6], cr[6];
        float ev[6] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y};
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            cl[q] = fminf(kFltMax, wmin32(left ? ev[q] : qnan));
            cr[q] = fminf(kFltMax, wmin32(right ? ev[q] : qnan));
            cl[3 + q] = fmaxf(-kFltMax, wmax32(left ? ev[3 + q] : qnan));
            cr[3 + q] = fmaxf(-kFltMax, wmax32(right ? ev[3 + q] : qnan));
        }
        uint32_t p = npairs++;
        uint32_t l = 1 + 2 * p, rr = 2 + 2 * p;
        if (sp + 2 > kSmall + 2 || rr >= 2 * count) {
            if (lane == 0)
                atomicOr(err, 4u);
            break;
        }
        if (lane == l || lane == rr) {
            bool isl = lane == l;
            nb0 = isl ? cl[0] : cr[0]; nb1 = isl ? cl[1] : cr[1]; nb2 = isl ? cl[2] : cr[2];
            nb3 = isl ? cl[3] : cr[3]; nb4 = isl ? cl[4] : cr[4]; nb5 = isl ? cl[5] : cr[5];
            nlf = isl ? fl : fl + lc;
            ncnt = isl ? lc : cn - lc;
        }
        if (lane == j) {
            nlf = p;
            ncnt = 0;
        }
        if (lane == sp)
            stk = rr;
        if (lane == sp + 1)
            stk = l;
        sp += 2;
    }
    if (lane < count)
        ec[first + lane] = c;
    if (lane < 1 + 2 * npairs) {
        SNode o;
        o.bmin[0] = nb0; o.bmin[1] = nb1; o.bmin[2] = nb2;
        o.bmax[0] = nb3; o.bmax[1] = nb4; o.bmax[2] = nb5;
        o.count = ncnt;
        o.lf = ncnt > 0 ? first + nlf : nlf;
        snodes[2 * (size_t)first + lane] = o;
    }
    if (lane == 0)
        pairs_out[r] = npairs;
}

// Place the small subtrees: local node 0 at root_final, local pair p at
// 2 + 2*(rank + p).  Internal nodes point at their pair's final index.  Each
// node's level (its subtree root's level plus its depth in the subtree; a
// pair always comes after its parent) goes to `level`, and the deepest level
// + 1 to *depth.
__global__ __launch_bounds__(64) void k_small_emit(const SNode* __restrict__ snodes, const BNode* __restrict__ bn,
                                                   const uint32_t* __restrict__ small_roots,
                                                   const uint32_t* __restrict__ pairs,
                                                   const uint32_t* __restrict__ root_final,
                                                   const uint32_t* __restrict__ rank,
                                                   const uint32_t* __restrict__ root_level, uint32_t nroots,
                                                   bvht_node* __restrict__ nodes, uint32_t* __restrict__ level,
                                                   uint32_t node_count, uint32_t* __restrict__ err,
                                                   uint32_t* __restrict__ depth)
{
    __shared__ uint8_t ldep[64 * 2 * kSmall];  // depth of local node j of this lane's subtree
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t deepest = 0;  // deepest level + 1 of this lane's subtree
    if (r < nroots) {
        uint8_t* ld = ldep + threadIdx.x * 2 * kSmall;
        const SNode* s = snodes + 2 * (size_t)bn[small_roots[r]].first;
        uint32_t np = pairs[r], rk = rank[r], lv0 = root_level[r];
        ld[0] = 0;
        for (uint32_t j = 0; j < 1 + 2 * np && j < 2 * kSmall; ++j) {
            SNode sn = s[j];
            uint32_t fj = j == 0 ? root_final[r] : 2 + 2 * (rk + (j - 1) / 2) + (j - 1) % 2;
            uint32_t dj = ld[j];
            bvht_node o;
            o.bmin[0] = sn.bmin[0]; o.bmin[1] = sn.bmin[1]; o.bmin[2] = sn.bmin[2];
            o.bmax[0] = sn.bmax[0]; o.bmax[1] = sn.bmax[1]; o.bmax[2] = sn.bmax[2];
            o.count = sn.count;
            o.left_first = sn.count > 0 ? sn.lf : 2 + 2 * (rk + sn.lf);
            if (sn.count == 0) {
                uint32_t c0 = 1 + 2 * sn.lf;  // local index of the left child
                if (c0 <= j || c0 + 1 >= 2 * kSmall) {
                    atomicOr(err, 8u);
                    continue;
                }
                ld[c0] = (uint8_t)(dj + 1);
                ld[c0 + 1] = (uint8_t)(dj + 1);
            }
            if (fj >= node_count) {
                atomicOr(err, 8u);
                continue;
            }
            nodes[fj] = o;
            level[fj] = lv0 + dj;
            deepest = max(deepest, lv0 + dj + 1);
        }
    }
    // one atomic per wave (every lane takes part)
    for (int off = 32; off > 0; off >>= 1)
        deepest = max(deepest, (uint32_t)__shfl_xor((int)deepest, off, 64));
    if (threadIdx.x == 0 && deepest)
        atomicMax(depth, deepest);
}

// Structural check of the finished node array (the same rules the upload
// path checks on the host): children after their parent and inside
// nodes_used, leaf ranges inside the triangle array.
__global__ void k_validate_tree(const bvht_node* __restrict__ nodes, const Ctl* __restrict__ ctl, uint32_t n,
                                uint32_t* __restrict__ err)
{
    uint32_t nodes_used = ctl->nodes_used;
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

// ------------------------------------------------- device-driven levels --
//
// The level loop needs no host round trip: one workgroup turns the level's
// node list into tasks and chunks (k_level_begin), the chunk and task
// kernels run a fixed grid over the counts in Ctl, and one workgroup turns
// the decisions into the next level's list, the small roots and their
// BNodes (k_level_end).  A level with no node makes every kernel return at
// once, so the host enqueues several levels between read-backs.

// Exclusive prefix of v over a kCtlWG-thread workgroup in thread order;
// `total` gets the sum.  lds holds kCtlWG / 64 + 1 words.
// Inclusive prefix over each 16-lane row (DPP row shifts; a lane whose
// source is outside its row adds 0).
__device__ __forceinline__ uint32_t row_scan_incl(uint32_t x)
{
    x += dpp_u<0x111>(x);
    x += dpp_u<0x112>(x);
    x += dpp_u<0x114>(x);
    x += dpp_u<0x118>(x);
    return x;
}

__device__ uint32_t block_scan_excl(uint32_t v, uint32_t* lds, uint32_t& total)
{
    static_assert(kCtlWG / 64 <= 16, "one 16-lane row scans the wave totals");
    uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    // the wave's inclusive prefix: rows, then the carries of the rows before
    uint32_t incl = row_scan_incl(v);
    uint32_t r0 = (uint32_t)__builtin_amdgcn_readlane((int)incl, 15);
    uint32_t r1 = (uint32_t)__builtin_amdgcn_readlane((int)incl, 31);
    uint32_t r2 = (uint32_t)__builtin_amdgcn_readlane((int)incl, 47);
    incl += lane < 16u ? 0u : (lane < 32u ? r0 : (lane < 48u ? r0 + r1 : r0 + r1 + r2));
    if (lane == 63u)
        lds[wave] = incl;
    __syncthreads();
    if (wave == 0) {  // exclusive prefix of the wave totals in lanes 0..15
        uint32_t x = lane < (uint32_t)(kCtlWG / 64) ? lds[lane] : 0u;
        uint32_t xi = row_scan_incl(x);
        if (lane < (uint32_t)(kCtlWG / 64))
            lds[lane] = xi - x;
        if (lane == 15u)
            lds[kCtlWG / 64] = xi;
    }
    __syncthreads();
    uint32_t r = lds[wave] + incl - v;
    total = lds[kCtlWG / 64];
    __syncthreads();
    return r;
}

// keys {min xyz} and {max xyz} -> box floats; empty ranges give (+MAX, -MAX)
__device__ __forceinline__ void keys_to_box_dev(const uint32_t* kmin, const uint32_t* kmax, float* box)
{
    for (int a = 0; a < 3; ++a) {
        box[a] = kmin[a] == kEmptyMin ? kFltMax : key2f(kmin[a]);
        box[3 + a] = kmax[a] == kEmptyMax ? -kFltMax : key2f(kmax[a]);
    }
}

// BNode 0 = the root over all n elements (update_node_bounds, bvh.rs:537);
// a root of more than kSmall elements starts the level list, else it is the
// only small root.
__global__ void k_root_init(const uint32_t* __restrict__ rootk, uint32_t n, BNode* __restrict__ bn,
                            uint32_t* __restrict__ lvl_all, uint32_t* __restrict__ small_roots,
                            uint32_t* __restrict__ level_start, Ctl* __restrict__ ctl)
{
    BNode r;
    r.first = 0;
    r.count = n;
    keys_to_box_dev(rootk, rootk + 3, r.box);
    r.left = kNone;
    Ctl c;
    c.C = 0;
    c.H = 1;
    c.lvl_off = 0;
    c.nlevels = 0;
    c.err = 0;
    c.nodes_used = 0;
    c.depth = 0;
    c.done = 0;
    if (n > kSmall) {
        r.small = kNone;
        lvl_all[0] = 0;
        c.L = 1;
        c.R = 0;
    } else {
        r.small = 0;
        small_roots[0] = 0;
        c.L = 0;
        c.R = 1;
    }
    bn[0] = r;
    level_start[0] = 0;
    *ctl = c;
}

// Tasks of the level (node range and box) and their chunks of at most
// `chunk` elements, in node order; node_chunk[s] = the first chunk of node s.
__global__ __launch_bounds__(kCtlWG) void k_level_begin(const BNode* __restrict__ bn,
                                                        const uint32_t* __restrict__ lvl_all, Ctl* __restrict__ ctl,
                                                        Task* __restrict__ tasks, uint32_t* __restrict__ node_chunk,
                                                        Chunk* __restrict__ chunks, uint32_t lcap, uint32_t ccap,
                                                        uint32_t* __restrict__ wgsum, uint32_t chunk)
{
    __shared__ uint32_t lds[kCtlWG / 64 + 1];
    uint32_t L = ctl->L;
    if (L == 0 || L > lcap) {  // more than lcap cannot happen: the nodes' ranges are disjoint
        if (threadIdx.x == 0) {
            if (L)
                atomicOr(&ctl->err, 256u);
            ctl->C = 0;
            ctl->L = 0;
        }
        return;
    }
    if (threadIdx.x < 2 * kLvlWGs)  // the level's range counts (k_part_prefix adds to them)
        wgsum[threadIdx.x] = 0;
    const uint32_t* lvl = lvl_all + ctl->lvl_off;
    // rounds of kCtlWG consecutive nodes, one per thread (coalesced), the
    // chunk offsets carried from round to round
    uint32_t carry = 0;
    for (uint32_t r0 = 0; r0 < L; r0 += kCtlWG) {
        uint32_t s = r0 + threadIdx.x;
        Task t;
        uint32_t nch = 0;
        if (s < L) {
            BNode nd = bn[lvl[s]];
            t.first = nd.first;
            t.count = nd.count;
            for (int k = 0; k < 6; ++k)
                t.box[k] = nd.box[k];
            tasks[s] = t;
            nch = (nd.count + chunk - 1) / chunk;
        }
        uint32_t round_total;
        uint32_t run = carry + block_scan_excl(nch, lds, round_total);
        if (s < L) {
            node_chunk[s] = run;
            uint32_t end = t.first + t.count;
            for (uint32_t b = t.first; b < end; b += chunk) {
                if (run < ccap)
                    chunks[run] = Chunk{s, b, min(b + chunk, end), 0};
                ++run;
            }
        }
        carry += round_total;
    }
    const uint32_t total = carry;
    if (threadIdx.x == 0) {
        node_chunk[L] = total;
        if (total > ccap) {  // cannot happen: chunks <= n / chunk + L
            atomicOr(&ctl->err, 256u);
            ctl->C = 0;
            ctl->L = 0;
        } else {
            ctl->C = total;
        }
    }
}

// Children of the level's split nodes: a BNode pair each (in node order),
// the large ones listed as the next level, the small ones as small roots.
// Each workgroup places its range from the counts of the ranges before it;
// the last workgroup to finish (every other one has read ctl by then)
// advances ctl to the next level.
__global__ __launch_bounds__(kCtlWG) void k_level_end(BNode* __restrict__ bn, uint32_t* __restrict__ lvl_all,
                                                      uint32_t* __restrict__ small_roots,
                                                      uint32_t* __restrict__ level_start,
                                                      const Task* __restrict__ tasks,
                                                      const Decision* __restrict__ dec,
                                                      const uint32_t* __restrict__ child_keys, Ctl* __restrict__ ctl,
                                                      const uint32_t* __restrict__ wgsum, int faithful,
                                                      uint32_t max_leaf, uint32_t hcap, uint32_t rcap,
                                                      uint32_t* __restrict__ cacc, uint32_t lcap)
{
    __shared__ uint32_t lds[kCtlWG / 64 + 1];
    __shared__ uint32_t base[4];  // splits and large children before this range; level totals
    uint32_t L = ctl->L;
    if (L == 0)
        return;
    uint32_t H0 = ctl->H, R0 = ctl->R, off = ctl->lvl_off, nlev = ctl->nlevels;
    const uint32_t* lvl = lvl_all + off;
    if (threadIdx.x < 64) {  // one wave scans the kLvlWGs range counts
        uint32_t lane = threadIdx.x;
        uint32_t xs = lane < kLvlWGs ? wgsum[2 * lane] : 0u, xl = lane < kLvlWGs ? wgsum[2 * lane + 1] : 0u;
        uint32_t is = row_scan_incl(xs), il = row_scan_incl(xl);
        uint32_t cs0 = (uint32_t)__builtin_amdgcn_readlane((int)is, 15), cs1 = (uint32_t)__builtin_amdgcn_readlane((int)is, 31),
                 cs2 = (uint32_t)__builtin_amdgcn_readlane((int)is, 47);
        uint32_t cl0 = (uint32_t)__builtin_amdgcn_readlane((int)il, 15), cl1 = (uint32_t)__builtin_amdgcn_readlane((int)il, 31),
                 cl2 = (uint32_t)__builtin_amdgcn_readlane((int)il, 47);
        is += lane < 16u ? 0u : (lane < 32u ? cs0 : (lane < 48u ? cs0 + cs1 : cs0 + cs1 + cs2));
        il += lane < 16u ? 0u : (lane < 32u ? cl0 : (lane < 48u ? cl0 + cl1 : cl0 + cl1 + cl2));
        if (lane == blockIdx.x) {
            base[0] = is - xs;
            base[1] = il - xl;
        }
        if (lane == 63u) {
            base[2] = is;
            base[3] = il;
        }
    }
    __syncthreads();
    const uint32_t tsplit = base[2], tlarge = base[3];
    const uint32_t tsmall = 2 * tsplit - tlarge;
    bool fits = (uint64_t)H0 + 2 * (uint64_t)tsplit <= hcap && (uint64_t)off + L + tlarge <= hcap &&
                (uint64_t)R0 + tsmall <= rcap;
    if (fits) {  // cannot fail: the tree has fewer than 2n nodes, n elements
        uint32_t s0, s1;
        level_range(L, s0, s1);
        // rounds of kCtlWG consecutive nodes, one per thread, placed in node
        // order by a scan of (splits, large children) packed in one word (at
        // most kCtlWG and 2 kCtlWG per round)
        uint32_t c_split = base[0], c_large = base[1];
        for (uint32_t r0 = s0; r0 < s1; r0 += kCtlWG) {
            uint32_t s = r0 + threadIdx.x;
            Task t;
            uint32_t nl = 0, big = 0;
            bool split = false;
            if (s < s1) {
                t = tasks[s];
                split = node_splits(dec[s], t.count, faithful, max_leaf, nl);
                if (split)
                    big = (nl > kSmall ? 1u : 0u) + (t.count - nl > kSmall ? 1u : 0u);
            }
            uint32_t round_total;
            uint32_t ex = block_scan_excl((split ? 1u : 0u) | (big << 16), lds, round_total);
            uint32_t esplit = c_split + (ex & 0xFFFFu), elarge = c_large + (ex >> 16);
            if (split) {
                uint32_t child = H0 + 2 * esplit, lpos = off + L + elarge;
                uint32_t rpos = R0 + 2 * esplit - elarge;  // small children before this node
                bn[lvl[s]].left = child;
                const uint32_t* ck = child_keys + kChildKeys * (size_t)s;
                for (uint32_t side = 0; side < 2; ++side) {
                    BNode c;
                    c.first = side ? t.first + nl : t.first;
                    c.count = side ? t.count - nl : nl;
                    keys_to_box_dev(ck + 3 * side, ck + 12 + 3 * side, c.box);
                    c.left = kNone;
                    if (c.count > kSmall) {
                        c.small = kNone;
                        // its centroid range, as the next level's slot
                        // lpos - (off + L) reads it
                        const uint32_t ns = lpos - (off + L);
                        if (ns < lcap)
                            for (uint32_t k = 0; k < 3; ++k) {
                                cacc[6 * (size_t)ns + k] = ck[6 + 3 * side + k];
                                cacc[6 * (size_t)ns + 3 + k] = ck[18 + 3 * side + k];
                            }
                        lvl_all[lpos++] = child + side;
                    } else {
                        c.small = rpos;
                        small_roots[rpos++] = child + side;
                    }
                    bn[child + side] = c;
                }
            }
            c_split += round_total & 0xFFFFu;
            c_large += round_total >> 16;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&ctl->done, 1u) == gridDim.x - 1) {  // the last workgroup
            __threadfence();
            if (fits) {
                ctl->H = H0 + 2 * tsplit;
                ctl->R = R0 + tsmall;
                ctl->lvl_off = off + L;
                ctl->L = tlarge;
                ctl->nlevels = nlev + 1;
                level_start[nlev + 1] = off + L;
            } else {
                atomicOr(&ctl->err, 128u);
                ctl->L = 0;
            }
            ctl->done = 0;
        }
    }
}

// Node numbering of the reference's depth-first pair allocation
// (bvh.rs:440-466): pass 1, deepest level first, counts each large node's
// internal nodes (a small root's are its pairs); pass 2, from the root,
// gives each internal node its pre-order rank r (children at 2 + 2r and
// 3 + 2r) and each child the rank its subtree starts at, writes the large
// nodes and hands the small roots their index and starting rank.
__global__ __launch_bounds__(kCtlWG) void k_number(const BNode* __restrict__ bn, const uint32_t* __restrict__ lvl_all,
                                                   const uint32_t* __restrict__ level_start, Ctl* __restrict__ ctl,
                                                   const uint32_t* __restrict__ pairs, uint32_t* __restrict__ ic,
                                                   uint32_t* __restrict__ start, uint32_t* __restrict__ fin,
                                                   uint32_t* __restrict__ rfinal, uint32_t* __restrict__ rrank,
                                                   uint32_t* __restrict__ rlevel, bvht_node* __restrict__ nodes,
                                                   uint32_t* __restrict__ level, uint32_t node_count)
{
    uint32_t D = ctl->nlevels;
    if (bn[0].small != kNone) {  // n <= kSmall: the root is the one small subtree
        if (threadIdx.x == 0) {
            rfinal[0] = 0;
            rrank[0] = 0;
            rlevel[0] = 0;
            level[1] = kNone;  // the unused node
            uint32_t total = pairs[0];
            if ((uint64_t)2 + 2 * (uint64_t)total > node_count)
                atomicOr(&ctl->err, 512u);
            else
                ctl->nodes_used = 2 + 2 * total;
        }
        return;
    }
    for (uint32_t l = D; l-- > 0;) {
        uint32_t a = level_start[l], b = level_start[l + 1];
        for (uint32_t s = a + threadIdx.x; s < b; s += kCtlWG) {
            uint32_t h = lvl_all[s];
            BNode nd = bn[h];
            uint32_t v = 0;
            if (nd.left != kNone) {
                uint32_t c0 = nd.left, c1 = nd.left + 1;
                uint32_t m0 = bn[c0].small, m1 = bn[c1].small;
                v = 1 + (m0 != kNone ? pairs[m0] : ic[c0]) + (m1 != kNone ? pairs[m1] : ic[c1]);
            }
            ic[h] = v;
        }
        __syncthreads();
    }
    uint32_t total = ic[0];
    if ((uint64_t)2 + 2 * (uint64_t)total > node_count) {
        if (threadIdx.x == 0)
            atomicOr(&ctl->err, 512u);
        return;
    }
    if (threadIdx.x == 0) {
        start[0] = 0;
        fin[0] = 0;
    }
    __syncthreads();
    for (uint32_t l = 0; l < D; ++l) {
        uint32_t a = level_start[l], b = level_start[l + 1];
        for (uint32_t s = a + threadIdx.x; s < b; s += kCtlWG) {
            uint32_t h = lvl_all[s];
            BNode nd = bn[h];
            uint32_t f = fin[h], r = start[h];
            bvht_node o;
            o.bmin[0] = nd.box[0];
            o.bmin[1] = nd.box[1];
            o.bmin[2] = nd.box[2];
            o.bmax[0] = nd.box[3];
            o.bmax[1] = nd.box[4];
            o.bmax[2] = nd.box[5];
            if (nd.left != kNone) {
                o.left_first = 2 + 2 * r;
                o.count = 0;
                uint32_t c0 = nd.left, c1 = nd.left + 1;
                uint32_t m0 = bn[c0].small, m1 = bn[c1].small;
                uint32_t r0 = r + 1, r1 = r + 1 + (m0 != kNone ? pairs[m0] : ic[c0]);
                fin[c0] = 2 + 2 * r;
                fin[c1] = 3 + 2 * r;
                start[c0] = r0;
                start[c1] = r1;
                if (m0 != kNone) {
                    rfinal[m0] = 2 + 2 * r;
                    rrank[m0] = r0;
                    rlevel[m0] = l + 1;
                }
                if (m1 != kNone) {
                    rfinal[m1] = 3 + 2 * r;
                    rrank[m1] = r1;
                    rlevel[m1] = l + 1;
                }
            } else {
                o.left_first = nd.first;
                o.count = nd.count;
            }
            if (f < node_count) {
                nodes[f] = o;
                level[f] = l;
            } else {
                atomicOr(&ctl->err, 512u);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        ctl->nodes_used = 2 + 2 * total;
        level[1] = kNone;  // the unused node
        atomicMax(&ctl->depth, D);
    }
}

// ------------------------------------------------------------ host side --

inline uint32_t nblk(size_t n, uint32_t wg) { return (uint32_t)((n + wg - 1) / wg); }

int32_t fill(hipStream_t st, uint32_t* p, uint32_t v, size_t n)
{
    if (n == 0)
        return BVHT_OK;
    hipLaunchKernelGGL(k_fill, dim3(nblk(n, 256)), dim3(256), 0, st, p, v, n);
    BVHT_HIP(hipGetLastError());
    return BVHT_OK;
}

#define RC(x)                    \
    do {                         \
        int32_t _rc = (x);       \
        if (_rc)                 \
            return _rc;          \
    } while (0)

// Device workspace of the build, owned by the context and kept between
// builds (one allocation, grown when a larger mesh comes): a build then
// allocates nothing but the BLAS itself.  Capacities for n_cap elements:
//   per element   input triangles, element records (two copies), partition
//                 positions, small-subtree nodes (2 per element), small roots
//   per BNode     2 n_cap + 1 (the tree has fewer than 2n nodes): the nodes,
//                 the level list, internal counts, ranks, final indices,
//                 level starts
//   per task      n_cap / 33 + 1 (a level's large nodes have disjoint ranges
//                 of more than 32 elements)
//   per chunk     n_cap / kChunkMin + tasks + 2
struct BuildWs {
    void* base = nullptr;
    size_t n_cap = 0;
    uint32_t hcap = 0, rcap = 0, lcap = 0, ccap = 0;
    float* tin;
    float4 *ec, *elo, *oc, *olo;
    float2 *ehi, *ohi;
    uint32_t *kpos, *sel;
    SNode* snodes;
    uint32_t *small_roots, *pairs, *rfinal, *rrank, *rlevel;
    BNode* bn;
    uint32_t *lvl_all, *ic, *start, *fin, *level_start;
    Task* tasks;
    uint32_t *cacc, *bacc, *child, *node_chunk;
    Decision* dec;
    Chunk* chunks;
    uint32_t* chunk_left;
    Ctl* ctl;
    uint32_t* rootk;
    uint32_t* wgsum;  // per-range (splits, large children) of the level
};

struct Carver {
    char* base;
    size_t off = 0;
    template <class T>
    void operator()(T*& out, size_t count)
    {
        if (base)
            out = reinterpret_cast<T*>(base + off);
        off += (sizeof(T) * count + 255) & ~(size_t)255;
    }
};

// Lays the buffers out from `base` (or, with base NULL, only adds up the
// size); returns the bytes used.
size_t layout(BuildWs& w, char* base)
{
    Carver carve{base};
    size_t n = w.n_cap;
    carve(w.tin, 9 * n);
    carve(w.ec, n);
    carve(w.elo, n);
    carve(w.ehi, n);
    carve(w.oc, n);
    carve(w.olo, n);
    carve(w.ohi, n);
    carve(w.kpos, n);
    carve(w.sel, n);
    carve(w.snodes, 2 * n);
    carve(w.small_roots, w.rcap);
    carve(w.pairs, w.rcap);
    carve(w.rfinal, w.rcap);
    carve(w.rrank, w.rcap);
    carve(w.rlevel, w.rcap);
    carve(w.bn, w.hcap);
    carve(w.lvl_all, w.hcap);
    carve(w.ic, w.hcap);
    carve(w.start, w.hcap);
    carve(w.fin, w.hcap);
    carve(w.level_start, (size_t)w.hcap + 1);
    carve(w.tasks, w.lcap);
    carve(w.cacc, 6 * (size_t)w.lcap);
    carve(w.bacc, (size_t)3 * kBins * 7 * w.lcap);
    carve(w.child, kChildKeys * (size_t)w.lcap);
    carve(w.node_chunk, (size_t)w.lcap + 1);
    carve(w.dec, w.lcap);
    carve(w.chunks, w.ccap);
    carve(w.chunk_left, w.ccap);
    carve(w.ctl, 1);
    carve(w.wgsum, 2 * (size_t)kLvlWGs);
    carve(w.rootk, 6);
    return carve.off;
}

int32_t ensure_ws(bvht_ctx ctx, uint32_t n, BuildWs** out)
{
    BuildWs* w = static_cast<BuildWs*>(ctx->build_ws);
    if (!w) {
        w = new (std::nothrow) BuildWs();
        if (!w)
            return set_error(BVHT_ENOMEM, "host allocation failed");
        ctx->build_ws = w;
    }
    if (w->n_cap < n) {
        BuildWs g;
        g.n_cap = std::max<size_t>(n, w->n_cap + w->n_cap / 2);
        g.n_cap = std::min<size_t>(g.n_cap, (size_t)1 << 30);
        BVHT_HIP(hipStreamSynchronize(ctx->stream));
        if (w->base)
            (void)hipFree(w->base);
        w->base = nullptr;
        w->n_cap = 0;
        g.hcap = (uint32_t)(2 * g.n_cap + 1);
        g.rcap = (uint32_t)g.n_cap;
        g.lcap = (uint32_t)(g.n_cap / (kSmall + 1) + 1);
        g.ccap = (uint32_t)(g.n_cap / kChunkMin + g.lcap + 2);
        size_t bytes = layout(g, nullptr);
        if (hipMalloc(&g.base, bytes) != hipSuccess)
            return set_error(BVHT_ENOMEM, "hipMalloc(%zu) of the BLAS build workspace failed", bytes);
        layout(g, static_cast<char*>(g.base));
        *w = g;
    }
    *out = w;
    return BVHT_OK;
}

constexpr uint32_t kChunkGrid = 2048;  // workgroups of the chunk kernels
constexpr uint32_t kTaskGrid = 512;    // 64-thread workgroups of the per-node kernels

// Chunk size for n elements: the power of two at or below 2 sqrt(n), within
// [kChunkMin, kChunkMax] -- so about sqrt(n) / 2 chunks over the root level.
// Larger chunks mean fewer per-chunk flushes of partial sums (global
// atomics), smaller ones more workgroups to hide latency.  Measured best of
// 256 / 512 / 1024 / 2048 / 4096: 2048 for 1.3M elements, 512 for 82k, 256
// for 12.6k.
uint32_t chunk_elems(uint32_t n)
{
    const double target = 2.0 * std::sqrt((double)n);
    uint32_t c = kChunkMin;
    while (c < kChunkMax && 2.0 * c <= target)
        c *= 2;
    return c;
}

void enqueue_level(hipStream_t st, BuildWs& w, int faithful, uint32_t max_leaf, uint32_t chunk)
{
    Ctl* ctl = w.ctl;
    uint32_t* err = &w.ctl->err;
    uint32_t cg = std::min<uint32_t>(kChunkGrid, w.ccap);
    hipLaunchKernelGGL(k_level_begin, dim3(1), dim3(kCtlWG), 0, st, w.bn, w.lvl_all, ctl, w.tasks, w.node_chunk,
                       w.chunks, w.lcap, w.ccap, w.wgsum, chunk);
    hipLaunchKernelGGL(k_init_acc, dim3(1024), dim3(256), 0, st, ctl, w.bacc, w.child);
    hipLaunchKernelGGL(k_bins, dim3(cg), dim3(kWG), 0, st, w.ec, w.elo, w.ehi, w.chunks, ctl, w.cacc, w.bacc,
                       faithful);
    hipLaunchKernelGGL(k_decide, dim3(kTaskGrid), dim3(64), 0, st, w.tasks, ctl, w.cacc, w.bacc, w.dec, faithful,
                       max_leaf);
    hipLaunchKernelGGL(k_part_count, dim3(cg), dim3(kWG), 0, st, w.ec, w.chunks, ctl, w.dec, w.chunk_left);
    hipLaunchKernelGGL(k_part_prefix, dim3(kTaskGrid), dim3(64), 0, st, w.node_chunk, ctl, w.chunk_left, w.dec,
                       w.tasks, faithful, max_leaf, w.wgsum);
    hipLaunchKernelGGL((k_part_scatter<false>), dim3(cg), dim3(kWG), 0, st, w.ec, w.elo, w.ehi, w.chunks, ctl,
                       w.tasks, w.dec, w.chunk_left, w.kpos, w.sel, w.oc, w.olo, w.ohi, err);
    hipLaunchKernelGGL((k_part_scatter<true>), dim3(cg), dim3(kWG), 0, st, w.ec, w.elo, w.ehi, w.chunks, ctl,
                       w.tasks, w.dec, w.chunk_left, w.kpos, w.sel, w.oc, w.olo, w.ohi, err);
    hipLaunchKernelGGL(k_part_commit, dim3(cg), dim3(kWG), 0, st, w.oc, w.olo, w.ohi, w.chunks, ctl, w.tasks, w.dec,
                       w.ec, w.elo, w.ehi, w.child);
    hipLaunchKernelGGL(k_level_end, dim3(kLvlWGs), dim3(kCtlWG), 0, st, w.bn, w.lvl_all, w.small_roots,
                       w.level_start, w.tasks, w.dec, w.child, ctl, w.wgsum, faithful, max_leaf, w.hcap, w.rcap,
                       w.cacc, w.lcap);
}

int32_t build_impl(bvht_ctx ctx, float* h_tris, uint32_t n, int faithful, uint32_t max_leaf, bvht_blas* out)
{
    hipStream_t st = ctx->stream;
    BuildWs* wp = nullptr;
    RC(ensure_ws(ctx, n, &wp));
    BuildWs& w = *wp;
    bvht_blas b = nullptr;
    RC(bvht::blas_alloc(ctx, n, 2 * n, &b));
    struct Guard {
        bvht_blas* b;
        ~Guard() { if (*b) bvht_blas_destroy(*b); }
    } guard{&b};
    b->d_perm = static_cast<uint32_t*>(bvht::blas_buffer(ctx, sizeof(uint32_t) * (size_t)n));
    if (!b->d_perm)
        return set_error(BVHT_ENOMEM, "device allocation of the %u-entry build permutation failed", n);

    RC(bvht::host_to_device(ctx, w.tin, h_tris, sizeof(float) * 9 * (size_t)n));
    BVHT_HIP(hipMemsetAsync(b->d_nodes, 0, sizeof(bvht_node) * 2 * (size_t)n, st));
    hipLaunchKernelGGL(k_init_elems, dim3(nblk(n, 256)), dim3(256), 0, st, w.tin, n, w.ec, w.elo, w.ehi);
    // root box: update_node_bounds over everything (bvh.rs:537)
    RC(fill(st, w.rootk, kEmptyMin, 3));
    RC(fill(st, w.rootk + 3, kEmptyMax, 3));
    hipLaunchKernelGGL(k_range_box, dim3(std::min<uint32_t>(nblk(n, kWG), 1024)), dim3(kWG), 0, st, w.elo, w.ehi, 0u,
                       n, w.rootk);
    hipLaunchKernelGGL(k_root_init, dim3(1), dim3(1), 0, st, w.rootk, n, w.bn, w.lvl_all, w.small_roots,
                       w.level_start, w.ctl);
    // the root's centroid range (slot 0 of the first level; k_level_end
    // writes those of the later levels' nodes)
    RC(fill(st, w.cacc, kEmptyMin, 3));
    RC(fill(st, w.cacc + 3, kEmptyMax, 3));
    hipLaunchKernelGGL(k_centroid_root, dim3(std::min<uint32_t>(nblk(n, kWG), 1024)), dim3(kWG), 0, st, w.ec, n,
                       w.cacc);
    BVHT_HIP(hipGetLastError());

    // Large nodes, level by level, read back every few levels: a level with
    // no node costs only its launches (about 25 us for its 11), a read-back
    // a host round trip.  With forced splits (SAH mode, max_leaf > 0) the
    // large nodes span about 1.15-1.25 log2(n / kSmall) levels (19 for the
    // 1.3M-triangle sphere, 14 for 82k triangles, 10 for 12.6k), so the first
    // read-back comes after 1.15 log2(n / kSmall) levels, then after 2, 4,
    // 8, ... more.  Without them the tree stops early (6 levels on those
    // spheres and on bigben.tri, 7 on unity.tri, 10 on the 12.6k soup in the
    // reference mode): the first read-back comes after at most 7, then after
    // 4, 8, ... more.
    Ctl h{};
    const bool forced = !faithful && max_leaf > 0;
    const uint32_t predicted =
        std::max<uint32_t>(1u, (uint32_t)(1.15 * std::log2(std::max(1.0, (double)n / kSmall))));
    const uint32_t first_levels = forced ? predicted : std::min<uint32_t>(7u, predicted);
    const uint32_t chunk = chunk_elems(n);
    for (uint32_t batch = 0;; ++batch) {
        uint32_t levels = batch == 0 ? first_levels : std::min<uint32_t>(32u, (forced ? 1u : 2u) << batch);
        for (uint32_t k = 0; k < levels; ++k)
            enqueue_level(st, w, faithful, max_leaf, chunk);
        BVHT_HIP(hipGetLastError());
        BVHT_HIP(hipMemcpyAsync(&h, w.ctl, sizeof h, hipMemcpyDeviceToHost, st));
        BVHT_HIP(hipStreamSynchronize(st));
        if (h.err || h.L == 0)
            break;
    }
    if (h.err)
        return set_error(BVHT_EDEVICE, "BLAS build self-check failed (flags 0x%x): inconsistent partition or tree",
                         h.err);

    // small subtrees; then the permuted triangles, which depend on nothing
    // else, go back to the caller's array (build_for, bvh.rs:524, partitions
    // the slice it is given in place) on the copy stream while numbering,
    // placement and validation run
    uint32_t R = h.R;
    if (R)
        hipLaunchKernelGGL(k_small_wave, dim3(nblk(R, kSmallWavesPerWG)), dim3(64 * kSmallWavesPerWG), 0, st, w.ec,
                           w.elo, w.ehi, w.bn, w.small_roots, R, w.snodes, w.pairs, faithful, max_leaf,
                           &w.ctl->err);
    hipLaunchKernelGGL(k_gather_tris, dim3(nblk(n, 256)), dim3(256), 0, st, w.tin, w.ec, n, b->d_tris, b->d_perm,
                       &w.ctl->err);
    BVHT_HIP(hipGetLastError());
    RC(bvht::mark_for_copy(ctx));
    hipLaunchKernelGGL(k_number, dim3(1), dim3(kCtlWG), 0, st, w.bn, w.lvl_all, w.level_start, w.ctl, w.pairs, w.ic,
                       w.start, w.fin, w.rfinal, w.rrank, w.rlevel, b->d_nodes, b->d_level, 2 * n);
    if (R)
        hipLaunchKernelGGL(k_small_emit, dim3(nblk(R, 64)), dim3(64), 0, st, w.snodes, w.bn, w.small_roots, w.pairs,
                           w.rfinal, w.rrank, w.rlevel, R, b->d_nodes, b->d_level, 2 * n, &w.ctl->err,
                           &w.ctl->depth);
    hipLaunchKernelGGL(k_validate_tree, dim3(nblk(2 * (size_t)n, 256)), dim3(256), 0, st, b->d_nodes, w.ctl, n,
                       &w.ctl->err);
    BVHT_HIP(hipGetLastError());
    // (the control block's read-back comes after the copy: a copy into
    // pageable memory waits for the work queued before it)
    const size_t tri_bytes = sizeof(float) * 9 * (size_t)n;
    int32_t rc = bvht::device_to_host_beside(ctx, h_tris, b->d_tris, tri_bytes);
    BVHT_HIP(hipMemcpyAsync(&h, w.ctl, sizeof h, hipMemcpyDeviceToHost, st));
    BVHT_HIP(hipStreamSynchronize(st));
    if (rc == BVHT_OK) {
        if (h.err)
            rc = set_error(BVHT_EDEVICE, "BLAS build self-check failed (flags 0x%x): inconsistent partition or tree",
                           h.err);
        else if (h.nodes_used < 2 || h.nodes_used > 2 * n)
            rc = set_error(BVHT_EDEVICE, "BLAS build: %u nodes used of %u", h.nodes_used, 2 * n);
        else if (h.depth == 0 || h.depth > h.nodes_used)
            rc = set_error(BVHT_EDEVICE, "BLAS build: depth %u out of range", h.depth);
        if (rc == BVHT_OK) {
            b->nodes_used = h.nodes_used;
            rc = compute_blas_topology(b, nullptr, h.depth);  // the levels are already in place
        }
        if (rc != BVHT_OK) {
            // the build failed after the copy: the caller's array gets its
            // triangles back in their original order (still in the upload
            // buffer); the first error stays the one reported
            const std::string msg = bvht_last_error();
            (void)bvht::device_to_host(ctx, h_tris, w.tin, tri_bytes);
            set_error(rc, "%s", msg.c_str());
        }
    }
    if (rc != BVHT_OK)
        return rc;
    *out = b;
    b = nullptr;
    return BVHT_OK;
}

}  // namespace

namespace bvht {

void release_build_ws(bvht_ctx ctx)
{
    BuildWs* w = static_cast<BuildWs*>(ctx->build_ws);
    if (!w)
        return;
    if (w->base)
        (void)hipFree(w->base);
    delete w;
    ctx->build_ws = nullptr;
}

}  // namespace bvht

extern "C" int32_t bvht_blas_build(bvht_ctx ctx, float* tris, uint32_t n, int32_t mode, uint32_t max_leaf,
                                   bvht_blas* out)
{
    BVHT_REQUIRE(ctx && tris && out, "NULL argument");
    BVHT_REQUIRE(n > 0, "empty triangle array (the reference panics on an empty mesh)");
    BVHT_REQUIRE(n <= (1u << 30), "too many triangles (%u)", n);
    BVHT_REQUIRE(mode == BVHT_BUILD_REFERENCE || mode == BVHT_BUILD_SAH, "unknown build mode %d", mode);
    return build_impl(ctx, tris, n, mode == BVHT_BUILD_REFERENCE ? 1 : 0, max_leaf, out);
}
