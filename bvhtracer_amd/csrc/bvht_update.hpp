// bvht_update.hpp -- per-frame structure updates shared by the single-structure
// kernels and the batched multi-scene kernels: BLAS refit, object bounds and
// the TLAS rebuild.
#pragma once

#include "bvht_device.hpp"
#include "bvht_trace.hpp"

namespace bvht {

// Bvh::refit (bvh.rs:469-493) for node k, run one tree level at a time from
// the deepest level up (the levels are separate launches, so a node's
// children are final when it is processed): a leaf recomputes its box from
// its triangles starting at (+MAX, -MAX) (update_node_bounds, bvh.rs:317-330),
// a branch takes the union of its two children.  min/max are exact and every
// child has a larger index than its parent (bvh.rs:601-612), so the result
// equals the reference's reverse index sweep bit for bit.
// Leaves with more than kBigLeaf triangles (the reference builder makes
// leaves of thousands) are refit by a whole wave each (refit_big_leaf), before
// the level passes; refit_node leaves them alone.
constexpr uint32_t kBigLeaf = 64;

__device__ __forceinline__ void refit_node(bvht_node* __restrict__ nodes, const float* __restrict__ tris, uint32_t k)
{
    bvht_node nd = nodes[k];
    if (nd.count > kBigLeaf)
        return;
    V3 lo, hi;
    if (nd.count > 0) {
        lo = mk3(kFltMax, kFltMax, kFltMax);
        hi = mk3(-kFltMax, -kFltMax, -kFltMax);
        for (uint32_t j = 0; j < nd.count; ++j) {
            const float* t = tris + 9 * (size_t)(nd.left_first + j);
            V3 v0 = mk3(t[0], t[1], t[2]), v1 = mk3(t[3], t[4], t[5]), v2 = mk3(t[6], t[7], t[8]);
            lo = min3(lo, v0);
            lo = min3(lo, v1);
            lo = min3(lo, v2);
            hi = max3(hi, v0);
            hi = max3(hi, v1);
            hi = max3(hi, v2);
        }
    } else {
        bvht_node a = nodes[nd.left_first], c = nodes[nd.left_first + 1];
        lo = min3(mk3(a.bmin[0], a.bmin[1], a.bmin[2]), mk3(c.bmin[0], c.bmin[1], c.bmin[2]));
        hi = max3(mk3(a.bmax[0], a.bmax[1], a.bmax[2]), mk3(c.bmax[0], c.bmax[1], c.bmax[2]));
    }
    float4* o = reinterpret_cast<float4*>(nodes + k);
    o[0] = make_float4(lo.x, lo.y, lo.z, __uint_as_float(nd.left_first));
    o[1] = make_float4(hi.x, hi.y, hi.z, __uint_as_float(nd.count));
}

// Leaf k's box from its triangles by the 64 lanes of one wave: lanes take
// every 64th triangle, then min/max are combined across the wave (exact, so
// the order does not change the box; only the sign of a zero bound can
// differ from the sequential sweep).  Every lane of the wave must call it.
__device__ __forceinline__ void refit_big_leaf(bvht_node* __restrict__ nodes, const float* __restrict__ tris,
                                               uint32_t k, uint32_t lane)
{
    bvht_node nd = nodes[k];
    V3 lo = mk3(kFltMax, kFltMax, kFltMax);
    V3 hi = mk3(-kFltMax, -kFltMax, -kFltMax);
    for (uint32_t j = lane; j < nd.count; j += 64) {
        const float* t = tris + 9 * (size_t)(nd.left_first + j);
        V3 v0 = mk3(t[0], t[1], t[2]), v1 = mk3(t[3], t[4], t[5]), v2 = mk3(t[6], t[7], t[8]);
        lo = min3(min3(min3(lo, v0), v1), v2);
        hi = max3(max3(max3(hi, v0), v1), v2);
    }
    for (int m = 32; m >= 1; m >>= 1) {
        lo = min3(lo, mk3(__shfl_xor(lo.x, m, 64), __shfl_xor(lo.y, m, 64), __shfl_xor(lo.z, m, 64)));
        hi = max3(hi, mk3(__shfl_xor(hi.x, m, 64), __shfl_xor(hi.y, m, 64), __shfl_xor(hi.z, m, 64)));
    }
    if (lane == 0) {
        float4* o = reinterpret_cast<float4*>(nodes + k);
        o[0] = make_float4(lo.x, lo.y, lo.z, __uint_as_float(nd.left_first));
        o[1] = make_float4(hi.x, hi.y, hi.z, __uint_as_float(nd.count));
    }
}

// SceneObject::set_transform (scene_object.rs:60-75) for object i: world box of
// the 8 corners of its BLAS root box under M, and the traced rows of M^-1.
__device__ __forceinline__ void object_bounds(const bvht_instance* __restrict__ objs,
                                              const BlasRef* __restrict__ blases, float* __restrict__ bounds,
                                              ObjRows* __restrict__ rows, uint32_t i)
{
    const bvht_instance& o = objs[i];
    const float4* bn = blases[o.blas].nodes;
    float4 a = bn[0], b = bn[1];
    V3 lo = mk3(kFltMax, kFltMax, kFltMax), hi = mk3(-kFltMax, -kFltMax, -kFltMax);
    for (int c = 0; c < 8; ++c) {
        V3 p = mk3((c & 1) ? b.x : a.x, (c & 2) ? b.y : a.y, (c & 4) ? b.z : a.z);
        V3 w = mat_mul4(o.m, p, 1.0f);
        lo = min3(lo, w);
        hi = max3(hi, w);
    }
    float* bo = bounds + 6 * (size_t)i;
    bo[0] = lo.x; bo[1] = lo.y; bo[2] = lo.z;
    bo[3] = hi.x; bo[4] = hi.y; bo[5] = hi.z;
    ObjRows r;
    r.r0 = make_float4(o.minv[0], o.minv[4], o.minv[8], o.minv[12]);
    r.r1 = make_float4(o.minv[1], o.minv[5], o.minv[9], o.minv[13]);
    r.r2 = make_float4(o.minv[2], o.minv[6], o.minv[10], o.minv[14]);
    r.blas = o.blas;
    r.pad[0] = r.pad[1] = r.pad[2] = 0;
    rows[i] = r;
}

constexpr uint32_t kRebuildWG = 256;

// Tlas::find_best_match (tlas.rs:179-202) as a block-wide argmin: the
// sequential scan keeps the first b whose area is strictly below the running
// minimum (starting at f32::MAX), i.e. the smallest index among the minimal
// areas that are < f32::MAX.
__device__ __forceinline__ int32_t block_best_match(const bvht_tnode* nodes, const int32_t* list, int32_t count, int32_t a,
                                    float* s_val, int32_t* s_idx)
{
    float best = kFltMax;
    int32_t bi = -1;
    const bvht_tnode& na = nodes[list[a]];
    for (int32_t b = threadIdx.x; b < count; b += kRebuildWG) {
        if (b == a)
            continue;
        const bvht_tnode& nb = nodes[list[b]];
        V3 mx = max3(mk3(na.bmax[0], na.bmax[1], na.bmax[2]), mk3(nb.bmax[0], nb.bmax[1], nb.bmax[2]));
        V3 mn = min3(mk3(na.bmin[0], na.bmin[1], na.bmin[2]), mk3(nb.bmin[0], nb.bmin[1], nb.bmin[2]));
        V3 e = sub3(mx, mn);
        float sa = e.x * e.y + e.y * e.z + e.z * e.x;
        if (sa < best) {
            best = sa;
            bi = b;
        }
    }
    s_val[threadIdx.x] = best;
    s_idx[threadIdx.x] = bi;
    __syncthreads();
    for (uint32_t off = kRebuildWG / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            float v2 = s_val[threadIdx.x + off];
            int32_t i2 = s_idx[threadIdx.x + off];
            float v1 = s_val[threadIdx.x];
            int32_t i1 = s_idx[threadIdx.x];
            bool take = i2 >= 0 && (i1 < 0 || v2 < v1 || (!(v1 < v2) && i2 < i1));
            if (take) {
                s_val[threadIdx.x] = v2;
                s_idx[threadIdx.x] = i2;
            }
        }
        __syncthreads();
    }
    int32_t r = s_idx[0];
    __syncthreads();
    return r;
}

// The merge loop of Tlas::rebuild (tlas.rs:204-250) for 1 <= n <= 64 objects
// by one wave, without barriers.  Lane i holds list entry i in registers: the
// node index, its box, link field, object index and subtree height (entries
// past count keep their last values, as the reference's list does).
// find_best_match is a butterfly minimum over the areas < f32::MAX (as
// order-preserving unsigned keys) followed by a ballot for the first lane
// holding it: the sequential scan's first minimum.  Lane 0 writes the merged
// nodes, the root copy in node 0 and the results; the height of the root is
// the depth that walking the tree from node 0 gives.  The loop is capped (it
// ends after at most n merges and chain steps, far below the cap).
__device__ __forceinline__ float lane_f(float v, int32_t l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ int32_t lane_i(int32_t v, int32_t l) { return __builtin_amdgcn_readlane(v, l); }

__device__ void wave_rebuild_loop(const float* __restrict__ bounds, bvht_tnode* nodes, int32_t* list, int32_t n,
                                  int32_t* out_a, int32_t* out_used, int32_t* out_fail, uint32_t* out_depth)
{
    const int32_t lane = (int32_t)(threadIdx.x & 63u);
    int32_t li = 0, lr = 0, obj = 0, h = 1;
    V3 lo = mk3(0.0f, 0.0f, 0.0f), hi = lo;
    if (lane < n) {  // leaves 1..n in object order
        const float* bb = bounds + 6 * (size_t)lane;
        li = lane + 1;
        obj = lane;
        lo = mk3(bb[0], bb[1], bb[2]);
        hi = mk3(bb[3], bb[4], bb[5]);
    }
    auto best_match = [&](int32_t count, int32_t x) -> int32_t {
        V3 xl = mk3(lane_f(lo.x, x), lane_f(lo.y, x), lane_f(lo.z, x));
        V3 xh = mk3(lane_f(hi.x, x), lane_f(hi.y, x), lane_f(hi.z, x));
        // candidate areas as order-preserving unsigned keys (an area is
        // never -0, so equal areas give equal keys); no candidate: ~0
        uint32_t key = 0xFFFFFFFFu;
        if (lane < count && lane != x) {
            V3 e = sub3(max3(xh, hi), min3(xl, lo));
            float sa = e.x * e.y + e.y * e.z + e.z * e.x;
            if (sa < kFltMax) {
                uint32_t u = __float_as_uint(sa);
                key = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
            }
        }
        uint32_t mk = key, ul = (uint32_t)lane;  // butterfly minimum, VALU exchanges
        mk = min(mk, xor_lane<32>(mk, ul));
        mk = min(mk, xor_lane<16>(mk, ul));
        mk = min(mk, xor_lane<8>(mk, ul));
        mk = min(mk, xor_lane<4>(mk, ul));
        mk = min(mk, xor_lane<2>(mk, ul));
        mk = min(mk, xor_lane<1>(mk, ul));
        mk = __builtin_amdgcn_readfirstlane(mk);
        if (mk == 0xFFFFFFFFu)
            return -1;
        // the first lane holding the minimum: the sequential scan's choice
        return (int32_t)__ffsll((long long)__ballot(key == mk)) - 1;
    };
    int32_t count = n, used = n + 1, a = 0, fail = 0;
    int32_t b = best_match(count, a);
    for (int32_t guard = 0; count > 1; ++guard) {
        if (b < 0 || guard > 8 * 64 * 64) {  // every area >= f32::MAX (the reference panics)
            fail = 1;
            break;
        }
        int32_t c = best_match(count, b);
        if (a == c) {
            int32_t ia = lane_i(li, a), ib = lane_i(li, b);
            V3 mn = min3(mk3(lane_f(lo.x, a), lane_f(lo.y, a), lane_f(lo.z, a)),
                         mk3(lane_f(lo.x, b), lane_f(lo.y, b), lane_f(lo.z, b)));
            V3 mx = max3(mk3(lane_f(hi.x, a), lane_f(hi.y, a), lane_f(hi.z, a)),
                         mk3(lane_f(hi.x, b), lane_f(hi.y, b), lane_f(hi.z, b)));
            int32_t hn = 1 + max(lane_i(h, a), lane_i(h, b));
            int32_t lrn = (int32_t)((uint32_t)ia + ((uint32_t)ib << 16));
            if (lane == 0) {
                bvht_tnode nn;
                nn.left_right = (uint32_t)lrn;
                nn.bmin[0] = mn.x; nn.bmin[1] = mn.y; nn.bmin[2] = mn.z;
                nn.bmax[0] = mx.x; nn.bmax[1] = mx.y; nn.bmax[2] = mx.z;
                nn.blas = 0;
                nodes[used] = nn;
            }
            // list[a] = used, then list[b] = list[count - 1] (which is the new
            // entry when a == count - 1)
            if (lane == a) {
                li = used;
                lo = mn;
                hi = mx;
                lr = lrn;
                obj = 0;
                h = hn;
            }
            int32_t t = count - 1;
            int32_t tli = lane_i(li, t), tlr = lane_i(lr, t), tobj = lane_i(obj, t), th_ = lane_i(h, t);
            V3 tlo = mk3(lane_f(lo.x, t), lane_f(lo.y, t), lane_f(lo.z, t));
            V3 thi = mk3(lane_f(hi.x, t), lane_f(hi.y, t), lane_f(hi.z, t));
            if (lane == b) {
                li = tli;
                lr = tlr;
                obj = tobj;
                h = th_;
                lo = tlo;
                hi = thi;
            }
            used += 1;
            count -= 1;
            b = best_match(count, a);
        } else {
            a = b;
            b = c;
        }
    }
    if (lane < n)
        list[lane] = li;
    // the root (entry a) copied to node 0, and its height
    V3 rl = mk3(lane_f(lo.x, a), lane_f(lo.y, a), lane_f(lo.z, a));
    V3 rh = mk3(lane_f(hi.x, a), lane_f(hi.y, a), lane_f(hi.z, a));
    int32_t rlr = lane_i(lr, a), robj = lane_i(obj, a), rh_ = lane_i(h, a);
    if (lane == 0) {
        if (!fail) {
            bvht_tnode r;
            r.bmin[0] = rl.x; r.bmin[1] = rl.y; r.bmin[2] = rl.z;
            r.bmax[0] = rh.x; r.bmax[1] = rh.y; r.bmax[2] = rh.z;
            r.left_right = (uint32_t)rlr;
            r.blas = (uint32_t)robj;
            nodes[0] = r;
        }
        *out_a = a;
        *out_used = used;
        *out_fail = fail;
        *out_depth = (uint32_t)rh_;
    }
}

__device__ __forceinline__ void meta_out(uint32_t* meta, uint32_t used, uint32_t depth, uint32_t status)
{
    meta[0] = used;
    meta[1] = depth;
    meta[2] = status;
}

// Tlas::rebuild (tlas.rs:204-250) + TlasBuilder::build_for (tlas.rs:273-279):
// 2n default nodes (empty box), leaves 1..n in object order, mutual-best-match
// merges with LeftRightIndex::new(a, b) = a + (b << 16), root = copy of the
// last node.  meta = {nodes_used, depth, status}.
__device__ void tlas_rebuild_block(const float* __restrict__ bounds, uint32_t n, bvht_tnode* nodes, int32_t* list,
                                   uint32_t* meta)
{
    __shared__ float s_val[kRebuildWG];
    __shared__ int32_t s_idx[kRebuildWG];
    __shared__ int32_t sh_a, sh_b, sh_count, sh_used, sh_merged, sh_fail;
    __shared__ uint32_t sh_depth;
    for (uint32_t i = threadIdx.x; i < 2 * n; i += kRebuildWG) {
        bvht_tnode d;
        d.bmin[0] = d.bmin[1] = d.bmin[2] = kFltMax;
        d.bmax[0] = d.bmax[1] = d.bmax[2] = -kFltMax;
        d.left_right = 0;
        d.blas = 0;
        nodes[i] = d;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += kRebuildWG) {
        list[i] = (int32_t)(i + 1);
        bvht_tnode l;
        const float* b = bounds + 6 * (size_t)i;
        l.bmin[0] = b[0]; l.bmin[1] = b[1]; l.bmin[2] = b[2];
        l.bmax[0] = b[3]; l.bmax[1] = b[4]; l.bmax[2] = b[5];
        l.left_right = 0;
        l.blas = i;
        nodes[i + 1] = l;
    }
    if (threadIdx.x == 0) {
        sh_a = 0;
        sh_count = (int32_t)n;
        sh_used = (int32_t)n + 1;
        sh_fail = 0;
    }
    __syncthreads();
    if (n >= 1 && n <= 64) {
        if (threadIdx.x < 64)
            wave_rebuild_loop(bounds, nodes, list, (int32_t)n, &sh_a, &sh_used, &sh_fail, &sh_depth);
        __syncthreads();
        if (threadIdx.x == 0)
            meta_out(meta, (uint32_t)sh_used, sh_fail ? 0u : sh_depth, sh_fail ? 1u : 0u);
        return;
    } else {
    int32_t b0 = block_best_match(nodes, list, sh_count, sh_a, s_val, s_idx);
    if (threadIdx.x == 0)
        sh_b = b0;
    __syncthreads();
    while (sh_count > 1) {
        if (sh_b < 0) {  // every candidate area was >= f32::MAX (the reference panics)
            if (threadIdx.x == 0)
                sh_fail = 1;
            __syncthreads();
            break;
        }
        int32_t c = block_best_match(nodes, list, sh_count, sh_b, s_val, s_idx);
        if (threadIdx.x == 0) {
            int32_t a = sh_a, b = sh_b;
            if (a == c) {
                int32_t ia = list[a], ib = list[b];
                bvht_tnode na = nodes[ia], nb = nodes[ib];
                bvht_tnode nn;
                nn.left_right = (uint32_t)ia + ((uint32_t)ib << 16);
                V3 mn = min3(mk3(na.bmin[0], na.bmin[1], na.bmin[2]), mk3(nb.bmin[0], nb.bmin[1], nb.bmin[2]));
                V3 mx = max3(mk3(na.bmax[0], na.bmax[1], na.bmax[2]), mk3(nb.bmax[0], nb.bmax[1], nb.bmax[2]));
                nn.bmin[0] = mn.x; nn.bmin[1] = mn.y; nn.bmin[2] = mn.z;
                nn.bmax[0] = mx.x; nn.bmax[1] = mx.y; nn.bmax[2] = mx.z;
                nn.blas = 0;
                nodes[sh_used] = nn;
                list[a] = sh_used;
                sh_used += 1;
                list[b] = list[sh_count - 1];
                sh_count -= 1;
                sh_merged = 1;
            } else {
                sh_a = b;
                sh_b = c;
                sh_merged = 0;
            }
        }
        __syncthreads();
        if (sh_merged) {
            int32_t nb = block_best_match(nodes, list, sh_count, sh_a, s_val, s_idx);
            if (threadIdx.x == 0)
                sh_b = nb;
            __syncthreads();
        }
    }
    }  // more than 64 objects
    if (threadIdx.x == 0) {
        if (!sh_fail) {
            nodes[0] = nodes[list[sh_a]];
            // depth: iterative walk, list[] reused as (node, level) stack
            uint32_t best = 0, sp = 0;
            uint32_t* st = reinterpret_cast<uint32_t*>(list);
            st[sp++] = 0;
            st[sp++] = 1;
            while (sp) {
                uint32_t lvl = st[--sp], ni = st[--sp];
                best = max(best, lvl);
                uint32_t lr = nodes[ni].left_right;
                if (lr != 0 && sp + 4 <= 2 * max(n, 4u) + 4) {  // the list holds 2 max(n, 4) + 4 words
                    st[sp++] = lr >> 16;
                    st[sp++] = lvl + 1;
                    st[sp++] = lr & 0xFFFFu;
                    st[sp++] = lvl + 1;
                }
            }
            meta[0] = (uint32_t)sh_used;
            meta[1] = best;
            meta[2] = 0;
        } else {
            meta[0] = (uint32_t)sh_used;
            meta[1] = 0;
            meta[2] = 1;
        }
    }
}


}  // namespace bvht
