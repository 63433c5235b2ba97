// bvht_trace.hpp -- the per-lane traversal loops (BLAS and two-level TLAS).
#pragma once

#include "bvht_device.hpp"

namespace bvht {

// LDS (address space 3) pointer, so stack accesses compile to ds_read/ds_write
// instead of flat accesses.
typedef __attribute__((address_space(3))) uint32_t lds_u32;

// 64 bytes at a wave-uniform, 4-byte aligned address, loaded once for the
// wave through the scalar data cache into 16 SGPRs.  The wait is inside the
// same asm block, so the registers hold the data when the block ends.
typedef int int16v __attribute__((ext_vector_type(16)));
__device__ __forceinline__ int16v scalar_load64(const void* p)
{
    int16v r;
    asm volatile("s_load_dwordx16 %0, %1, 0x0\n\ts_waitcnt lgkmcnt(0)" : "=s"(r) : "s"(p));
    return r;
}

// Per-lane traversal stack: the first LDS_DEPTH entries live in LDS at
// lds[slot * WG] (lds already offset by the lane: the lanes of a wave hit
// distinct banks), deeper entries in a per-ray column of a global buffer
// (spill[(slot - LDS_DEPTH) * stride]).  `cap` bounds the total; a push beyond
// it is dropped rather than written out of bounds, and sets kStatusStack in
// *status, which the host reports as BVHT_ESTACK at the next fence.  cap comes
// from the tree depth, so only a caller's smaller stack limit (or a wrong
// depth bound) can drop a push.
template <int LDS_DEPTH, int WG>
struct Stack {
    lds_u32* lds;
    uint32_t* spill;
    size_t stride;
    uint32_t cap;
    uint32_t* status;

    __device__ __forceinline__ void push(uint32_t& sp, uint32_t v)
    {
        if (sp < (uint32_t)LDS_DEPTH) {
            lds[sp * WG] = v;
        } else {
            if (sp >= cap) {
                *status = 1u;  // kStatusStack
                return;
            }
            spill[(size_t)(sp - LDS_DEPTH) * stride] = v;
        }
        ++sp;
    }
    __device__ __forceinline__ uint32_t pop(uint32_t& sp)
    {
        --sp;
        return sp < (uint32_t)LDS_DEPTH ? lds[sp * WG] : spill[(size_t)(sp - LDS_DEPTH) * stride];
    }
};

// Trace kernel configurations: workgroup size x LDS stack entries per lane.
// The X-macro instantiates every kernel variant; the host picks one per
// launch from the context's configuration (set by the autotuner or the
// depth-based default).
#define BVHT_TRACE_CONFIGS(X) \
    X(64, 8) X(64, 16) X(64, 24) X(64, 32) \
    X(128, 8) X(128, 16) X(128, 24) X(128, 32) \
    X(256, 8) X(256, 16) X(256, 24) X(256, 32)

// A BLAS stack entry carries the pushed node's (left_first, count) itself
// when they fit -- bit 31 set, count in bits 25..30, left_first in bits
// 0..24 -- so a pop needs no node fetch; otherwise it is the node index
// (bit 31 clear: node indices stay below 2^31) and the pop reads the node.
// The visiting order is unchanged.
constexpr uint32_t kPackFlag = 0x80000000u;
constexpr uint32_t kPackShift = 25;
constexpr uint32_t kPackMask = (1u << kPackShift) - 1u;

__device__ __forceinline__ uint32_t pack_entry(uint32_t idx, uint32_t lf, uint32_t cnt)
{
    return (cnt < 64u && lf <= kPackMask) ? (kPackFlag | (cnt << kPackShift) | lf) : idx;
}

// Lanes of a wave that must still be descending for the node-pair loop to
// take another step (see traverse_blas).
#ifndef BVHT_DESCEND_MIN
#define BVHT_DESCEND_MIN 24
#endif
constexpr uint32_t kDescendMin = BVHT_DESCEND_MIN;

struct BlasResult {
    float t, u, v;
    uint32_t prim;
    bool found;
};

// cnt value of a lane that has finished its node (both children missed) or
// leaf and pops next.  A leaf's count is below the triangle count, so it
// never takes this value.
constexpr uint32_t kPopNext = 0xFFFFFFFFu;

// Bvh::intersect_subtree (bvh.rs:242-305) for one ray, starting at node 0.
// The stack is shared with an enclosing TLAS traversal: this call only uses
// slots at or above sp_base and leaves the stack as it found it.  PACKED:
// every node of the tree fits a packed stack entry (no leaf of 64 or more
// triangles, left_first below 2^25), so entries are packed without a check.
// FINITE: the ray's t is not above FLT_MAX, so every distance below the
// current nearest hit is below FLT_MAX and the child choice reduces to
// comparisons of slab_key values (same decisions as the general form).
template <bool STATS, bool PACKED, bool FINITE = false, class StackT>
__device__ __forceinline__ BlasResult traverse_blas(const float4* __restrict__ nodes,
                                                    const float* __restrict__ tris, const DRay& ray,
                                                    StackT& st, uint32_t sp_base, uint32_t& pairs,
                                                    uint32_t& tests)
{
    BlasResult res{0.0f, 0.0f, 0.0f, 0u, false};
    float closest = ray.t;
    uint32_t sp = sp_base;
    uint32_t lf = __float_as_uint(nodes[0].w);
    uint32_t cnt = __float_as_uint(nodes[1].w);
    while (true) {
        // Node-pair steps while the lane is at an interior node (cnt == 0).
        // After each step the wave leaves this loop once fewer than
        // kDescendMin of its lanes are still descending, instead of holding
        // the lanes that wait at a leaf or a pop until the longest descent
        // ends; the lanes still descending resume here on the next pass.
        // Every lane's own sequence of node steps, leaf tests and pops is
        // that of the one-loop traversal.
        while (cnt == 0) {
            if (STATS)
                ++pairs;
            float4 a0, a1, b0, b1;
            const uint32_t lf0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)lf);
            if (__builtin_amdgcn_ballot_w64(lf != lf0) == 0) {
                // every descending lane is at the same pair: one scalar load
                // of its 64 bytes, which leaves the vector memory path (the
                // texture address and data units) to the other waves
                const int16v p = scalar_load64(nodes + 2 * (size_t)lf0);
                a0 = make_float4(__int_as_float(p[0]), __int_as_float(p[1]), __int_as_float(p[2]),
                                 __int_as_float(p[3]));
                a1 = make_float4(__int_as_float(p[4]), __int_as_float(p[5]), __int_as_float(p[6]),
                                 __int_as_float(p[7]));
                b0 = make_float4(__int_as_float(p[8]), __int_as_float(p[9]), __int_as_float(p[10]),
                                 __int_as_float(p[11]));
                b1 = make_float4(__int_as_float(p[12]), __int_as_float(p[13]), __int_as_float(p[14]),
                                 __int_as_float(p[15]));
            } else {
                const float4* c = nodes + 2 * (size_t)lf;
                a0 = c[0];
                a1 = c[1];
                b0 = c[2];
                b1 = c[3];
            }
            bool left_near, near_hit, far_hit;
            if (FINITE) {
                float kl = slab_key(a0, a1, ray);
                float kr = slab_key(b0, b1, ray);
                // keys are never NaN, so the nearer and farther key are
                // selects on left_near (fmaxf would add two canonicalizes)
                left_near = kl < kr;
                near_hit = (left_near ? kl : kr) < closest;
                far_hit = (left_near ? kr : kl) < closest;
            } else {
                float dl, dr;
                bool hl = slab(a0, a1, ray, closest, dl);
                bool hr = slab(b0, b1, ray, closest, dr);
                float kl = hl ? dl : kFltMax;
                float kr = hr ? dr : kFltMax;
                left_near = kl < kr;
                near_hit = left_near ? hl : hr;
                far_hit = left_near ? hr : hl;
            }
            if (near_hit) {
                if (far_hit) {
                    uint32_t flf = __float_as_uint(left_near ? b0.w : a0.w);
                    uint32_t fcnt = __float_as_uint(left_near ? b1.w : a1.w);
                    st.push(sp, PACKED ? (kPackFlag | (fcnt << kPackShift) | flf)
                                       : pack_entry(left_near ? lf + 1 : lf, flf, fcnt));
                }
                uint32_t nlf = __float_as_uint(left_near ? a0.w : b0.w);
                cnt = __float_as_uint(left_near ? a1.w : b1.w);
                lf = nlf;
            } else {
                cnt = kPopNext;
            }
            if ((uint32_t)__popcll(__builtin_amdgcn_ballot_w64(cnt == 0)) < kDescendMin)
                break;
        }
        if (cnt != 0 && cnt != kPopNext) {
            for (uint32_t k = 0; k < cnt; ++k) {
                float t, u, v;
                if (STATS)
                    ++tests;
                if (tri_test(tris + 9 * (size_t)(lf + k), ray, t, u, v) && t < closest) {
                    closest = t;
                    res.t = t;
                    res.u = u;
                    res.v = v;
                    res.prim = lf + k;
                    res.found = true;
                }
            }
            cnt = kPopNext;
        }
        if (cnt == kPopNext) {
            if (sp == sp_base)
                break;
            uint32_t e = st.pop(sp);
            if (PACKED || (e & kPackFlag)) {
                lf = e & kPackMask;
                cnt = (e >> kPackShift) & 63u;
            } else {
                lf = __float_as_uint(nodes[2 * (size_t)e].w);
                cnt = __float_as_uint(nodes[2 * (size_t)e + 1].w);
            }
        }
        // One loop latch for every lane: without this convergent point the
        // compiler gives the no-pop path a back edge of its own and nests the
        // loops so that pops wait for every lane of the wave to finish its
        // node or leaf.
        __builtin_amdgcn_wave_barrier();
    }
    if (!(closest < kFltMax))
        res.found = false;
    return res;
}

// Per-BLAS device pointers used by the TLAS kernels.
struct BlasRef {
    const float4* nodes;
    const float* tris;
    uint64_t ntris;
};

// Object record as the TLAS kernels read it: the 3x4 row block of M^-1
// ({m[0][i], m[1][i], m[2][i], m[3][i]} for rows i = 0..2) and the BLAS id.
struct ObjRows {
    float4 r0, r1, r2;
    uint32_t blas;
    uint32_t pad[3];
};

// Whether a hit record is bounced off: a hit whose indices lie inside the
// structure (prim < prim_limit and inst < inst_limit; for TLAS hits, with
// rows/blases given, prim below the triangle count of the object's BLAS).
// Hits outside are skipped, never dereferenced.
__device__ __forceinline__ bool bounce_hit(const bvht_hit& h, uint32_t prim_limit, uint32_t inst_limit,
                                           const ObjRows* rows, const BlasRef* blases)
{
    if (h.prim == BVHT_MISS || h.inst >= inst_limit)
        return false;
    if (rows)
        return h.prim < blases[rows[h.inst].blas].ntris;
    return h.prim < prim_limit;
}

struct TlasResult {
    float t, u, v;
    uint32_t prim, inst;
    bool found;
};

// Tlas::intersect (tlas.rs:123-177): the TLAS node loop with
// SceneObject::intersect at the leaves (scene_object.rs:77-89).  Left child =
// high 16 bits of the LeftRightIndex word (tlas.rs:24-32).
//
// Both levels run in one loop: a lane is either in the TLAS (world ray) or in
// the BLAS of the object whose leaf it reached (object-space ray), and the
// node-pair steps of the two levels share one step body, so the lanes of a
// wave descend their instances' BLASes together instead of one instance
// visit at a time.  Each lane's sequence of TLAS steps, object visits (with
// the BLAS traversal of traverse_blas inside) and pops is that of the nested
// traversal; a visit's nearest hit is committed when its BLAS stack empties,
// under the same conditions as the nested form (found, t < FLT_MAX, t below
// the committed distance).
template <bool STATS, bool PACKED = false, bool FINITE = false, class StackT>
__device__ __forceinline__ TlasResult traverse_tlas(const float4* __restrict__ tnodes, const ObjRows* __restrict__ objs,
                                                    const BlasRef* __restrict__ blases, const DRay& ray,
                                                    StackT& st, uint32_t& pairs, uint32_t& tests,
                                                    uint32_t& tpairs, uint32_t& visits)
{
    TlasResult res{0.0f, 0.0f, 0.0f, 0u, 0u, false};
    float closest = ray.t;  // committed nearest distance (world)
    // Current level: in_blas = false -> TLAS with the world ray; true -> the
    // BLAS of object `obj` with the object-space ray cr.
    bool in_blas = false;
    DRay cr = ray;
    float cclosest = closest;  // culling distance of the current level
    const float4* bnodes = tnodes;
    const float* btris = nullptr;
    uint32_t obj = 0, sp = 0, sp_base = 0;
    float bt = 0.0f, bu = 0.0f, bv = 0.0f;
    uint32_t bprim = 0;
    bool bfound = false;
    // Current node: TLAS -> (left_right, object index), leaf when x == 0;
    // BLAS -> (left_first, count), leaf when y > 0.
    uint32_t x = __float_as_uint(tnodes[0].w);
    uint32_t y = __float_as_uint(tnodes[1].w);
    bool pend = false;
    while (true) {
        while (!pend && (in_blas ? y == 0 : x != 0)) {
            if (STATS) {
                if (in_blas)
                    ++pairs;
                else
                    ++tpairs;
            }
            const float4* pl;
            const float4* pr;
            if (in_blas) {
                pl = bnodes + 2 * (size_t)x;
                pr = pl + 2;
            } else {
                pl = tnodes + 2 * (size_t)((x & 0xFFFF0000u) >> 16);
                pr = tnodes + 2 * (size_t)(x & 0x0000FFFFu);
            }
            float4 a0 = pl[0], a1 = pl[1], b0 = pr[0], b1 = pr[1];
            bool left_near, near_hit, far_hit;
            if (FINITE) {
                float kl = slab_key(a0, a1, cr);
                float kr = slab_key(b0, b1, cr);
                left_near = kl < kr;
                near_hit = (left_near ? kl : kr) < cclosest;
                far_hit = (left_near ? kr : kl) < cclosest;
            } else {
                float dl, dr;
                bool hl = slab(a0, a1, cr, cclosest, dl);
                bool hr = slab(b0, b1, cr, cclosest, dr);
                float kl = hl ? dl : kFltMax;
                float kr = hr ? dr : kFltMax;
                left_near = kl < kr;
                near_hit = left_near ? hl : hr;
                far_hit = left_near ? hr : hl;
            }
            if (near_hit) {
                if (far_hit) {
                    uint32_t e;
                    if (in_blas) {
                        uint32_t flf = __float_as_uint(left_near ? b0.w : a0.w);
                        uint32_t fcnt = __float_as_uint(left_near ? b1.w : a1.w);
                        e = PACKED ? (kPackFlag | (fcnt << kPackShift) | flf)
                                   : pack_entry(left_near ? x + 1 : x, flf, fcnt);
                    } else
                        e = left_near ? (x & 0x0000FFFFu) : ((x & 0xFFFF0000u) >> 16);
                    st.push(sp, e);
                }
                uint32_t nx = __float_as_uint(left_near ? a0.w : b0.w);
                y = __float_as_uint(left_near ? a1.w : b1.w);
                x = nx;
            } else {
                pend = true;
            }
            if ((uint32_t)__popcll(__builtin_amdgcn_ballot_w64(!pend && (in_blas ? y == 0 : x != 0))) < kDescendMin)
                break;
        }
        if (!pend && in_blas && y > 0) {
            for (uint32_t k = 0; k < y; ++k) {
                float t, u, v;
                if (STATS)
                    ++tests;
                if (tri_test(btris + 9 * (size_t)(x + k), cr, t, u, v) && t < cclosest) {
                    cclosest = t;
                    bt = t;
                    bu = u;
                    bv = v;
                    bprim = x + k;
                    bfound = true;
                }
            }
            pend = true;
        }
        if (!pend && !in_blas && x == 0) {
            // object leaf: enter its BLAS with the object-space ray
            if (STATS)
                ++visits;
            const ObjRows* ob = objs + y;
            float4 q0 = ob->r0, q1 = ob->r1, q2 = ob->r2;
            uint32_t model = ob->blas;
            V3 om = rows_mul(q0, q1, q2, ray.o, 1.0f);
            V3 dm = rows_mul(q0, q1, q2, ray.d, 0.0f);
            cr = make_ray(om, dm, closest);
            BlasRef br = blases[model];
            bnodes = br.nodes;
            btris = br.tris;
            obj = y;
            sp_base = sp;
            cclosest = closest;
            bfound = false;
            in_blas = true;
            x = __float_as_uint(bnodes[0].w);
            y = __float_as_uint(bnodes[1].w);
        }
        if (pend) {
            if (in_blas && sp == sp_base) {
                // the object's BLAS traversal is over: commit its nearest hit
                if (bfound && cclosest < kFltMax && cclosest < closest) {
                    closest = cclosest;
                    res.t = bt;
                    res.u = bu;
                    res.v = bv;
                    res.prim = bprim;
                    res.inst = obj;
                    res.found = true;
                }
                in_blas = false;
                cr = ray;
                cclosest = closest;
            }
            if (!in_blas && sp == 0)
                break;
            uint32_t e = st.pop(sp);
            if (!in_blas) {
                x = __float_as_uint(tnodes[2 * (size_t)e].w);
                y = __float_as_uint(tnodes[2 * (size_t)e + 1].w);
            } else if (PACKED || (e & kPackFlag)) {
                x = e & kPackMask;
                y = (e >> kPackShift) & 63u;
            } else {
                x = __float_as_uint(bnodes[2 * (size_t)e].w);
                y = __float_as_uint(bnodes[2 * (size_t)e + 1].w);
            }
            pend = false;
        }
        // one loop latch for every lane (see traverse_blas)
        __builtin_amdgcn_wave_barrier();
    }
    if (!(closest < kFltMax))
        res.found = false;
    return res;
}

}  // namespace bvht
