// bvht_tlas.hip -- scene objects, the TLAS rebuild and the two-level trace.
//
// Objects (SceneObject, scene_object.rs:11-18) carry M, M^-1 and a BLAS id.
// Their world boxes come from the current BLAS root box through the 8
// transformed corners (set_transform, scene_object.rs:60-75).  The TLAS is
// rebuilt by agglomerative clustering (Tlas::rebuild, tlas.rs:179-250),
// replayed by one workgroup: the clustering loop is inherently sequential,
// each find_best_match scan is a block-wide argmin.  N is the object count
// (16 per scene in the instancing workload), so the whole rebuild is a few
// microseconds.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "bvht_device.hpp"
#include "bvht_internal.hpp"
#include "bvht_trace.hpp"
#include "bvht_update.hpp"

using namespace bvht;

namespace {

inline uint32_t nblk(size_t n, uint32_t wg) { return (uint32_t)((n + wg - 1) / wg); }

__global__ void k_object_bounds(const bvht_instance* __restrict__ objs, uint32_t n, const BlasRef* __restrict__ blases,
                                float* __restrict__ bounds, ObjRows* __restrict__ rows)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        object_bounds(objs, blases, bounds, rows, i);
}

__global__ __launch_bounds__(kRebuildWG) void k_tlas_rebuild(const float* __restrict__ bounds, uint32_t n,
                                                             bvht_tnode* nodes, int32_t* list, uint32_t* meta)
{
    tlas_rebuild_block(bounds, n, nodes, list, meta);
}

template <int WG, int LDS_DEPTH, bool STATS, bool PACKED>
__global__ __launch_bounds__(WG) void k_trace_tlas(const float4* __restrict__ tnodes, const ObjRows* __restrict__ rows,
                                                   const BlasRef* __restrict__ blases,
                                                   const bvht_ray* __restrict__ rays, bvht_hit* __restrict__ hits,
                                                   uint32_t n, const uint32_t* __restrict__ d_n,
                                                   bvht_stats* __restrict__ stats, uint32_t* __restrict__ spill,
                                                   uint32_t cap, uint32_t* __restrict__ status,
                                                   const uint32_t* __restrict__ sched,
                                                   const uint32_t* __restrict__ sched_n, uint32_t* __restrict__ cost)
{
    __shared__ uint32_t lds[LDS_DEPTH * WG];
    // 64-ray group of this wave: slot order, or the order in `sched` (a
    // permutation of the groups); with `cost`, lane 0 records the group's
    // duration for the next call's order
    const uint64_t t0 = cost ? (uint64_t)wall_clock64() : 0;
    const uint32_t n_groups = (n + 63) / 64;
    const uint32_t slot = blockIdx.x * (WG / 64) + (threadIdx.x >> 6);
    const uint32_t group = (sched && slot < n_groups && *sched_n == n_groups) ? sched[slot] : slot;
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t i = group * 64 + lane;
    uint32_t count = d_n ? min(*d_n, n) : n;
    if (i >= count)
        return;
    Stack<LDS_DEPTH, WG> st{(lds_u32*)(lds + threadIdx.x), spill + i, (size_t)n, cap, status};
    DRay ray = load_ray(rays, i);
    uint32_t pairs = 0, tests = 0, tpairs = 0, visits = 0;
    // a wave whose rays all have t <= FLT_MAX (or NaN) takes the reduced child choice
    TlasResult r = __builtin_amdgcn_ballot_w64(ray.t > kFltMax) == 0
                       ? traverse_tlas<STATS, PACKED, true>(tnodes, rows, blases, ray, st, pairs, tests, tpairs, visits)
                       : traverse_tlas<STATS, PACKED, false>(tnodes, rows, blases, ray, st, pairs, tests, tpairs,
                                                             visits);
    store_hit(hits, i, r.found, r.t, r.u, r.v, r.prim, r.inst);
    if (STATS) {
        bvht_stats s;
        s.pairs = pairs;
        s.tests = tests;
        s.tlas_pairs = tpairs;
        s.inst_visits = visits;
        stats[i] = s;
    }
    if (cost && lane == 0) {
        uint64_t dt = (uint64_t)wall_clock64() - t0;
        cost[group] = dt > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)dt;
    }
}

typedef void (*trace_tlas_fn)(const float4*, const ObjRows*, const BlasRef*, const bvht_ray*, bvht_hit*, uint32_t,
                              const uint32_t*, bvht_stats*, uint32_t*, uint32_t, uint32_t*, const uint32_t*,
                              const uint32_t*, uint32_t*);

trace_tlas_fn pick_trace_tlas(int wg, int lds, bool stats, bool packed)
{
#define X(W, D)                                                                             \
    if (wg == W && lds == D) {                                                              \
        if (packed)                                                                         \
            return stats ? k_trace_tlas<W, D, true, true> : k_trace_tlas<W, D, false, true>;  \
        return stats ? k_trace_tlas<W, D, true, false> : k_trace_tlas<W, D, false, false>;    \
    }
    BVHT_TRACE_CONFIGS(X)
#undef X
    return nullptr;
}

// Bounce rays for TLAS hits: the geometric normal of the model-space
// triangle is carried to world space by the transpose of M^-1.
__global__ __launch_bounds__(256) void k_bounce_write_tlas(const ObjRows* __restrict__ rows, uint32_t n_obj,
                                                           const BlasRef* __restrict__ blases,
                                                           const bvht_ray* __restrict__ rays,
                                                           const bvht_hit* __restrict__ hits, uint32_t n,
                                                           const uint32_t* __restrict__ block_offsets, uint64_t seed,
                                                           bvht_ray* __restrict__ out)
{
    __shared__ uint32_t wave_counts[4];
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    bvht_hit h;
    bool hit = false;
    if (i < n) {
        h = hits[i];
        hit = bounce_hit(h, 0xFFFFFFFFu, n_obj, rows, blases);
    }
    uint64_t m = __ballot(hit);
    uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        wave_counts[wave] = __popcll(m);
    __syncthreads();
    uint32_t base = block_offsets[blockIdx.x];
    for (uint32_t w = 0; w < wave; ++w)
        base += wave_counts[w];
    uint32_t rank = __popcll(m & ((1ull << lane) - 1ull));
    if (!hit)
        return;
    ObjRows ob = rows[h.inst];
    const float* tri = blases[ob.blas].tris + 9 * (size_t)h.prim;
    V3 v0 = mk3(tri[0], tri[1], tri[2]);
    V3 e1 = sub3(mk3(tri[3], tri[4], tri[5]), v0);
    V3 e2 = sub3(mk3(tri[6], tri[7], tri[8]), v0);
    V3 nm = cross3(e1, e2);
    V3 nw = mk3(ob.r0.x * nm.x + ob.r1.x * nm.y + ob.r2.x * nm.z, ob.r0.y * nm.x + ob.r1.y * nm.y + ob.r2.y * nm.z,
                ob.r0.z * nm.x + ob.r1.z * nm.y + ob.r2.z * nm.z);
    DRay r = load_ray(rays, i);
    float4 a, b;
    bounce_ray(nw, add3(r.o, scl3(r.d, h.t)), r.d, seed, i, a, b);
    float4* o = reinterpret_cast<float4*>(out + base + rank);
    o[0] = a;
    o[1] = b;
}

void tlas_free(bvht_tlas t)
{
    (void)hipFree(t->d_blas_table);
    (void)hipFree(t->d_attr_table);
    (void)hipFree(t->d_objects);
    (void)hipFree(t->d_rows);
    (void)hipFree(t->d_bounds);
    (void)hipFree(t->d_nodes);
    (void)hipFree(t->d_list);
    (void)hipFree(t->d_meta);
    (void)hipFree(t->d_object_body);
    (void)hipFree(t->d_transform_init);
    delete t;
}

int32_t validate_objects(bvht_tlas t, const bvht_instance* objs, uint32_t count)
{
    for (uint32_t i = 0; i < count; ++i)
        if (objs[i].blas >= t->blases.size())
            return set_error(BVHT_EINVAL, "object %u references BLAS %u of %zu", i, objs[i].blas, t->blases.size());
    return BVHT_OK;
}

int32_t refresh(bvht_tlas t)
{
    hipLaunchKernelGGL(k_object_bounds, dim3(nblk(t->n, 64)), dim3(64), 0, t->ctx->stream, t->d_objects, t->n,
                       static_cast<const BlasRef*>(t->d_blas_table), t->d_bounds, static_cast<ObjRows*>(t->d_rows));
    BVHT_HIP(hipGetLastError());
    return BVHT_OK;
}

int32_t rebuild(bvht_tlas t, bool read_depth)
{
    hipLaunchKernelGGL(k_tlas_rebuild, dim3(1), dim3(kRebuildWG), 0, t->ctx->stream, t->d_bounds, t->n, t->d_nodes,
                       t->d_list, t->d_meta);
    BVHT_HIP(hipGetLastError());
    if (read_depth) {
        uint32_t meta[3];
        BVHT_HIP(hipMemcpyAsync(meta, t->d_meta, sizeof meta, hipMemcpyDeviceToHost, t->ctx->stream));
        BVHT_HIP(hipStreamSynchronize(t->ctx->stream));
        if (meta[2])
            return set_error(BVHT_EINVAL, "TLAS rebuild found no finite merge candidate (non-finite object bounds)");
        t->depth = meta[1];
    }
    return BVHT_OK;
}

int32_t launch_trace_tlas(bvht_tlas t, const bvht_ray* d_rays, uint32_t n, const uint32_t* d_n, bvht_hit* d_hits,
                          bvht_stats* d_stats)
{
    if (n == 0)
        return BVHT_OK;
    bvht_ctx ctx = t->ctx;
    uint32_t need =
        stack_cap(ctx, (t->depth > 0 ? t->depth - 1 : 0) + (t->max_blas_depth > 0 ? t->max_blas_depth - 1 : 0));
    int wg, lds;
    trace_config(ctx, need, &wg, &lds);
    uint32_t* spill = nullptr;
    if (need > (uint32_t)lds) {
        int32_t rc = ensure_spill(ctx, &ctx->spill, &ctx->spill_words, (size_t)(need - lds) * n);
        if (rc)
            return rc;
        spill = ctx->spill;
    }
    // packed stack entries when every BLAS the objects can enter fits them
    bool packed = true;
    for (bvht_blas b : t->blases)
        packed = packed && b->packed_stack;
    trace_tlas_fn fn = pick_trace_tlas(wg, lds, d_stats != nullptr, packed);
    if (!fn)
        return set_error(BVHT_EINVAL, "no TLAS trace kernel for workgroup %d, LDS depth %d", wg, lds);
    // longest-first order from the last call over the same TLAS, rays and
    // count (bvht_set_wave_schedule)
    TraceOrder o;
    int32_t rc = trace_order_begin(ctx, t, d_rays, n, d_n == nullptr && d_stats == nullptr, &o);
    if (rc)
        return rc;
    hipLaunchKernelGGL(fn, dim3(nblk(n, wg)), dim3(wg), 0, ctx->stream, reinterpret_cast<const float4*>(t->d_nodes),
                       static_cast<const ObjRows*>(t->d_rows), static_cast<const BlasRef*>(t->d_blas_table), d_rays,
                       d_hits, n, d_n, d_stats, spill, need, ctx->d_status, o.sched, o.sched_n, o.cost);
    BVHT_HIP(hipGetLastError());
    return trace_order_end(ctx, t, d_rays, n, o);
}

}  // namespace

int32_t bvht::tlas_refresh(bvht_tlas t)
{
    return refresh(t);
}

extern "C" {

int32_t bvht_tlas_create(bvht_ctx ctx, const bvht_blas* blases, uint32_t blas_count, const bvht_instance* objects,
                         uint32_t n, bvht_tlas* out)
{
    BVHT_REQUIRE(ctx && out && blases && objects, "NULL argument");
    BVHT_REQUIRE(n > 0, "empty object list (the reference panics on an empty TLAS)");
    BVHT_REQUIRE(n <= 32767, "%u objects exceed the 16-bit LeftRightIndex (at most 32767)", n);
    BVHT_REQUIRE(blas_count > 0, "no BLAS given");
    bvht_tlas t = new (std::nothrow) bvht_tlas_s();
    if (!t)
        return set_error(BVHT_ENOMEM, "host allocation failed");
    t->ctx = ctx;
    t->n = n;
    std::vector<BlasRef> table(blas_count);
    for (uint32_t i = 0; i < blas_count; ++i) {
        if (!blases[i]) {
            delete t;
            return set_error(BVHT_EINVAL, "BLAS %u is NULL", i);
        }
        t->blases.push_back(blases[i]);
        table[i].nodes = reinterpret_cast<const float4*>(blases[i]->d_nodes);
        table[i].tris = blases[i]->d_tris;
        table[i].ntris = blases[i]->n_tris;
        t->max_blas_depth = std::max(t->max_blas_depth, blases[i]->depth);
    }
    int32_t rc = validate_objects(t, objects, n);
    if (rc) {
        delete t;
        return rc;
    }
    bool ok = hipMalloc(&t->d_blas_table, sizeof(BlasRef) * blas_count) == hipSuccess &&
              hipMalloc(&t->d_objects, sizeof(bvht_instance) * n) == hipSuccess &&
              hipMalloc(&t->d_rows, sizeof(ObjRows) * n) == hipSuccess &&
              hipMalloc(&t->d_bounds, sizeof(float) * 6 * n) == hipSuccess &&
              hipMalloc(&t->d_nodes, sizeof(bvht_tnode) * 2 * n) == hipSuccess &&
              hipMalloc(&t->d_list, sizeof(int32_t) * (std::max<uint32_t>(n, 4) * 2 + 4)) == hipSuccess &&
              hipMalloc(&t->d_meta, sizeof(uint32_t) * 4) == hipSuccess;
    if (!ok) {
        tlas_free(t);
        return set_error(BVHT_ENOMEM, "device allocation for a TLAS of %u objects failed", n);
    }
    hipStream_t st = ctx->stream;
    if (hipMemcpyAsync(t->d_blas_table, table.data(), sizeof(BlasRef) * blas_count, hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        hipMemcpyAsync(t->d_objects, objects, sizeof(bvht_instance) * n, hipMemcpyHostToDevice, st) != hipSuccess) {
        tlas_free(t);
        return set_error(BVHT_EDEVICE, "TLAS upload failed");
    }
    rc = refresh(t);
    if (!rc)
        rc = rebuild(t, true);
    if (rc) {
        tlas_free(t);
        return rc;
    }
    *out = t;
    return BVHT_OK;
}

int32_t bvht_tlas_set_objects(bvht_ctx ctx, bvht_tlas tlas, const bvht_instance* objects, uint32_t first,
                              uint32_t count)
{
    BVHT_REQUIRE(ctx && tlas && objects, "NULL argument");
    BVHT_REQUIRE((uint64_t)first + count <= tlas->n, "objects [%u, +%u) out of range (%u)", first, count, tlas->n);
    int32_t rc = validate_objects(tlas, objects, count);
    if (rc)
        return rc;
    BVHT_HIP(hipMemcpyAsync(tlas->d_objects + first, objects, sizeof(bvht_instance) * count, hipMemcpyHostToDevice,
                            ctx->stream));
    BVHT_HIP(hipStreamSynchronize(ctx->stream));
    return refresh(tlas);
}

int32_t bvht_tlas_get_objects(bvht_ctx ctx, bvht_tlas tlas, bvht_instance* out, uint32_t first, uint32_t count)
{
    BVHT_REQUIRE(ctx && tlas && (count == 0 || out), "NULL argument");
    BVHT_REQUIRE((uint64_t)first + count <= tlas->n, "objects [%u, +%u) out of range (%u)", first, count, tlas->n);
    BVHT_HIP(hipMemcpyAsync(out, tlas->d_objects + first, sizeof(bvht_instance) * count, hipMemcpyDeviceToHost,
                            ctx->stream));
    BVHT_HIP(hipStreamSynchronize(ctx->stream));
    return BVHT_OK;
}

int32_t bvht_tlas_refresh_bounds(bvht_ctx ctx, bvht_tlas tlas)
{
    BVHT_REQUIRE(ctx && tlas, "NULL argument");
    return refresh(tlas);
}

int32_t bvht_tlas_rebuild(bvht_ctx ctx, bvht_tlas tlas)
{
    BVHT_REQUIRE(ctx && tlas, "NULL argument");
    // the BLAS depths may have changed if a BLAS was rebuilt; refits keep them
    for (bvht_blas b : tlas->blases)
        tlas->max_blas_depth = std::max(tlas->max_blas_depth, b->depth);
    return rebuild(tlas, true);
}

int32_t bvht_tlas_rebuild_async(bvht_ctx ctx, bvht_tlas tlas)
{
    BVHT_REQUIRE(ctx && tlas, "NULL argument");
    BVHT_REQUIRE(tlas->n <= 64, "asynchronous rebuild needs at most 64 objects (%u)", tlas->n);
    for (bvht_blas b : tlas->blases)
        tlas->max_blas_depth = std::max(tlas->max_blas_depth, b->depth);
    // a TLAS over n objects has at most n levels: traces size their stack from that bound
    tlas->depth = tlas->n;
    return rebuild(tlas, false);
}

int32_t bvht_tlas_download(bvht_ctx ctx, bvht_tlas tlas, bvht_tnode* nodes, uint32_t* nodes_used)
{
    BVHT_REQUIRE(ctx && tlas, "NULL argument");
    if (nodes)
        BVHT_HIP(hipMemcpyAsync(nodes, tlas->d_nodes, sizeof(bvht_tnode) * 2 * tlas->n, hipMemcpyDeviceToHost,
                                ctx->stream));
    uint32_t meta[3] = {0, 0, 0};
    BVHT_HIP(hipMemcpyAsync(meta, tlas->d_meta, sizeof meta, hipMemcpyDeviceToHost, ctx->stream));
    BVHT_HIP(hipStreamSynchronize(ctx->stream));
    if (nodes_used)
        *nodes_used = meta[0];
    return BVHT_OK;
}

int32_t bvht_tlas_destroy(bvht_tlas tlas)
{
    if (!tlas)
        return BVHT_OK;
    (void)hipStreamSynchronize(tlas->ctx->stream);
    tlas_free(tlas);
    return BVHT_OK;
}

int32_t bvht_trace_tlas(bvht_ctx ctx, bvht_tlas tlas, const bvht_ray* d_rays, uint32_t n, bvht_hit* d_hits,
                        bvht_stats* d_stats)
{
    BVHT_REQUIRE(ctx && tlas, "NULL argument");
    BVHT_REQUIRE(n == 0 || (d_rays && d_hits), "NULL ray or hit buffer");
    return launch_trace_tlas(tlas, d_rays, n, nullptr, d_hits, d_stats);
}

int32_t bvht_trace_tlas_count(bvht_ctx ctx, bvht_tlas tlas, const bvht_ray* d_rays, const uint32_t* d_count,
                              uint32_t max_n, bvht_hit* d_hits, bvht_stats* d_stats)
{
    BVHT_REQUIRE(ctx && tlas && d_count, "NULL argument");
    BVHT_REQUIRE(max_n == 0 || (d_rays && d_hits), "NULL ray or hit buffer");
    return launch_trace_tlas(tlas, d_rays, max_n, d_count, d_hits, d_stats);
}

int32_t bvht_bounce_tlas(bvht_ctx ctx, bvht_tlas tlas, const bvht_ray* d_rays, const bvht_hit* d_hits, uint32_t n,
                         uint64_t seed, bvht_ray* d_out, uint32_t* d_count)
{
    BVHT_REQUIRE(ctx && tlas && d_count, "NULL argument");
    BVHT_REQUIRE(n == 0 || (d_rays && d_hits && d_out), "NULL buffer");
    uint32_t* bc = nullptr;
    int32_t rc = hit_block_offsets(ctx, d_hits, n, 0xFFFFFFFFu, tlas->n, tlas->d_rows, tlas->d_blas_table, d_count,
                                   &bc);
    if (rc || n == 0)
        return rc;
    hipLaunchKernelGGL(k_bounce_write_tlas, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream,
                       static_cast<const ObjRows*>(tlas->d_rows), tlas->n,
                       static_cast<const BlasRef*>(tlas->d_blas_table),
                       d_rays, d_hits, n, bc, seed, d_out);
    BVHT_HIP(hipGetLastError());
    return BVHT_OK;
}

int32_t bvht_bounce_tlas_sorted(bvht_ctx ctx, bvht_tlas tlas, const bvht_ray* d_rays, const bvht_hit* d_hits,
                                uint32_t n, uint64_t seed, bvht_ray* d_out, uint32_t* d_src, uint32_t* d_count)
{
    BVHT_REQUIRE(ctx && tlas && d_count && tlas->ctx == ctx, "NULL argument or TLAS of another context");
    BVHT_REQUIRE(n == 0 || (d_rays && d_hits && d_out), "NULL buffer");
    return bvht::launch_bounce_sorted(ctx, nullptr, 0xFFFFFFFFu, tlas->d_rows, tlas->d_blas_table, tlas->n, d_rays,
                                      d_hits, n, seed, d_out, d_src, d_count);
}

}  // extern "C"
