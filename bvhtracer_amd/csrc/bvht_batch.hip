// bvht_batch.hip -- many scenes per launch, and hipGraph capture of frames.
//
// A batch is a set of scenes (TLASes, each over its own BLASes) updated and
// traced with one launch per stage instead of one per scene:
//   * refit: every distinct BLAS of the batch, one launch per tree level
//     (the deepest level of all BLASes first), one thread per node;
//   * object bounds: every object of every scene, one launch;
//   * TLAS rebuild: one workgroup per scene, one launch;
//   * trace: a persistent kernel.  The rays of all scenes sit in one buffer,
//     scene s owning [ray_off[s], ray_off[s+1]); they are cut into work items
//     of `rays_per_item` rays; each wave takes the next item from an atomic
//     counter until the items run out (every wave reaches that exit), so
//     scenes of very different sizes share the machine without per-scene
//     launch tails.  Stack spill columns belong to lanes, not rays.
// No reference counterpart: the reference traces one scene, one ray at a
// time.  Per-ray results equal the per-scene trace exactly (the traversal is
// the same function).
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

struct BatchScene {
    const float4* tnodes;
    const ObjRows* rows;
    const BlasRef* blases;
    const bvht_instance* objects;
    float* bounds;
    ObjRows* rows_w;
    bvht_tnode* nodes_w;
    int32_t* list;
    uint32_t* meta;
    uint32_t n_obj;
    uint32_t pad;
};

struct BatchBlas {
    bvht_node* nodes;
    const float* tris;
};

__device__ __forceinline__ uint32_t upper_index(const uint32_t* prefix, uint32_t n, uint32_t x)
{
    // largest s in [0, n) with prefix[s] <= x (prefix ascending, prefix[0] = 0)
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (prefix[mid] <= x)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// One tree level of every BLAS of the batch: entries are (BLAS, node) pairs.
__global__ void k_batch_refit_level(const BatchBlas* __restrict__ blases, const uint2* __restrict__ entries,
                                    uint32_t count)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count)
        return;
    uint2 e = entries[i];
    BatchBlas bb = blases[e.x];
    refit_node(bb.nodes, bb.tris, e.y);
}

// Consecutive small levels in one workgroup, deepest first, a barrier
// between levels (one workgroup shares one CU, so its own global writes are
// visible to it after the barrier): level j is entries[begin[j], end[j]).
constexpr uint32_t kRunWG = 1024;
constexpr uint32_t kRunLevels = 32;
struct EntryRun {
    uint32_t levels;
    uint32_t begin[kRunLevels], end[kRunLevels];
};

__global__ __launch_bounds__(kRunWG) void k_batch_refit_levels(const BatchBlas* __restrict__ blases,
                                                               const uint2* __restrict__ entries, EntryRun run)
{
    for (uint32_t j = 0; j < run.levels; ++j) {
        for (uint32_t i = run.begin[j] + threadIdx.x; i < run.end[j]; i += kRunWG) {
            uint2 e = entries[i];
            BatchBlas bb = blases[e.x];
            refit_node(bb.nodes, bb.tris, e.y);
        }
        __syncthreads();
    }
}

// Big leaves of every BLAS of the batch, one wave per (BLAS, leaf) entry.
__global__ __launch_bounds__(256) void k_batch_refit_big(const BatchBlas* __restrict__ blases,
                                                         const uint2* __restrict__ entries, uint32_t count)
{
    uint32_t w = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= count)
        return;
    uint2 e = entries[w];
    BatchBlas bb = blases[e.x];
    refit_big_leaf(bb.nodes, bb.tris, e.y, threadIdx.x & 63u);
}

__global__ void k_batch_bounds(const BatchScene* __restrict__ scenes, const uint32_t* __restrict__ obj_prefix,
                               uint32_t n_scenes)
{
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= obj_prefix[n_scenes])
        return;
    uint32_t s = upper_index(obj_prefix, n_scenes, k);
    BatchScene sc = scenes[s];
    object_bounds(sc.objects, sc.blases, sc.bounds, sc.rows_w, k - obj_prefix[s]);
}

__global__ __launch_bounds__(kRebuildWG) void k_batch_rebuild(const BatchScene* __restrict__ scenes)
{
    BatchScene sc = scenes[blockIdx.x];
    tlas_rebuild_block(sc.bounds, sc.n_obj, sc.nodes_w, sc.list, sc.meta);
}

template <int WG, int LDS_DEPTH, int RPI>
__global__ __launch_bounds__(WG) void k_batch_trace(const BatchScene* __restrict__ scenes,
                                                    const uint32_t* __restrict__ item_prefix,
                                                    const uint32_t* __restrict__ ray_off, uint32_t n_scenes,
                                                    const bvht_ray* __restrict__ rays, bvht_hit* __restrict__ hits,
                                                    uint32_t* __restrict__ counter, uint32_t* __restrict__ spill,
                                                    uint32_t cap, uint32_t* __restrict__ status)
{
    __shared__ uint32_t lds[LDS_DEPTH * WG];
    uint32_t lane = threadIdx.x & 63;
    uint32_t total = item_prefix[n_scenes];
    uint32_t gtid = blockIdx.x * WG + threadIdx.x;
    Stack<LDS_DEPTH, WG> st{(lds_u32*)(lds + threadIdx.x), spill + gtid, (size_t)gridDim.x * WG, cap, status};
    while (true) {
        uint32_t item = 0;
        if (lane == 0)
            item = atomicAdd(counter, 1u);
        // readfirstlane makes the item (and so the loop exit) wave-uniform to
        // the compiler; with a divergent exit it may thread lane-0 code at
        // the end of the body into this fetch and leave lanes 1-63 in a loop
        // of their own
        item = __builtin_amdgcn_readfirstlane(__shfl(item, 0));
        if (item >= total)
            break;
        uint32_t s = upper_index(item_prefix, n_scenes, item);
        uint32_t base = ray_off[s] + (item - item_prefix[s]) * RPI;
        uint32_t end = min(ray_off[s + 1], base + RPI);
        BatchScene sc = scenes[s];
        for (uint32_t r = base + lane; r < end; r += 64) {
            DRay ray = load_ray(rays, r);
            uint32_t pairs = 0, tests = 0, tpairs = 0, visits = 0;
            TlasResult res =
                __builtin_amdgcn_ballot_w64(ray.t > kFltMax) == 0
                    ? traverse_tlas<false, false, true>(sc.tnodes, sc.rows, sc.blases, ray, st, pairs, tests, tpairs,
                                                        visits)
                    : traverse_tlas<false, false, false>(sc.tnodes, sc.rows, sc.blases, ray, st, pairs, tests, tpairs,
                                                         visits);
            store_hit(hits, r, res.found, res.t, res.u, res.v, res.prim, res.inst);
        }
    }
}

#define BVHT_BATCH_CONFIGS(X) \
    X(64, 16, 64) X(64, 16, 128) X(64, 16, 256) X(64, 24, 64) X(64, 24, 128) X(64, 24, 256) \
    X(256, 16, 64) X(256, 16, 128) X(256, 16, 256) X(256, 24, 64) X(256, 24, 128) X(256, 24, 256)

typedef void (*batch_trace_fn)(const BatchScene*, const uint32_t*, const uint32_t*, uint32_t, const bvht_ray*,
                               bvht_hit*, uint32_t*, uint32_t*, uint32_t, uint32_t*);

batch_trace_fn pick_batch(int wg, int lds, int rpi)
{
#define X(W, D, R)                               \
    if (wg == W && lds == D && rpi == R)         \
        return k_batch_trace<W, D, R>;
    BVHT_BATCH_CONFIGS(X)
#undef X
    return nullptr;
}

// y' = y + amp * sin(kx * x + w * frame) from the original vertices, with the
// frame number read from device memory (so a captured graph advances it).
__global__ void k_deform_wave(const float* __restrict__ orig, float* __restrict__ out, uint32_t n_vert, float amp,
                              float kx, float w, const uint32_t* __restrict__ frame)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_vert)
        return;
    float phase = w * (float)(*frame);
    float x = orig[3 * (size_t)i], y = orig[3 * (size_t)i + 1], z = orig[3 * (size_t)i + 2];
    out[3 * (size_t)i] = x;
    out[3 * (size_t)i + 1] = y + amp * sinf(kx * x + phase);
    out[3 * (size_t)i + 2] = z;
}

__global__ void k_add_u32(uint32_t* p, uint32_t v) { *p += v; }

}  // namespace

struct bvht_batch_s {
    bvht_ctx ctx = nullptr;
    std::vector<bvht_tlas> scenes;
    std::vector<bvht_blas> blases;
    BatchScene* d_scenes = nullptr;
    BatchBlas* d_blases = nullptr;
    uint2* d_refit = nullptr;           // (BLAS, node) of every used node, grouped by level
    uint2* d_refit_big = nullptr;       // (BLAS, leaf) of every big leaf
    uint32_t n_refit_big = 0;
    std::vector<uint32_t> refit_off;     // level L occupies [refit_off[L], refit_off[L+1])
    uint32_t* d_obj_prefix = nullptr;
    uint32_t total_nodes = 0, total_objs = 0;
    uint32_t cap = 0;  // traversal stack bound: max over scenes of (objects - 1) + (BLAS depth - 1)
    // trace tables (bvht_batch_set_rays)
    uint32_t* d_items = nullptr;
    uint32_t* d_ray_off = nullptr;
    uint32_t* d_counter = nullptr;
    uint32_t* d_spill = nullptr;
    size_t spill_words = 0;
    uint32_t n_rays = 0;
    int wg = 64, lds = 24, rpi = 128;
    uint32_t grid = 0;
};

struct bvht_graph_s {
    bvht_ctx ctx = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

void batch_free(bvht_batch b)
{
    (void)hipFree(b->d_scenes);
    (void)hipFree(b->d_blases);
    (void)hipFree(b->d_refit);
    (void)hipFree(b->d_refit_big);
    (void)hipFree(b->d_obj_prefix);
    (void)hipFree(b->d_items);
    (void)hipFree(b->d_ray_off);
    (void)hipFree(b->d_counter);
    (void)hipFree(b->d_spill);
    delete b;
}

int32_t batch_grid(bvht_batch b)
{
    batch_trace_fn fn = pick_batch(b->wg, b->lds, b->rpi);
    if (!fn)
        return set_error(BVHT_EINVAL, "unsupported batched trace configuration %dx%dx%d", b->wg, b->lds, b->rpi);
    int dev = b->ctx->device, cus = 0, per_cu = 0;
    BVHT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    BVHT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(fn), b->wg, 0));
    b->grid = (uint32_t)std::max(1, cus * std::max(per_cu, 1));
    size_t need = b->cap > (uint32_t)b->lds ? (size_t)(b->cap - b->lds) * b->grid * b->wg : 0;
    if (need)
        return ensure_spill(b->ctx, &b->d_spill, &b->spill_words, need);
    return BVHT_OK;
}

}  // namespace

extern "C" {

int32_t bvht_batch_create(bvht_ctx ctx, const bvht_tlas* scenes, uint32_t n_scenes, bvht_batch* out)
{
    BVHT_REQUIRE(ctx && scenes && out, "NULL argument");
    BVHT_REQUIRE(n_scenes > 0, "empty batch");
    bvht_batch b = new (std::nothrow) bvht_batch_s();
    if (!b)
        return set_error(BVHT_ENOMEM, "host allocation failed");
    b->ctx = ctx;
    std::vector<BatchScene> hs(n_scenes);
    std::vector<uint32_t> obj_prefix(n_scenes + 1, 0);
    for (uint32_t s = 0; s < n_scenes; ++s) {
        bvht_tlas t = scenes[s];
        if (!t || t->ctx != ctx) {
            delete b;
            return set_error(BVHT_EINVAL, "scene %u is NULL or belongs to another context", s);
        }
        b->scenes.push_back(t);
        BatchScene& d = hs[s];
        d.tnodes = reinterpret_cast<const float4*>(t->d_nodes);
        d.rows = static_cast<const ObjRows*>(t->d_rows);
        d.blases = static_cast<const BlasRef*>(t->d_blas_table);
        d.objects = t->d_objects;
        d.bounds = t->d_bounds;
        d.rows_w = static_cast<ObjRows*>(t->d_rows);
        d.nodes_w = t->d_nodes;
        d.list = t->d_list;
        d.meta = t->d_meta;
        d.n_obj = t->n;
        d.pad = 0;
        obj_prefix[s + 1] = obj_prefix[s] + t->n;
        uint32_t cap = (t->n - 1) + (t->max_blas_depth > 0 ? t->max_blas_depth - 1 : 0);
        b->cap = std::max(b->cap, cap);
        for (bvht_blas bl : t->blases)
            if (std::find(b->blases.begin(), b->blases.end(), bl) == b->blases.end())
                b->blases.push_back(bl);
    }
    std::vector<BatchBlas> hb(b->blases.size());
    uint32_t max_depth = 0;
    for (size_t k = 0; k < b->blases.size(); ++k) {
        bvht_blas bl = b->blases[k];
        hb[k].nodes = bl->d_nodes;
        hb[k].tris = bl->d_tris;
        max_depth = std::max(max_depth, bl->depth);
    }
    // combined level table: level L lists (BLAS, node) for every BLAS's level-L nodes
    std::vector<std::vector<uint32_t>> orders(b->blases.size());
    b->refit_off.assign(max_depth + 1, 0);
    for (size_t k = 0; k < b->blases.size(); ++k) {
        bvht_blas bl = b->blases[k];
        orders[k].resize(bl->level_off.back());
        if (hipMemcpy(orders[k].data(), bl->d_level_order, sizeof(uint32_t) * orders[k].size(),
                      hipMemcpyDeviceToHost) != hipSuccess) {
            delete b;
            return set_error(BVHT_EDEVICE, "batch: level table download failed");
        }
        for (uint32_t L = 0; L < bl->depth; ++L)
            b->refit_off[L + 1] += bl->level_off[L + 1] - bl->level_off[L];
    }
    for (uint32_t L = 0; L < max_depth; ++L)
        b->refit_off[L + 1] += b->refit_off[L];
    b->total_nodes = b->refit_off[max_depth];
    std::vector<uint2> entries(b->total_nodes);
    std::vector<uint32_t> cursor(b->refit_off.begin(), b->refit_off.end() - 1);
    for (size_t k = 0; k < b->blases.size(); ++k) {
        bvht_blas bl = b->blases[k];
        for (uint32_t L = 0; L < bl->depth; ++L)
            for (uint32_t i = bl->level_off[L]; i < bl->level_off[L + 1]; ++i)
                entries[cursor[L]++] = make_uint2((uint32_t)k, orders[k][i]);
    }
    std::vector<uint2> big;
    for (size_t k = 0; k < b->blases.size(); ++k) {
        bvht_blas bl = b->blases[k];
        std::vector<uint32_t> lst(bl->n_big_leaves);
        if (!lst.empty() && hipMemcpy(lst.data(), bl->d_big_leaves, sizeof(uint32_t) * lst.size(),
                                      hipMemcpyDeviceToHost) != hipSuccess) {
            delete b;
            return set_error(BVHT_EDEVICE, "batch: big-leaf list download failed");
        }
        for (uint32_t leaf : lst)
            big.push_back(make_uint2((uint32_t)k, leaf));
    }
    b->n_refit_big = (uint32_t)big.size();
    b->total_objs = obj_prefix.back();
    bool ok = hipMalloc(&b->d_scenes, sizeof(BatchScene) * n_scenes) == hipSuccess &&
              hipMalloc(&b->d_blases, sizeof(BatchBlas) * hb.size()) == hipSuccess &&
              hipMalloc(&b->d_refit, sizeof(uint2) * std::max<size_t>(entries.size(), 1)) == hipSuccess &&
              hipMalloc(&b->d_refit_big, sizeof(uint2) * std::max<size_t>(big.size(), 1)) == hipSuccess &&
              hipMalloc(&b->d_obj_prefix, sizeof(uint32_t) * obj_prefix.size()) == hipSuccess &&
              hipMalloc(&b->d_items, sizeof(uint32_t) * (n_scenes + 1)) == hipSuccess &&
              hipMalloc(&b->d_ray_off, sizeof(uint32_t) * (n_scenes + 1)) == hipSuccess &&
              hipMalloc(&b->d_counter, sizeof(uint32_t)) == hipSuccess;
    if (!ok) {
        batch_free(b);
        return set_error(BVHT_ENOMEM, "device allocation for a batch of %u scenes failed", n_scenes);
    }
    hipStream_t st = ctx->stream;
    if (hipMemcpyAsync(b->d_scenes, hs.data(), sizeof(BatchScene) * n_scenes, hipMemcpyHostToDevice, st) ||
        hipMemcpyAsync(b->d_blases, hb.data(), sizeof(BatchBlas) * hb.size(), hipMemcpyHostToDevice, st) ||
        hipMemcpyAsync(b->d_refit, entries.data(), sizeof(uint2) * entries.size(), hipMemcpyHostToDevice, st) ||
        (!big.empty() &&
         hipMemcpyAsync(b->d_refit_big, big.data(), sizeof(uint2) * big.size(), hipMemcpyHostToDevice, st)) ||
        hipMemcpyAsync(b->d_obj_prefix, obj_prefix.data(), sizeof(uint32_t) * obj_prefix.size(),
                       hipMemcpyHostToDevice, st) ||
        hipStreamSynchronize(st)) {
        batch_free(b);
        return set_error(BVHT_EDEVICE, "batch table upload failed");
    }
    int32_t rc = batch_grid(b);
    if (rc) {
        batch_free(b);
        return rc;
    }
    *out = b;
    return BVHT_OK;
}

int32_t bvht_batch_destroy(bvht_batch b)
{
    if (!b)
        return BVHT_OK;
    (void)hipStreamSynchronize(b->ctx->stream);
    batch_free(b);
    return BVHT_OK;
}

int32_t bvht_batch_refit(bvht_ctx ctx, bvht_batch b)
{
    BVHT_REQUIRE(ctx && b && b->ctx == ctx, "NULL argument or foreign context");
    if (b->n_refit_big)
        hipLaunchKernelGGL(k_batch_refit_big, dim3(nblk(b->n_refit_big, 4)), dim3(256), 0, ctx->stream, b->d_blases,
                           b->d_refit_big, b->n_refit_big);
    // a level of more than kRunWG entries gets a launch of its own; runs of
    // consecutive smaller levels share one single-workgroup launch
    EntryRun run{};
    auto flush = [&]() {
        if (run.levels)
            hipLaunchKernelGGL(k_batch_refit_levels, dim3(1), dim3(kRunWG), 0, ctx->stream, b->d_blases, b->d_refit,
                               run);
        run.levels = 0;
    };
    for (size_t L = b->refit_off.size() - 1; L-- > 0;) {
        uint32_t off = b->refit_off[L], cnt = b->refit_off[L + 1] - off;
        if (cnt == 0)
            continue;
        if (cnt > kRunWG) {
            flush();
            hipLaunchKernelGGL(k_batch_refit_level, dim3(nblk(cnt, 256)), dim3(256), 0, ctx->stream, b->d_blases,
                               b->d_refit + off, cnt);
            continue;
        }
        run.begin[run.levels] = off;
        run.end[run.levels] = off + cnt;
        if (++run.levels == kRunLevels)
            flush();
    }
    flush();
    BVHT_HIP(hipGetLastError());
    return BVHT_OK;
}

int32_t bvht_batch_rebuild(bvht_ctx ctx, bvht_batch b)
{
    BVHT_REQUIRE(ctx && b && b->ctx == ctx, "NULL argument or foreign context");
    // the depth of a rebuilt TLAS is not read back: per-scene traces size
    // their stack from the bound (a TLAS over n objects has at most n levels)
    for (bvht_tlas t : b->scenes) {
        t->depth = t->n;
        for (bvht_blas bl : t->blases)
            t->max_blas_depth = std::max(t->max_blas_depth, bl->depth);
    }
    hipLaunchKernelGGL(k_batch_bounds, dim3(nblk(b->total_objs, 64)), dim3(64), 0, ctx->stream, b->d_scenes,
                       b->d_obj_prefix, (uint32_t)b->scenes.size());
    hipLaunchKernelGGL(k_batch_rebuild, dim3((uint32_t)b->scenes.size()), dim3(kRebuildWG), 0, ctx->stream,
                       b->d_scenes);
    BVHT_HIP(hipGetLastError());
    return BVHT_OK;
}

int32_t bvht_batch_set_rays(bvht_ctx ctx, bvht_batch b, const uint32_t* ray_offsets, uint32_t rays_per_item)
{
    BVHT_REQUIRE(ctx && b && ray_offsets && b->ctx == ctx, "NULL argument or foreign context");
    uint32_t n = (uint32_t)b->scenes.size();
    BVHT_REQUIRE(ray_offsets[0] == 0, "ray_offsets[0] must be 0");
    std::vector<uint32_t> items(n + 1, 0);
    for (uint32_t s = 0; s < n; ++s) {
        BVHT_REQUIRE(ray_offsets[s + 1] >= ray_offsets[s], "ray_offsets must be non-decreasing");
        uint32_t cnt = ray_offsets[s + 1] - ray_offsets[s];
        items[s + 1] = items[s] + (cnt + rays_per_item - 1) / rays_per_item;
    }
    if (rays_per_item != (uint32_t)b->rpi) {
        BVHT_REQUIRE(pick_batch(b->wg, b->lds, (int)rays_per_item) != nullptr,
                     "unsupported rays_per_item %u (64, 128 or 256)", rays_per_item);
        b->rpi = (int)rays_per_item;
    }
    BVHT_HIP(hipMemcpyAsync(b->d_items, items.data(), sizeof(uint32_t) * (n + 1), hipMemcpyHostToDevice, ctx->stream));
    BVHT_HIP(hipMemcpyAsync(b->d_ray_off, ray_offsets, sizeof(uint32_t) * (n + 1), hipMemcpyHostToDevice,
                            ctx->stream));
    BVHT_HIP(hipStreamSynchronize(ctx->stream));
    b->n_rays = ray_offsets[n];
    return batch_grid(b);
}

int32_t bvht_batch_config(bvht_ctx ctx, bvht_batch b, int32_t wg, int32_t lds_depth, int32_t rays_per_item)
{
    BVHT_REQUIRE(ctx && b && b->ctx == ctx, "NULL argument or foreign context");
    BVHT_REQUIRE(pick_batch(wg, lds_depth, rays_per_item) != nullptr,
                 "unsupported batched trace configuration %dx%dx%d", wg, lds_depth, rays_per_item);
    BVHT_REQUIRE(rays_per_item == b->rpi || b->n_rays == 0,
                 "rays_per_item changes the work items: call bvht_batch_set_rays again");
    b->wg = wg;
    b->lds = lds_depth;
    b->rpi = rays_per_item;
    return batch_grid(b);
}

int32_t bvht_batch_trace(bvht_ctx ctx, bvht_batch b, const bvht_ray* d_rays, bvht_hit* d_hits)
{
    BVHT_REQUIRE(ctx && b && b->ctx == ctx, "NULL argument or foreign context");
    if (b->n_rays == 0)
        return BVHT_OK;
    BVHT_REQUIRE(d_rays && d_hits, "NULL ray or hit buffer");
    int32_t rc = fill_u32(ctx, b->d_counter, 0, 1);
    if (rc)
        return rc;
    batch_trace_fn fn = pick_batch(b->wg, b->lds, b->rpi);
    hipLaunchKernelGGL(fn, dim3(b->grid), dim3(b->wg), 0, ctx->stream, b->d_scenes, b->d_items, b->d_ray_off,
                       (uint32_t)b->scenes.size(), d_rays, d_hits, b->d_counter, b->d_spill,
                       stack_cap(ctx, b->cap), ctx->d_status);
    BVHT_HIP(hipGetLastError());
    return BVHT_OK;
}

int32_t bvht_deform_wave(bvht_ctx ctx, bvht_blas blas, const float* d_orig, float amp, float kx, float w,
                         const uint32_t* d_frame)
{
    BVHT_REQUIRE(ctx && blas && d_orig && d_frame, "NULL argument");
    uint32_t nv = 3 * blas->n_tris;
    hipLaunchKernelGGL(k_deform_wave, dim3(nblk(nv, 256)), dim3(256), 0, ctx->stream, d_orig, blas->d_tris, nv, amp,
                       kx, w, d_frame);
    BVHT_HIP(hipGetLastError());
    return BVHT_OK;
}

int32_t bvht_counter_add(bvht_ctx ctx, uint32_t* d_value, uint32_t v)
{
    BVHT_REQUIRE(ctx && d_value, "NULL argument");
    hipLaunchKernelGGL(k_add_u32, dim3(1), dim3(1), 0, ctx->stream, d_value, v);
    BVHT_HIP(hipGetLastError());
    return BVHT_OK;
}

int32_t bvht_graph_begin(bvht_ctx ctx)
{
    BVHT_REQUIRE(ctx, "ctx is NULL");
    BVHT_REQUIRE(!ctx->capturing, "a graph capture is already open on this context");
    BVHT_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    ctx->capturing = true;
    return BVHT_OK;
}

int32_t bvht_graph_end(bvht_ctx ctx, bvht_graph* out)
{
    BVHT_REQUIRE(ctx && out, "NULL argument");
    bvht_graph g = new (std::nothrow) bvht_graph_s();
    if (!g)
        return set_error(BVHT_ENOMEM, "host allocation failed");
    g->ctx = ctx;
    ctx->capturing = false;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g->graph);
    if (e != hipSuccess) {
        delete g;
        return set_error(BVHT_EDEVICE, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
    }
    e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g->graph);
        delete g;
        return set_error(BVHT_EDEVICE, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
    }
    ++ctx->live_graphs;
    *out = g;
    return BVHT_OK;
}

int32_t bvht_graph_launch(bvht_ctx ctx, bvht_graph g)
{
    BVHT_REQUIRE(ctx && g, "NULL argument");
    BVHT_HIP(hipGraphLaunch(g->exec, ctx->stream));
    return BVHT_OK;
}

int32_t bvht_graph_destroy(bvht_graph g)
{
    if (!g)
        return BVHT_OK;
    (void)hipGraphExecDestroy(g->exec);
    (void)hipGraphDestroy(g->graph);
    --g->ctx->live_graphs;
    delete g;
    return BVHT_OK;
}

}  // extern "C"
