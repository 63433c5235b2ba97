// bvht_core.hip -- context, BLAS storage, refit and the BLAS trace kernels
// for gfx950 (MI355X).
//
// Traversal kernels run one ray per lane (wave64).  Each lane keeps its
// traversal stack in LDS, laid out slot-major (slot * WG + lane) so the 64
// lanes of a wave hit 64 distinct banks; entries deeper than the LDS depth
// spill to a per-ray region of a global buffer.  The visiting order is the
// reference's exactly (bvh.rs:242-305): no root test, both children tested
// against the current closest t, nearer first (ties to the right child), the
// farther pushed only if hit, no re-cull on pop, first hit wins among equal t.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "bvht_device.hpp"
#include "bvht_internal.hpp"
#include "bvht_trace.hpp"
#include "bvht_update.hpp"

using namespace bvht;

namespace bvht {

static thread_local std::string g_error;

int32_t set_error(int32_t code, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

// A graph captured on ctx (or a capture in progress) holds the addresses of
// the context's buffers as kernel arguments: they must not move then.
static int32_t refuse_move(bvht_ctx ctx, const char* what, size_t bytes)
{
    return set_error(BVHT_EINVAL,
                     "%s must grow to %zu bytes, but %s on this context holds its address: make one eager call of "
                     "the largest size before capturing, or destroy the graphs first",
                     what, bytes, ctx->capturing ? "the open graph capture" : "a captured graph");
}

int32_t ensure_scratch(bvht_ctx ctx, size_t bytes)
{
    if (ctx->scratch_bytes >= bytes)
        return BVHT_OK;
    if (ctx->live_graphs > 0 || ctx->capturing)
        return refuse_move(ctx, "the context scratch", bytes);
    if (ctx->scratch) {
        BVHT_HIP(hipStreamSynchronize(ctx->stream));
        BVHT_HIP(hipFree(ctx->scratch));
        ctx->scratch = nullptr;
        ctx->scratch_bytes = 0;
    }
    size_t sz = std::max<size_t>(bytes, 1 << 20);
    if (hipMalloc(&ctx->scratch, sz) != hipSuccess)
        return set_error(BVHT_ENOMEM, "hipMalloc(%zu) for scratch failed", sz);
    ctx->scratch_bytes = sz;
    return BVHT_OK;
}

int32_t ensure_spill(bvht_ctx ctx, uint32_t** buf, size_t* words, size_t need)
{
    if (*words >= need)
        return BVHT_OK;
    if (ctx->live_graphs > 0 || ctx->capturing)
        return refuse_move(ctx, "a traversal stack spill buffer", need * sizeof(uint32_t));
    if (*buf) {
        BVHT_HIP(hipStreamSynchronize(ctx->stream));
        BVHT_HIP(hipFree(*buf));
        *buf = nullptr;
        *words = 0;
    }
    if (hipMalloc(reinterpret_cast<void**>(buf), need * sizeof(uint32_t)) != hipSuccess)
        return set_error(BVHT_ENOMEM, "hipMalloc(%zu) for the traversal spill failed", need * 4);
    *words = need;
    return BVHT_OK;
}

int32_t take_status(bvht_ctx ctx)
{
    // the stream is idle, so every kernel's store to the mapped word is visible
    uint32_t s = __atomic_exchange_n(ctx->h_status, 0u, __ATOMIC_ACQ_REL);
    if (s & kStatusStack)
        return set_error(BVHT_ESTACK,
                         "a trace dropped traversal stack entries at its stack bound: hits since the last fence may "
                         "be wrong (raise the stack limit, bvht_set_stack_limit)");
    return BVHT_OK;
}

}  // namespace bvht

// ---- fixed scratch of bvht_trace_camera_bounce_noalloc --------------------
// The allocation-free camera pass keeps its per-(bin, wave) counts and scan
// tile sums here: 24 counts per 64 pixels for up to kNoAllocMaxPixels pixels,
// plus one sum per 4096 counts.  Each GPU holds its own copy (6.0 MiB) for
// the life of the process.  Host code never names the array: it uses the
// address hipGetSymbolAddress gives on each device (noalloc_scratch).
constexpr uint32_t kNoAllocMaxPixels = 1u << 22;
constexpr size_t kNoAllocCounts = 24 * (size_t)(kNoAllocMaxPixels / 64);
constexpr size_t kNoAllocWords = kNoAllocCounts + (kNoAllocCounts + 4095) / 4096;
static __device__ __align__(256) uint32_t scratch[kNoAllocWords];

namespace bvht {

static std::mutex g_noalloc_mu;
static std::vector<void*> g_noalloc_ptr;  // device address of `scratch` on each device

// Called with `device` current (context creation).
static int32_t noalloc_scratch(int device, void** out)
{
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if ((size_t)device >= g_noalloc_ptr.size())
        g_noalloc_ptr.resize((size_t)device + 1, nullptr);
    if (!g_noalloc_ptr[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(BVHT_EDEVICE, "hipGetSymbolAddress of the fixed camera-pass scratch failed: %s",
                             hipGetErrorString(e));
        g_noalloc_ptr[device] = p;
    }
    *out = g_noalloc_ptr[device];
    return BVHT_OK;
}

}  // namespace bvht

// ------------------------------------------------------------- kernels --

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

// parent[] of every used node; interior nodes name their two children.
__global__ void k_parents(const bvht_node* __restrict__ nodes, uint32_t nodes_used, uint32_t* __restrict__ parent)
{
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nodes_used || k == 1)
        return;
    bvht_node nd = nodes[k];
    if (nd.count == 0) {
        parent[nd.left_first] = k;
        parent[nd.left_first + 1] = k;
    }
}

// Level of every used node (steps to the root) and the tree depth.  The
// depth is reduced over the wave first: one atomic per wave, not per node (a
// single counter shared by millions of nodes serialises otherwise).
__global__ __launch_bounds__(256) void k_levels(uint32_t nodes_used, const uint32_t* __restrict__ parent,
                                                uint32_t* __restrict__ level, uint32_t* __restrict__ depth)
{
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t m = 0;  // depth below this node's level + 1, 0 for no node
    if (k < nodes_used) {
        if (k == 1) {
            level[k] = BVHT_MISS;
        } else {
            uint32_t d = 0, p = parent[k];
            while (p != BVHT_MISS && d < nodes_used) {
                ++d;
                p = parent[p];
            }
            level[k] = d;
            m = d + 1;
        }
    }
    // every lane of the wave takes part (no early return above)
    for (int off = 32; off > 0; off >>= 1)
        m = max(m, (uint32_t)__shfl_xor((int)m, off, 64));
    if ((threadIdx.x & 63u) == 0 && m)
        atomicMax(depth, m);
}

// Per-level node counts and the level-grouped node order.  Levels below
// kLevelsLds are counted in LDS first, so each workgroup issues one global
// atomic per level instead of one per node (a few dozen counters shared by
// millions of nodes serialise otherwise); deeper levels (degenerate trees)
// go straight to the global counters.  Workgroups stride over the nodes.
// The order within a level is arbitrary: the nodes of one level are refit
// independently.  A level at or beyond `depth` (an inconsistent level array)
// is neither counted nor placed, which the caller's count check reports.
constexpr uint32_t kLevelsLds = 64;

__global__ void k_level_hist(uint32_t nodes_used, uint32_t depth, const uint32_t* __restrict__ level,
                             uint32_t* __restrict__ hist)
{
    __shared__ uint32_t sh[kLevelsLds];
    for (uint32_t i = threadIdx.x; i < kLevelsLds; i += blockDim.x)
        sh[i] = 0;
    __syncthreads();
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nodes_used; k += gridDim.x * blockDim.x) {
        if (k == 1)
            continue;
        uint32_t L = level[k];
        if (L < kLevelsLds && L < depth)
            atomicAdd(&sh[L], 1u);
        else if (L < depth)
            atomicAdd(&hist[L], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kLevelsLds; i += blockDim.x)
        if (sh[i])
            atomicAdd(&hist[i], sh[i]);
}

__global__ void k_level_scatter(uint32_t nodes_used, uint32_t depth, const uint32_t* __restrict__ level,
                                uint32_t* __restrict__ cursor, uint32_t* __restrict__ order)
{
    __shared__ uint32_t cnt[kLevelsLds], base[kLevelsLds];
    // each workgroup takes one contiguous range of nodes: counted first, so
    // it claims each level's slots with one global atomic, then placed
    uint32_t per = (nodes_used + gridDim.x - 1) / gridDim.x;
    uint32_t k0 = min(nodes_used, blockIdx.x * per), k1 = min(nodes_used, k0 + per);
    for (uint32_t i = threadIdx.x; i < kLevelsLds; i += blockDim.x)
        cnt[i] = 0;
    __syncthreads();
    for (uint32_t k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
        if (k == 1)
            continue;
        uint32_t L = level[k];
        if (L < kLevelsLds && L < depth)
            atomicAdd(&cnt[L], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kLevelsLds; i += blockDim.x) {
        base[i] = cnt[i] ? atomicAdd(&cursor[i], cnt[i]) : 0u;
        cnt[i] = 0;
    }
    __syncthreads();
    for (uint32_t k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
        if (k == 1)
            continue;
        uint32_t L = level[k];
        if (L < depth)
            order[L < kLevelsLds ? base[L] + atomicAdd(&cnt[L], 1u) : atomicAdd(&cursor[L], 1u)] = k;
    }
}

// Leaves with more than kBigLeaf triangles: counted, then listed.  `wide`
// (when given) is set if some leaf has 64 or more triangles, too many for a
// packed stack entry.
__global__ void k_big_leaves(const bvht_node* __restrict__ nodes, uint32_t nodes_used, uint32_t* __restrict__ counter,
                             uint32_t* __restrict__ list, uint32_t* __restrict__ wide)
{
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nodes_used || k == 1)
        return;
    uint32_t c = nodes[k].count;
    if (wide && c >= 64u)
        *wide = 1u;
    if (c <= kBigLeaf)
        return;
    uint32_t slot = atomicAdd(counter, 1u);
    if (list)
        list[slot] = k;
}

// Bvh::refit of the big leaves, one wave per leaf.
__global__ __launch_bounds__(256) void k_refit_big_leaves(bvht_node* __restrict__ nodes, const float* __restrict__ tris,
                                                          const uint32_t* __restrict__ list, uint32_t count)
{
    uint32_t w = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w < count)
        refit_big_leaf(nodes, tris, list[w], threadIdx.x & 63u);
}

// Bvh::refit, one tree level (see refit_node).
__global__ void k_refit_level(bvht_node* __restrict__ nodes, const float* __restrict__ tris,
                              const uint32_t* __restrict__ order, uint32_t count)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count)
        refit_node(nodes, tris, order[i]);
}

// Consecutive small levels of Bvh::refit in one workgroup: level j of the
// run is order[begin[j], end[j]), the run goes deepest first, and a barrier
// separates the levels (all threads of a workgroup share one CU, so its own
// global writes are visible to it after the barrier).
constexpr uint32_t kRunWG = 1024;     // threads of the one-workgroup pass
constexpr uint32_t kRunLevels = 32;   // levels per run
struct LevelRun {
    uint32_t levels;
    uint32_t begin[kRunLevels], end[kRunLevels];
};

__global__ __launch_bounds__(kRunWG) void k_refit_levels(bvht_node* __restrict__ nodes, const float* __restrict__ tris,
                                                         const uint32_t* __restrict__ order, LevelRun run)
{
    for (uint32_t j = 0; j < run.levels; ++j) {
        for (uint32_t i = run.begin[j] + threadIdx.x; i < run.end[j]; i += kRunWG)
            refit_node(nodes, tris, order[i]);
        __syncthreads();
    }
}

template <int WG, int LDS_DEPTH, bool STATS, bool PACKED>
__global__ __launch_bounds__(WG) void k_trace_blas(const float4* __restrict__ nodes, const float* __restrict__ tris,
                                                   const bvht_ray* __restrict__ rays, bvht_hit* __restrict__ hits,
                                                   uint32_t n, const uint32_t* __restrict__ d_n,
                                                   bvht_stats* __restrict__ stats, uint32_t* __restrict__ spill,
                                                   uint32_t cap, uint32_t* __restrict__ status,
                                                   const uint32_t* __restrict__ sched,
                                                   const uint32_t* __restrict__ sched_n, uint32_t* __restrict__ cost)
{
    __shared__ uint32_t lds[LDS_DEPTH * WG];
    // 64-ray group of this wave: slot order, or the order in `sched` when it
    // was made for this many groups; with `cost`, lane 0 records the group's
    // duration (bvht_set_wave_schedule)
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
    uint32_t pairs = 0, tests = 0;
    // a wave whose rays all have t <= FLT_MAX (or NaN) takes the reduced child choice
    BlasResult r = __builtin_amdgcn_ballot_w64(ray.t > kFltMax) == 0
                       ? traverse_blas<STATS, PACKED, true>(nodes, tris, ray, st, 0u, pairs, tests)
                       : traverse_blas<STATS, PACKED, false>(nodes, tris, ray, st, 0u, pairs, tests);
    store_hit(hits, i, r.found, r.t, r.u, r.v, r.prim, 0);
    if (STATS) {
        bvht_stats s;
        s.pairs = pairs;
        s.tests = tests;
        s.tlas_pairs = 0;
        s.inst_visits = 0;
        stats[i] = s;
    }
    if (cost && lane == 0) {
        uint64_t dt = (uint64_t)wall_clock64() - t0;
        cost[group] = dt > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)dt;
    }
}

typedef void (*trace_blas_fn)(const float4*, const float*, const bvht_ray*, bvht_hit*, uint32_t, const uint32_t*,
                              bvht_stats*, uint32_t*, uint32_t, uint32_t*, const uint32_t*, const uint32_t*,
                              uint32_t*);

static trace_blas_fn pick_trace_blas(int wg, int lds, bool stats, bool packed)
{
#define X(W, D)                                                                             \
    if (wg == W && lds == D) {                                                              \
        if (packed)                                                                         \
            return stats ? k_trace_blas<W, D, true, true> : k_trace_blas<W, D, false, true>;  \
        return stats ? k_trace_blas<W, D, true, false> : k_trace_blas<W, D, false, false>;    \
    }
    BVHT_TRACE_CONFIGS(X)
#undef X
    return nullptr;
}

// Camera::get_ray_eye + get_ray_world (camera.rs:994-1010) with the
// per-pixel mapping of PathTracer::evaluate (renderer.rs:353-362): ray j of a
// width x height frame in row or 8x8-tile order, world origin and direction.
__device__ __forceinline__ void camera_ray(const bvht_camera& cam, uint32_t width, uint32_t height, int32_t order,
                                           uint32_t j, V3& ow, V3& dw)
{
    uint32_t x, y;
    if (order == 1) {
        uint32_t tile = j >> 6, w = j & 63u, tiles_x = width >> 3;
        x = (tile % tiles_x) * 8 + (w & 7u);
        y = (tile / tiles_x) * 8 + (w >> 3);
    } else {
        x = j % width;
        y = j / width;
    }
    float u = (float)x / (float)width;
    float v = (float)y / (float)height;
    V3 origin = mk3(0.0f, 0.0f, 0.0f);
    V3 tl = mk3(cam.tl[0], cam.tl[1], cam.tl[2]);
    V3 tr = mk3(cam.tr[0], cam.tr[1], cam.tr[2]);
    V3 bl = mk3(cam.bl[0], cam.bl[1], cam.bl[2]);
    V3 pixel = add3(add3(add3(origin, tl), scl3(sub3(tr, tl), u)), scl3(sub3(bl, tl), v));
    V3 p = sub3(pixel, origin);
    float nrm = sqrtf(dot3(p, p));
    V3 dir = mk3(p.x / nrm, p.y / nrm, p.z / nrm);
    ow = mat_mul4(cam.view_inv, origin, 1.0f);
    dw = mat_mul4(cam.view_inv, dir, 0.0f);
}

__global__ void k_raygen(bvht_camera cam, uint32_t width, uint32_t height, int32_t order,
                         bvht_ray* __restrict__ rays)
{
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= width * height)
        return;
    V3 ow, dw;
    camera_ray(cam, width, height, order, j, ow, dw);
    float4* out = reinterpret_cast<float4*>(rays + j);
    out[0] = make_float4(ow.x, ow.y, ow.z, kFltMax);
    out[1] = make_float4(dw.x, dw.y, dw.z, 0.0f);
}

// Cosine-weighted bounce about the geometric normal of the hit triangle
// (flipped against the incoming direction), origin offset 1e-4 along it.
__device__ __forceinline__ void make_bounce(const float* tri, V3 o, V3 d, float t, uint64_t seed, uint32_t idx,
                                            float4& a, float4& b)
{
    V3 v0 = mk3(tri[0], tri[1], tri[2]);
    V3 e1 = sub3(mk3(tri[3], tri[4], tri[5]), v0);
    V3 e2 = sub3(mk3(tri[6], tri[7], tri[8]), v0);
    bounce_ray(cross3(e1, e2), add3(o, scl3(d, t)), d, seed, idx, a, b);
}

// Order-preserving compaction of hits: per-block counts, one-block scan, write.
__global__ __launch_bounds__(256) void k_hit_counts(const bvht_hit* __restrict__ hits, uint32_t n,
                                                    uint32_t prim_limit, uint32_t inst_limit,
                                                    const ObjRows* __restrict__ rows,
                                                    const BlasRef* __restrict__ blases,
                                                    uint32_t* __restrict__ block_counts)
{
    __shared__ uint32_t wave_counts[4];
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    bool h = false;
    if (i < n) {
        bvht_hit hh = hits[i];
        h = bounce_hit(hh, prim_limit, inst_limit, rows, blases);
    }
    uint64_t m = __ballot(h);
    if ((threadIdx.x & 63) == 0)
        wave_counts[threadIdx.x >> 6] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0)
        block_counts[blockIdx.x] = wave_counts[0] + wave_counts[1] + wave_counts[2] + wave_counts[3];
}

// Exclusive scan of a flat u32 array in place, in two launches: per-tile
// sums, then each tile rescans itself on top of the sum of the tiles before
// it.  A tile is 4096 entries, 16 consecutive ones per thread of a 256-thread
// block.  The total goes to *total (written by the last tile).
constexpr uint32_t kScanTile = 4096;

__device__ __forceinline__ void load16(const uint32_t* __restrict__ a, size_t i, uint32_t m, uint32_t (&v)[16])
{
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        size_t k = i + 4 * q;
        if (k + 4 <= m) {
            uint4 x = *reinterpret_cast<const uint4*>(a + k);
            v[4 * q] = x.x, v[4 * q + 1] = x.y, v[4 * q + 2] = x.z, v[4 * q + 3] = x.w;
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e)
                v[4 * q + e] = k + e < m ? a[k + e] : 0u;
        }
    }
}

__device__ __forceinline__ uint32_t block_sum_256(uint32_t v, uint32_t* red)
{
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = v;
    __syncthreads();
    uint32_t s = red[0] + red[1] + red[2] + red[3];
    __syncthreads();
    return s;
}

__global__ __launch_bounds__(256) void k_tile_sums(const uint32_t* __restrict__ a, uint32_t m,
                                                   uint32_t* __restrict__ sums)
{
    __shared__ uint32_t red[4];
    uint32_t v[16];
    load16(a, (size_t)blockIdx.x * kScanTile + 16 * threadIdx.x, m, v);
    uint32_t s = 0;
#pragma unroll
    for (int e = 0; e < 16; ++e)
        s += v[e];
    s = block_sum_256(s, red);
    if (threadIdx.x == 0)
        sums[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_tile_scan(uint32_t* __restrict__ a, uint32_t m,
                                                   const uint32_t* __restrict__ sums, uint32_t* __restrict__ total)
{
    __shared__ uint32_t red[4];
    __shared__ uint32_t wsum[4];
    uint32_t carry = 0;
    for (uint32_t k = threadIdx.x; k < blockIdx.x; k += 256)
        carry += sums[k];
    carry = block_sum_256(carry, red);
    size_t i = (size_t)blockIdx.x * kScanTile + 16 * threadIdx.x;
    uint32_t v[16];
    load16(a, i, m, v);
    uint32_t own = 0;
#pragma unroll
    for (int e = 0; e < 16; ++e)
        own += v[e];
    uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = own;  // inclusive scan over the wave
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off);
        if (lane >= (uint32_t)off)
            x += y;
    }
    if (lane == 63)
        wsum[w] = x;
    __syncthreads();
    uint32_t run = carry + x - own;
    for (uint32_t k = 0; k < w; ++k)
        run += wsum[k];
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        uint32_t c = v[e];
        v[e] = run;
        run += c;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        size_t k = i + 4 * q;
        if (k + 4 <= m) {
            *reinterpret_cast<uint4*>(a + k) = make_uint4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e)
                if (k + e < m)
                    a[k + e] = v[4 * q + e];
        }
    }
    if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0)
        *total = carry + wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

__global__ __launch_bounds__(256) void k_bounce_write(const float* __restrict__ tris, uint32_t ntris,
                                                      const bvht_ray* __restrict__ rays,
                                                      const bvht_hit* __restrict__ hits, uint32_t n,
                                                      const uint32_t* __restrict__ block_offsets,
                                                      uint64_t seed, bvht_ray* __restrict__ out)
{
    __shared__ uint32_t wave_counts[4];
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    bvht_hit h;
    bool hit = false;
    if (i < n) {
        h = hits[i];
        hit = bounce_hit(h, ntris, 0xFFFFFFFFu, nullptr, nullptr);
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
    DRay r = load_ray(rays, i);
    float4 a, b;
    make_bounce(tris + 9 * (size_t)h.prim, r.o, r.d, h.t, seed, i, a, b);
    float4* o = reinterpret_cast<float4*>(out + base + rank);
    o[0] = a;
    o[1] = b;
}

// ---- bounce rays grouped by direction ------------------------------------
// 24 direction bins: dominant axis (first of equal magnitudes) and its sign,
// then the signs of the other two components (in axis order after it).  A
// wave whose rays share a bin takes the same near/far child decisions more
// often; the traced rays and their hits are the same as in ray order.
constexpr uint32_t kDirBins = 24;

__device__ __forceinline__ uint32_t dir_bin(float dx, float dy, float dz)
{
    float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
    uint32_t axis = (ax >= ay && ax >= az) ? 0u : (ay >= az ? 1u : 2u);
    float c0 = axis == 0 ? dx : (axis == 1 ? dy : dz);
    float c1 = axis == 0 ? dy : (axis == 1 ? dz : dx);
    float c2 = axis == 0 ? dz : (axis == 1 ? dx : dy);
    return (axis * 2u + (c0 < 0.0f ? 1u : 0u)) * 4u + (c1 >= 0.0f ? 2u : 0u) + (c2 >= 0.0f ? 1u : 0u);
}

// Bounce ray of hit i for a BLAS trace (rows == nullptr) or a TLAS trace
// (normal carried to world space by the transpose of the object's M^-1);
// false where the hit is a miss or out of range.
struct BounceSrc {
    const float* tris;
    uint32_t ntris;
    const ObjRows* rows;
    const BlasRef* blases;
    uint32_t n_obj;
};

__device__ __forceinline__ bool bounce_for(const BounceSrc& s, const bvht_ray* __restrict__ rays, uint32_t n,
                                           const bvht_hit* __restrict__ hits, uint32_t i, uint64_t seed, float4& a,
                                           float4& b)
{
    if (i >= n)
        return false;
    bvht_hit h = hits[i];
    if (!bounce_hit(h, s.ntris, s.rows ? s.n_obj : 0xFFFFFFFFu, s.rows, s.blases))
        return false;
    DRay r = load_ray(rays, i);
    if (!s.rows) {
        make_bounce(s.tris + 9 * (size_t)h.prim, r.o, r.d, h.t, seed, i, a, b);
        return true;
    }
    ObjRows ob = s.rows[h.inst];
    const float* tri = s.blases[ob.blas].tris + 9 * (size_t)h.prim;
    V3 v0 = mk3(tri[0], tri[1], tri[2]);
    V3 nm = cross3(sub3(mk3(tri[3], tri[4], tri[5]), v0), sub3(mk3(tri[6], tri[7], tri[8]), v0));
    V3 nw = mk3(ob.r0.x * nm.x + ob.r1.x * nm.y + ob.r2.x * nm.z, ob.r0.y * nm.x + ob.r1.y * nm.y + ob.r2.y * nm.z,
                ob.r0.z * nm.x + ob.r1.z * nm.y + ob.r2.z * nm.z);
    bounce_ray(nw, add3(r.o, scl3(r.d, h.t)), r.d, seed, i, a, b);
    return true;
}

// Per (bin, 256-ray block) counts and each wave's rank inside its bin.
__device__ __forceinline__ void bin_ranks(uint32_t bin, bool valid, uint32_t (*wc)[kDirBins], uint32_t& rank)
{
    uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    unsigned long long below = (1ull << lane) - 1ull;
    rank = 0;
#pragma unroll 1
    for (uint32_t k = 0; k < kDirBins; ++k) {
        unsigned long long m = __ballot(valid && bin == k);
        if (lane == 0)
            wc[wave][k] = (uint32_t)__popcll(m);
        if (valid && bin == k)
            rank = (uint32_t)__popcll(m & below);
    }
}

__global__ __launch_bounds__(256) void k_bounce_bin_counts(BounceSrc src, const bvht_ray* __restrict__ rays,
                                                           const bvht_hit* __restrict__ hits, uint32_t n,
                                                           uint64_t seed, uint32_t nb, uint32_t* __restrict__ counts)
{
    __shared__ uint32_t wc[4][kDirBins];
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    float4 a, b;
    bool valid = bounce_for(src, rays, n, hits, i, seed, a, b);
    uint32_t bin = valid ? dir_bin(b.x, b.y, b.z) : 0u, rank;
    bin_ranks(bin, valid, wc, rank);
    __syncthreads();
    if (threadIdx.x < kDirBins)
        counts[(size_t)threadIdx.x * nb + blockIdx.x] =
            wc[0][threadIdx.x] + wc[1][threadIdx.x] + wc[2][threadIdx.x] + wc[3][threadIdx.x];
}

__global__ __launch_bounds__(256) void k_bounce_bin_write(BounceSrc src, const bvht_ray* __restrict__ rays,
                                                          const bvht_hit* __restrict__ hits, uint32_t n,
                                                          uint64_t seed, uint32_t nb,
                                                          const uint32_t* __restrict__ offsets,
                                                          bvht_ray* __restrict__ out, uint32_t* __restrict__ src_index)
{
    __shared__ uint32_t wc[4][kDirBins];
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    float4 a, b;
    bool valid = bounce_for(src, rays, n, hits, i, seed, a, b);
    uint32_t bin = valid ? dir_bin(b.x, b.y, b.z) : 0u, rank;
    bin_ranks(bin, valid, wc, rank);
    __syncthreads();
    if (!valid)
        return;
    uint32_t pos = offsets[(size_t)bin * nb + blockIdx.x] + rank;
    for (uint32_t w = 0; w < (threadIdx.x >> 6); ++w)
        pos += wc[w][bin];
    float4* o = reinterpret_cast<float4*>(out + pos);
    o[0] = a;
    o[1] = b;
    if (src_index)
        src_index[pos] = i;
}

// ---- camera rays, trace and direction-grouped bounce rays in one pass ------
// Lane j makes camera ray j (camera_ray), traces it and writes its hit.  A hit
// spawns bounce ray j (make_bounce, the same ray bounce_for makes from the
// stored ray and hit), in direction bin `bin`; lane k < 24 of each wave
// writes the wave's count of bin k to counts[k * n_waves + wave].  A scan of
// those bin-major counts gives every (bin, wave) its first output slot, so
// the rays come out exactly as bvht_bounce_sorted gives them: bins in order,
// ray order inside a bin.  Two ways to place them:
// - STAGED: the kernel also writes the bounce ray to a staging buffer in ray
//   order with key[j] = bin << 8 | rank among the wave's rays of that bin
//   (0xFFFFFFFF: no bounce), and k_bounce_scatter moves it to its slot;
// - not STAGED (no per-ray scratch): k_camera_bounce_regroup makes each
//   bounce ray again from the stored hit and the regenerated camera ray, and
//   ranks it with the same ballots.
//
// Wave order: the wave at grid position `slot` runs camera-ray group
// sched[slot] (rays 64 g .. 64 g + 63), or group slot without a schedule,
// and with `cost` its lane 0 records the wave's duration in cost[group]
// (s_memrealtime, 10 ns ticks).  Which wave runs a group changes nothing
// else: its rays, hits, counts and bounce rays are those of the group.
template <int WG, int LDS_DEPTH, bool PACKED, bool STAGED>
__global__ __launch_bounds__(WG) void k_trace_camera_bounce(const float4* __restrict__ nodes,
                                                            const float* __restrict__ tris, uint32_t ntris,
                                                            bvht_camera cam, uint32_t width, uint32_t height,
                                                            int32_t order, uint64_t seed, bvht_hit* __restrict__ hits,
                                                            float4* __restrict__ staged, uint32_t* __restrict__ key,
                                                            uint32_t* __restrict__ counts, uint32_t n_waves,
                                                            uint32_t* __restrict__ spill, uint32_t cap,
                                                            uint32_t* __restrict__ status,
                                                            const uint32_t* __restrict__ sched,
                                                            const uint32_t* __restrict__ sched_n,
                                                            uint32_t* __restrict__ cost)
{
    __shared__ uint32_t lds[LDS_DEPTH * WG];
    uint64_t t0 = (uint64_t)wall_clock64();
    uint32_t n = width * height;
    uint32_t slot = blockIdx.x * (WG / 64) + (threadIdx.x >> 6);
    uint32_t group = (sched && slot < n_waves && *sched_n == n_waves) ? sched[slot] : slot;
    uint32_t j = group * 64 + (threadIdx.x & 63u);
    bool valid = false;
    uint32_t bin = 0;
    if (j < n) {
        Stack<LDS_DEPTH, WG> st{(lds_u32*)(lds + threadIdx.x), spill + j, (size_t)n, cap, status};
        V3 ow, dw;
        camera_ray(cam, width, height, order, j, ow, dw);
        DRay ray = make_ray(ow, dw, kFltMax);
        uint32_t pairs = 0, tests = 0;
        BlasResult r = traverse_blas<false, PACKED, true>(nodes, tris, ray, st, 0u, pairs, tests);
        store_hit(hits, j, r.found, r.t, r.u, r.v, r.prim, 0);
        if (r.found && r.prim < ntris) {
            float4 a, b;
            make_bounce(tris + 9 * (size_t)r.prim, ray.o, ray.d, r.t, seed, j, a, b);
            if (STAGED) {
                staged[2 * (size_t)j] = a;
                staged[2 * (size_t)j + 1] = b;
            }
            bin = dir_bin(b.x, b.y, b.z);
            valid = true;
        }
    }
    uint32_t lane = threadIdx.x & 63u;
    unsigned long long below = (1ull << lane) - 1ull;
    uint32_t rank = 0, wave_count = 0;
    // not unrolled: 24 live ballot masks would cost a wave slot in SGPRs
#pragma unroll 1
    for (uint32_t k = 0; k < kDirBins; ++k) {
        unsigned long long m = __ballot(valid && bin == k);
        if (STAGED && valid && bin == k)
            rank = (uint32_t)__popcll(m & below);
        if (lane == k)
            wave_count = (uint32_t)__popcll(m);
    }
    if (group < n_waves && lane < kDirBins)
        counts[(size_t)lane * n_waves + group] = wave_count;
    if (STAGED && j < n)
        key[j] = valid ? (bin << 8 | rank) : 0xFFFFFFFFu;
    if (cost && lane == 0 && group < n_waves) {
        uint64_t dt = (uint64_t)wall_clock64() - t0;
        cost[group] = dt > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)dt;
    }
}

__global__ __launch_bounds__(256) void k_bounce_scatter(const float4* __restrict__ staged,
                                                        const uint32_t* __restrict__ key, uint32_t n,
                                                        uint32_t n_waves, const uint32_t* __restrict__ offsets,
                                                        bvht_ray* __restrict__ out, uint32_t* __restrict__ src_index)
{
    uint32_t j = blockIdx.x * 256 + threadIdx.x;
    if (j >= n)
        return;
    uint32_t k = key[j];
    if (k == 0xFFFFFFFFu)
        return;
    uint32_t bin = k >> 8;
    uint32_t pos = offsets[(size_t)bin * n_waves + (j >> 6)] + (k & 255u);
    float4* o = reinterpret_cast<float4*>(out + pos);
    o[0] = staged[2 * (size_t)j];
    o[1] = staged[2 * (size_t)j + 1];
    if (src_index)
        src_index[pos] = j;
}

// Bounce ray j of a camera frame from its stored hit: the camera ray again
// (camera_ray is exact arithmetic on (j, camera), so it is the traced ray),
// the hit triangle and t, the same make_bounce as the camera pass; its slot
// is the scanned (bin, wave) offset plus its rank among the wave's rays of
// that bin.  Waves are the camera pass's (64 consecutive rays).
__global__ __launch_bounds__(256) void k_camera_bounce_regroup(const float* __restrict__ tris, uint32_t ntris,
                                                               bvht_camera cam, uint32_t width, uint32_t height,
                                                               int32_t order, uint64_t seed,
                                                               const bvht_hit* __restrict__ hits, uint32_t n_waves,
                                                               const uint32_t* __restrict__ offsets,
                                                               bvht_ray* __restrict__ out,
                                                               uint32_t* __restrict__ src_index)
{
    uint32_t n = width * height;
    uint32_t j = blockIdx.x * 256 + threadIdx.x;
    bool valid = false;
    uint32_t bin = 0;
    float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f), b = a;
    if (j < n) {
        const bvht_hit* h = hits + j;
        uint32_t prim = h->prim;
        if (prim != BVHT_MISS && prim < ntris) {
            V3 ow, dw;
            camera_ray(cam, width, height, order, j, ow, dw);
            make_bounce(tris + 9 * (size_t)prim, ow, dw, h->t, seed, j, a, b);
            bin = dir_bin(b.x, b.y, b.z);
            valid = true;
        }
    }
    uint32_t lane = threadIdx.x & 63u;
    unsigned long long below = (1ull << lane) - 1ull;
    uint32_t rank = 0;
#pragma unroll 1
    for (uint32_t k = 0; k < kDirBins; ++k) {
        unsigned long long m = __ballot(valid && bin == k);
        if (valid && bin == k)
            rank = (uint32_t)__popcll(m & below);
    }
    if (!valid)
        return;
    uint32_t pos = offsets[(size_t)bin * n_waves + (j >> 6)] + rank;
    float4* o = reinterpret_cast<float4*>(out + pos);
    o[0] = a;
    o[1] = b;
    if (src_index)
        src_index[pos] = j;
}

typedef void (*camera_bounce_fn)(const float4*, const float*, uint32_t, bvht_camera, uint32_t, uint32_t, int32_t,
                                 uint64_t, bvht_hit*, float4*, uint32_t*, uint32_t*, uint32_t, uint32_t*, uint32_t,
                                 uint32_t*, const uint32_t*, const uint32_t*, uint32_t*);

static camera_bounce_fn pick_camera_bounce(int wg, int lds, bool packed, bool staged)
{
#define X(W, D)                                                                                          \
    if (wg == W && lds == D) {                                                                           \
        if (staged)                                                                                      \
            return packed ? k_trace_camera_bounce<W, D, true, true> : k_trace_camera_bounce<W, D, false, true>; \
        return packed ? k_trace_camera_bounce<W, D, true, false> : k_trace_camera_bounce<W, D, false, false>;   \
    }
    BVHT_TRACE_CONFIGS(X)
#undef X
    return nullptr;
}

__global__ __launch_bounds__(256) void k_stream_copy(float4* __restrict__ dst, const float4* __restrict__ src,
                                                     size_t n)
{
    size_t stride = (size_t)gridDim.x * 256;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride)
        dst[i] = src[i];
}

__global__ void k_depth_acc(const bvht_hit* __restrict__ hits, uint32_t n, float* __restrict__ depth)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        depth[i] = hits[i].prim != BVHT_MISS ? hits[i].t : kFltMax;
}

// ------------------------------------------------------------ host side --

namespace bvht {

// Longest-first order of work units (64-ray groups, batch items) from the
// durations their last run recorded: a counting sort on a log scale (16 bins
// per octave, longest bin first).  The order inside a bin comes from atomics
// and is not fixed; the schedule is always a permutation of the units.
__device__ __forceinline__ uint32_t sched_bin(uint32_t ticks)
{
    // floor(16 log2(ticks + 1)) from the leading bit and the next four bits,
    // reversed so that long waves come first
    uint64_t v = (uint64_t)ticks + 1u;
    uint32_t msb = 63u - (uint32_t)__builtin_clzll(v);
    uint32_t frac = msb >= 4 ? (uint32_t)(v >> (msb - 4)) & 15u : (uint32_t)(v << (4 - msb)) & 15u;
    uint32_t b = msb * 16u + frac;
    return kSchedBins - 1 - (b < kSchedBins ? b : kSchedBins - 1);
}

__global__ __launch_bounds__(256) void k_sched_count(const uint32_t* __restrict__ cost, uint32_t units,
                                                     uint32_t* __restrict__ bins)
{
    __shared__ uint32_t h[kSchedBins];
    for (uint32_t k = threadIdx.x; k < kSchedBins; k += 256)
        h[k] = 0;
    __syncthreads();
    for (uint32_t g = blockIdx.x * 256 + threadIdx.x; g < units; g += gridDim.x * 256)
        atomicAdd(&h[sched_bin(cost[g])], 1u);
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < kSchedBins; k += 256)
        if (h[k])
            atomicAdd(&bins[k], h[k]);
}

__global__ __launch_bounds__(kSchedBins) void k_sched_offsets(uint32_t* __restrict__ bins)
{
    __shared__ uint32_t part[kSchedBins];
    uint32_t t = threadIdx.x;
    part[t] = bins[t];
    __syncthreads();
    for (uint32_t off = 1; off < kSchedBins; off <<= 1) {
        uint32_t v = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    bins[kSchedBins + t] = t ? part[t - 1] : 0u;  // cursor: exclusive prefix
}

// Each block ranks its units per bin in LDS and reserves one range per
// non-empty bin with a single global atomic (most units fall in a few bins,
// so one global atomic per unit would serialise on those counters).
__global__ __launch_bounds__(256) void k_sched_place(const uint32_t* __restrict__ cost, uint32_t units,
                                                     uint32_t* __restrict__ bins, uint32_t* __restrict__ sched)
{
    __shared__ uint32_t h[kSchedBins];
    for (uint32_t k = threadIdx.x; k < kSchedBins; k += 256)
        h[k] = 0;
    __syncthreads();
    uint32_t g = blockIdx.x * 256 + threadIdx.x;
    bool valid = g < units;
    uint32_t b = valid ? sched_bin(cost[g]) : 0u;
    uint32_t rank = valid ? atomicAdd(&h[b], 1u) : 0u;
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < kSchedBins; k += 256)
        if (h[k])
            h[k] = atomicAdd(&bins[kSchedBins + k], h[k]);
    __syncthreads();
    if (valid)
        sched[h[b] + rank] = g;
    if (blockIdx.x == 0 && threadIdx.x == 0)
        bins[2 * kSchedBins] = units;  // the order's unit count (sched_units)
}

static inline uint32_t blocks_for(size_t n, uint32_t wg) { return (uint32_t)((n + wg - 1) / wg); }

int32_t trace_order_begin(bvht_ctx ctx, const void* tree, const void* rays, uint32_t n, bool eligible, TraceOrder* o)
{
    *o = TraceOrder{};
    if (!ctx->tls_on || !eligible || n == 0)
        return BVHT_OK;
    uint32_t groups = (n + 63) / 64;
    if (groups > ctx->tls_cap) {
        if (ctx->live_graphs > 0 || ctx->capturing)
            return set_error(BVHT_EINVAL,
                             "the wave order must grow to %u groups, but a graph captured on this context holds its "
                             "address: make one eager call of the largest ray count first",
                             groups);
        BVHT_HIP(hipStreamSynchronize(ctx->stream));
        (void)hipFree(ctx->tls_cost);
        (void)hipFree(ctx->tls_sched);
        ctx->tls_cost = ctx->tls_sched = nullptr;
        ctx->tls_cap = 0;
        ctx->tls_valid = false;
        if (hipMalloc(&ctx->tls_cost, sizeof(uint32_t) * groups) != hipSuccess ||
            hipMalloc(&ctx->tls_sched, sizeof(uint32_t) * groups) != hipSuccess ||
            (!ctx->tls_bins && hipMalloc(&ctx->tls_bins, sizeof(uint32_t) * kSchedBinsWords) != hipSuccess))
            return set_error(BVHT_ENOMEM, "wave order for %u groups", groups);
        ctx->tls_cap = groups;
    }
    bool same = ctx->tls_valid && ctx->tls_tlas == tree && ctx->tls_rays == rays && ctx->tls_n == n;
    o->sched = same ? ctx->tls_sched : nullptr;
    o->sched_n = sched_units(ctx->tls_bins);
    o->cost = ctx->tls_cost;
    o->groups = groups;
    o->on = true;
    ctx->tls_calls = same ? ctx->tls_calls + 1 : 0;
    return BVHT_OK;
}

int32_t trace_order_end(bvht_ctx ctx, const void* tree, const void* rays, uint32_t n, const TraceOrder& o)
{
    if (!o.on)
        return BVHT_OK;
    // rebuilt after the first call and then every 8 calls, and in every
    // captured call (a graph replays what it captured)
    constexpr uint32_t kResched = 8;
    if (ctx->tls_calls % kResched == 0 || ctx->capturing) {
        int32_t rc = launch_longest_first(ctx, ctx->tls_cost, o.groups, ctx->tls_bins, ctx->tls_sched);
        if (rc)
            return rc;
    }
    ctx->tls_valid = true;
    ctx->tls_tlas = tree;
    ctx->tls_rays = rays;
    ctx->tls_n = n;
    return BVHT_OK;
}

int32_t launch_longest_first(bvht_ctx ctx, const uint32_t* d_cost, uint32_t n, uint32_t* d_bins, uint32_t* d_sched)
{
    if (n == 0)
        return BVHT_OK;
    int32_t rc = fill_u32(ctx, d_bins, 0u, kSchedBins);
    if (rc)
        return rc;
    hipLaunchKernelGGL(k_sched_count, dim3(std::min<uint32_t>(blocks_for(n, 256), 64u)), dim3(256), 0, ctx->stream,
                       d_cost, n, d_bins);
    hipLaunchKernelGGL(k_sched_offsets, dim3(1), dim3(kSchedBins), 0, ctx->stream, d_bins);
    hipLaunchKernelGGL(k_sched_place, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, d_cost, n, d_bins, d_sched);
    BVHT_HIP(hipGetLastError());
    return BVHT_OK;
}
static inline size_t align256(size_t bytes) { return (bytes + 255) & ~(size_t)255; }

// In-place exclusive scan of a[0, m) (m > 0, a 16-byte aligned), total to
// *total; sums holds blocks_for(m, kScanTile) entries.
uint32_t scan_sums_entries(uint32_t m) { return std::max<uint32_t>(blocks_for(m, kScanTile), 1); }

void launch_scan(bvht_ctx ctx, uint32_t* a, uint32_t m, uint32_t* sums, uint32_t* total)
{
    uint32_t tiles = blocks_for(m, kScanTile);
    hipLaunchKernelGGL(k_tile_sums, dim3(tiles), dim3(256), 0, ctx->stream, a, m, sums);
    hipLaunchKernelGGL(k_tile_scan, dim3(tiles), dim3(256), 0, ctx->stream, a, m, sums, total);
}

int32_t fill_u32(bvht_ctx ctx, uint32_t* p, uint32_t v, size_t n)
{
    if (n == 0)
        return BVHT_OK;
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, p, v, n);
    BVHT_HIP(hipGetLastError());
    return BVHT_OK;
}

// Parent array, node levels, depth and the level-grouped node order on the
// device (after a build or an upload); the per-level node counts come back to
// the host to drive the level-by-level refit.
// Host copies through the pinned staging: chunks of kStageChunk bytes
// alternate between the two halves.  Below kStageMin bytes, or while the
// stream is being captured, the runtime's own pageable copy is used.
constexpr size_t kStageChunk = size_t(8) << 20;
constexpr size_t kStageMin = size_t(4) << 20;
constexpr unsigned kStageThreads = 4;

// Helper threads of a context's staged copies: kStageThreads - 1 threads,
// each copying its slice of every chunk; the calling thread copies the first
// slice.  They start with the staging and stay until bvht_destroy (starting
// threads per chunk had cost about 85 us for three).
struct CopyPool {
    struct Slice {
        char* dst;
        const char* src;
        size_t n;
    };
    std::mutex m;
    std::condition_variable go, done;
    Slice slice[kStageThreads];
    uint64_t round = 0;  // rounds posted
    unsigned busy = 0;   // helpers still copying the current round
    bool stop = false;
    std::vector<std::thread> th;
};

static void copy_helper(CopyPool* p, unsigned id)
{
    uint64_t seen = 0;
    std::unique_lock<std::mutex> lk(p->m);
    for (;;) {
        p->go.wait(lk, [&] { return p->stop || p->round != seen; });
        if (p->stop)
            return;
        seen = p->round;
        const CopyPool::Slice sl = p->slice[id];
        lk.unlock();
        if (sl.n)
            std::memcpy(sl.dst, sl.src, sl.n);
        lk.lock();
        if (--p->busy == 0)
            p->done.notify_one();
    }
}

static CopyPool* start_copy_pool()
{
    CopyPool* p = new (std::nothrow) CopyPool;
    if (!p)
        return nullptr;
    unsigned want = std::min(kStageThreads, std::max(1u, std::thread::hardware_concurrency())) - 1;
    try {
        p->th.reserve(want);
        for (unsigned t = 0; t < want; ++t)
            p->th.emplace_back(copy_helper, p, t);
    } catch (...) {  // fewer helpers: the rounds use those that started
    }
    return p;
}

void stop_copy_pool(CopyPool* p)
{
    if (!p)
        return;
    {
        std::lock_guard<std::mutex> lk(p->m);
        p->stop = true;
    }
    p->go.notify_all();
    for (std::thread& t : p->th)
        if (t.joinable())
            t.join();
    delete p;
}

bool ensure_stage(bvht_ctx ctx)
{
    if (ctx->h_stage)
        return true;
    for (hipEvent_t& e : ctx->stage_ev) {
        if (!e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
            return false;
        if (hipEventRecord(e, ctx->stream) != hipSuccess)
            return false;
    }
    void* p = nullptr;
    if (hipHostMalloc(&p, 2 * kStageChunk, hipHostMallocDefault) != hipSuccess)
        return false;
    ctx->h_stage = static_cast<char*>(p);
    ctx->stage_half = kStageChunk;
    if (!ctx->copy_pool)
        ctx->copy_pool = start_copy_pool();
    return true;
}

// memcpy split over the context's helper threads and the calling one, in
// 4 KiB-aligned slices (one plain memcpy without helpers).
void split_memcpy(bvht_ctx ctx, char* dst, const char* src, size_t n)
{
    CopyPool* p = ctx->copy_pool;
    if (!p || p->th.empty()) {
        std::memcpy(dst, src, n);
        return;
    }
    const unsigned nt = 1 + (unsigned)p->th.size();
    const size_t per = ((n + nt - 1) / nt + 4095) & ~(size_t)4095;
    try {
        std::unique_lock<std::mutex> lk(p->m);
        for (unsigned t = 1; t < nt; ++t) {
            size_t a = std::min(n, t * per), b = std::min(n, a + per);
            p->slice[t - 1] = CopyPool::Slice{dst + a, src + a, b - a};
        }
        p->busy = nt - 1;
        ++p->round;
        lk.unlock();
        p->go.notify_all();
        std::memcpy(dst, src, std::min(n, per));
        lk.lock();
        p->done.wait(lk, [&] { return p->busy == 0; });
    } catch (...) {  // a lock that cannot be taken: the whole copy here
        std::memcpy(dst, src, n);
    }
}

int32_t host_to_device(bvht_ctx ctx, void* dst, const void* src, size_t bytes)
{
    hipStream_t st = ctx->stream;
    if (bytes < kStageMin || ctx->capturing || !ensure_stage(ctx)) {
        BVHT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
        return BVHT_OK;
    }
    const char* s = static_cast<const char*>(src);
    char* d = static_cast<char*>(dst);
    for (size_t off = 0, i = 0; off < bytes; off += ctx->stage_half, ++i) {
        size_t len = std::min(ctx->stage_half, bytes - off);
        char* half = ctx->h_stage + (i % 2) * ctx->stage_half;
        BVHT_HIP(hipEventSynchronize(ctx->stage_ev[i % 2]));  // the half's last transfer is done
        split_memcpy(ctx, half, s + off, len);
        BVHT_HIP(hipMemcpyAsync(d + off, half, len, hipMemcpyHostToDevice, st));
        BVHT_HIP(hipEventRecord(ctx->stage_ev[i % 2], st));
    }
    return BVHT_OK;
}

static int32_t device_to_host_on(bvht_ctx ctx, hipStream_t st, void* dst, const void* src, size_t bytes)
{
    if (bytes < kStageMin || ctx->capturing || !ensure_stage(ctx)) {
        BVHT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
        BVHT_HIP(hipStreamSynchronize(st));
        return BVHT_OK;
    }
    const char* s = static_cast<const char*>(src);
    char* d = static_cast<char*>(dst);
    const size_t H = ctx->stage_half;
    size_t nch = (bytes + H - 1) / H;
    // the DMA of chunk i + 1 runs while the host copies chunk i out (the
    // stream orders it after any earlier transfer through the same half)
    auto issue = [&](size_t i) -> hipError_t {
        hipError_t e = hipMemcpyAsync(ctx->h_stage + (i % 2) * H, s + i * H, std::min(H, bytes - i * H),
                                      hipMemcpyDeviceToHost, st);
        return e != hipSuccess ? e : hipEventRecord(ctx->stage_ev[i % 2], st);
    };
    BVHT_HIP(issue(0));
    for (size_t i = 0; i < nch; ++i) {
        if (i + 1 < nch)
            BVHT_HIP(issue(i + 1));
        BVHT_HIP(hipEventSynchronize(ctx->stage_ev[i % 2]));
        split_memcpy(ctx, d + i * H, ctx->h_stage + (i % 2) * H, std::min(H, bytes - i * H));
    }
    return BVHT_OK;
}

int32_t device_to_host(bvht_ctx ctx, void* dst, const void* src, size_t bytes)
{
    return device_to_host_on(ctx, ctx->stream, dst, src, bytes);
}

int32_t mark_for_copy(bvht_ctx ctx)
{
    if (!ctx->copy_ev)
        BVHT_HIP(hipEventCreateWithFlags(&ctx->copy_ev, hipEventDisableTiming));
    BVHT_HIP(hipEventRecord(ctx->copy_ev, ctx->stream));
    return BVHT_OK;
}

// One copy stream per device for the process (default priority), made on
// first use and kept.  Streams map onto a few hardware queues
// (GPU_MAX_HW_QUEUES, 4 by default): a copy stream per context would make
// the streams of several contexts share queues, so that their work no
// longer overlaps, and a stream made per copy cost about 0.2 ms per build.
static hipStream_t process_copy_stream(int device)
{
    static std::mutex m;
    static hipStream_t streams[64] = {};
    if (device < 0 || device >= 64)
        return nullptr;
    std::lock_guard<std::mutex> lk(m);
    if (!streams[device] && hipStreamCreateWithFlags(&streams[device], hipStreamNonBlocking) != hipSuccess)
        streams[device] = nullptr;
    return streams[device];
}

int32_t device_to_host_beside(bvht_ctx ctx, void* dst, const void* src, size_t bytes)
{
    // a copy below the staging threshold goes through the runtime's pageable
    // path, which measured slower from a second stream (82k-triangle build
    // 1.92 -> 2.06 ms): it runs on ctx's stream after all queued work
    hipStream_t cs = nullptr;
    if (bytes < kStageMin || ctx->capturing || !ctx->copy_ev || !ensure_stage(ctx) ||
        !(cs = process_copy_stream(ctx->device)))
        return device_to_host_on(ctx, ctx->stream, dst, src, bytes);
    // every earlier transfer through this context's staging was queued on
    // ctx's stream before the mark, and device_to_host_on returns only once
    // its own transfers are done
    BVHT_HIP(hipStreamWaitEvent(cs, ctx->copy_ev, 0));
    return device_to_host_on(ctx, cs, dst, src, bytes);
}

int32_t compute_blas_topology(bvht_blas b, const bvht_node*, uint32_t known_depth)
{
    bvht_ctx ctx = b->ctx;
    hipStream_t st = ctx->stream;
    int32_t rc = fill_u32(ctx, b->d_parent, BVHT_MISS, b->node_count);
    if (rc)
        return rc;
    uint32_t nb = blocks_for(b->nodes_used, 256);
    hipLaunchKernelGGL(k_parents, dim3(nb), dim3(256), 0, st, b->d_nodes, b->nodes_used, b->d_parent);
    // scratch: [0] depth, then the per-level counts, then the big-leaf count
    // and the wide-leaf flag
    rc = ensure_scratch(ctx, sizeof(uint32_t) * (2 * (size_t)b->nodes_used + 4));
    if (rc)
        return rc;
    uint32_t* d_depth = static_cast<uint32_t*>(ctx->scratch);
    uint32_t depth = known_depth;
    if (depth == 0) {
        rc = fill_u32(ctx, d_depth, 0, 1);
        if (rc)
            return rc;
        hipLaunchKernelGGL(k_levels, dim3(nb), dim3(256), 0, st, b->nodes_used, b->d_parent, b->d_level, d_depth);
        BVHT_HIP(hipGetLastError());
        BVHT_HIP(hipMemcpyAsync(&depth, d_depth, 4, hipMemcpyDeviceToHost, st));
        BVHT_HIP(hipStreamSynchronize(st));
    }
    if (depth == 0 || depth > b->nodes_used)
        return set_error(BVHT_EDEVICE, "BLAS topology: depth %u out of range", depth);
    b->depth = depth;
    uint32_t* hist = d_depth + 1;
    uint32_t* big = hist + depth;  // big-leaf count, wide-leaf flag
    rc = fill_u32(ctx, hist, 0, depth + 2);
    if (rc)
        return rc;
    uint32_t nb_levels = std::min<uint32_t>(nb, 512);
    hipLaunchKernelGGL(k_level_hist, dim3(nb_levels), dim3(256), 0, st, b->nodes_used, depth, b->d_level, hist);
    // big leaves: counted here, listed below; and whether every node fits a
    // packed stack entry
    hipLaunchKernelGGL(k_big_leaves, dim3(nb), dim3(256), 0, st, b->d_nodes, b->nodes_used, big, nullptr, big + 1);
    BVHT_HIP(hipGetLastError());
    std::vector<uint32_t> h(depth + 2);
    BVHT_HIP(hipMemcpyAsync(h.data(), hist, sizeof(uint32_t) * (depth + 2), hipMemcpyDeviceToHost, st));
    BVHT_HIP(hipStreamSynchronize(st));
    b->level_off.assign(depth + 1, 0);
    for (uint32_t L = 0; L < depth; ++L)
        b->level_off[L + 1] = b->level_off[L] + h[L];
    if (b->level_off[depth] + 1 != b->nodes_used && !(b->nodes_used == 1))
        return set_error(BVHT_EDEVICE, "BLAS topology: %u nodes grouped, %u expected", b->level_off[depth],
                         b->nodes_used - 1);
    BVHT_HIP(hipMemcpyAsync(hist, b->level_off.data(), sizeof(uint32_t) * depth, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_level_scatter, dim3(nb_levels), dim3(256), 0, st, b->nodes_used, depth, b->d_level, hist,
                       b->d_level_order);
    BVHT_HIP(hipGetLastError());
    uint32_t nbig = h[depth];
    b->packed_stack = h[depth + 1] == 0 && b->nodes_used <= kPackMask + 1u && b->n_tris <= kPackMask + 1u;
    (void)hipFree(b->d_big_leaves);
    b->d_big_leaves = nullptr;
    b->n_big_leaves = nbig;
    if (nbig) {
        if (hipMalloc(&b->d_big_leaves, sizeof(uint32_t) * nbig) != hipSuccess)
            return set_error(BVHT_ENOMEM, "device allocation of the big-leaf list failed");
        rc = fill_u32(ctx, big, 0, 1);
        if (rc)
            return rc;
        hipLaunchKernelGGL(k_big_leaves, dim3(nb), dim3(256), 0, st, b->d_nodes, b->nodes_used, big,
                           b->d_big_leaves, nullptr);
        BVHT_HIP(hipGetLastError());
    }
    BVHT_HIP(hipStreamSynchronize(st));
    return BVHT_OK;
}

// Bvh::refit, deepest level first: a level of more than kRunWG nodes gets a
// launch of its own; runs of consecutive smaller levels (the top of the tree
// and sparse deep levels) share one single-workgroup launch.
int32_t launch_refit(bvht_ctx ctx, bvht_blas blas)
{
    if (blas->n_big_leaves)
        hipLaunchKernelGGL(k_refit_big_leaves, dim3(blocks_for(blas->n_big_leaves, 4)), dim3(256), 0, ctx->stream,
                           blas->d_nodes, blas->d_tris, blas->d_big_leaves, blas->n_big_leaves);
    LevelRun run{};
    auto flush = [&]() {
        if (run.levels)
            hipLaunchKernelGGL(k_refit_levels, dim3(1), dim3(kRunWG), 0, ctx->stream, blas->d_nodes, blas->d_tris,
                               blas->d_level_order, run);
        run.levels = 0;
    };
    for (size_t L = blas->depth; L-- > 0;) {
        uint32_t off = blas->level_off[L], cnt = blas->level_off[L + 1] - off;
        if (cnt == 0)
            continue;
        if (cnt > kRunWG) {
            flush();
            hipLaunchKernelGGL(k_refit_level, dim3(blocks_for(cnt, 256)), dim3(256), 0, ctx->stream, blas->d_nodes,
                               blas->d_tris, blas->d_level_order + off, cnt);
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

// Host-side structural validation of a caller-supplied node array, so that
// no kernel ever follows an out-of-range index: children must lie after
// their parent (which also rules out cycles) and inside nodes_used, and leaf
// ranges inside the triangle array.
static int32_t validate_nodes(const bvht_node* nodes, uint32_t node_count, uint32_t nodes_used, uint32_t n)
{
    if (nodes_used > node_count || nodes_used == 0)
        return set_error(BVHT_EINVAL, "nodes_used (%u) must be in [1, node_count=%u]", nodes_used, node_count);
    if (node_count < 2)
        return set_error(BVHT_EINVAL, "node array needs at least 2 entries (root + dummy)");
    for (uint32_t k = 0; k < nodes_used; ++k) {
        if (k == 1)
            continue;
        const bvht_node& nd = nodes[k];
        if (nd.count > 0) {
            if ((uint64_t)nd.left_first + nd.count > n)
                return set_error(BVHT_EINVAL, "leaf %u range [%u, +%u) exceeds %u triangles", k, nd.left_first,
                                 nd.count, n);
        } else {
            if (nd.left_first <= k || nd.left_first == 1 || (uint64_t)nd.left_first + 1 >= nodes_used)
                return set_error(BVHT_EINVAL, "branch %u has invalid children %u,%u (nodes_used %u)", k,
                                 nd.left_first, nd.left_first + 1, nodes_used);
        }
    }
    return BVHT_OK;
}

constexpr size_t kBlasPoolBytes = size_t(1) << 30;
constexpr size_t kBlasPoolEntries = 32;

void* blas_buffer(bvht_ctx ctx, size_t bytes)
{
    size_t best = ctx->blas_pool.size();
    for (size_t i = 0; i < ctx->blas_pool.size(); ++i) {
        size_t s = ctx->blas_pool[i].second;
        if (s >= bytes && s <= bytes + bytes / 4 && (best == ctx->blas_pool.size() || s < ctx->blas_pool[best].second))
            best = i;
    }
    if (best < ctx->blas_pool.size()) {
        void* p = ctx->blas_pool[best].first;
        ctx->blas_pool_bytes -= ctx->blas_pool[best].second;
        ctx->blas_pool.erase(ctx->blas_pool.begin() + best);
        return p;
    }
    void* p = nullptr;
    return hipMalloc(&p, bytes) == hipSuccess ? p : nullptr;
}

void blas_buffer_release(bvht_ctx ctx, void* p, size_t bytes)
{
    if (!p)
        return;
    if (ctx && ctx->blas_pool.size() < kBlasPoolEntries && ctx->blas_pool_bytes + bytes <= kBlasPoolBytes) {
        ctx->blas_pool.emplace_back(p, bytes);
        ctx->blas_pool_bytes += bytes;
        return;
    }
    (void)hipFree(p);
}

int32_t blas_alloc(bvht_ctx ctx, uint32_t n, uint32_t node_count, bvht_blas* out)
{
    bvht_blas b = new (std::nothrow) bvht_blas_s();
    if (!b)
        return set_error(BVHT_ENOMEM, "host allocation failed");
    b->ctx = ctx;
    b->n_tris = n;
    b->node_count = node_count;
    b->d_nodes = static_cast<bvht_node*>(blas_buffer(ctx, sizeof(bvht_node) * (size_t)node_count));
    b->d_tris = static_cast<float*>(blas_buffer(ctx, sizeof(float) * 9 * (size_t)n));
    b->d_parent = static_cast<uint32_t*>(blas_buffer(ctx, sizeof(uint32_t) * (size_t)node_count));
    b->d_level = static_cast<uint32_t*>(blas_buffer(ctx, sizeof(uint32_t) * (size_t)node_count));
    b->d_level_order = static_cast<uint32_t*>(blas_buffer(ctx, sizeof(uint32_t) * (size_t)node_count));
    bool ok = b->d_nodes && b->d_tris && b->d_parent && b->d_level && b->d_level_order;
    if (!ok) {
        bvht_blas_destroy(b);
        return set_error(BVHT_ENOMEM, "device allocation for a BLAS of %u triangles failed", n);
    }
    *out = b;
    return BVHT_OK;
}

// Trace configuration for a tree needing `need` stack entries: the context's
// configuration if set (autotuned), else 64-lane workgroups with the smallest
// LDS depth that holds the stack, at most 24 entries (deeper entries spill).
// Small LDS stacks keep 8-10 waves per SIMD resident, which matters more than
// avoiding the occasional spill (sweep on the 1.3M-triangle sphere: 64x24
// 1.155 ms, 64x16 1.163 ms, 256x32 1.473 ms for both passes of a frame).
void trace_config(bvht_ctx ctx, uint32_t need, int* wg, int* lds)
{
    if (ctx->trace_wg > 0) {
        *wg = ctx->trace_wg;
        *lds = ctx->trace_lds;
        return;
    }
    *wg = 64;
    *lds = need <= 8 ? 8 : (need <= 16 ? 16 : 24);
}

int32_t launch_trace_blas(bvht_ctx ctx, bvht_blas blas, const bvht_ray* d_rays, uint32_t n,
                          const uint32_t* d_n, bvht_hit* d_hits, bvht_stats* d_stats)
{
    if (n == 0)
        return BVHT_OK;
    // stack entries needed: at most depth - 1
    uint32_t need = stack_cap(ctx, blas->depth > 0 ? blas->depth - 1 : 0);
    int wg, lds;
    trace_config(ctx, need, &wg, &lds);
    uint32_t* spill = nullptr;
    if (need > (uint32_t)lds) {
        int32_t rc = ensure_spill(ctx, &ctx->spill, &ctx->spill_words, (size_t)(need - lds) * n);
        if (rc)
            return rc;
        spill = ctx->spill;
    }
    trace_blas_fn fn = pick_trace_blas(wg, lds, d_stats != nullptr, blas->packed_stack);
    if (!fn)
        return set_error(BVHT_EINVAL, "no trace kernel for workgroup %d, LDS depth %d", wg, lds);
    TraceOrder o;
    int32_t rc = trace_order_begin(ctx, blas, d_rays, n, d_n == nullptr && d_stats == nullptr, &o);
    if (rc)
        return rc;
    hipLaunchKernelGGL(fn, dim3(blocks_for(n, wg)), dim3(wg), 0, ctx->stream,
                       reinterpret_cast<const float4*>(blas->d_nodes), blas->d_tris, d_rays, d_hits, n, d_n, d_stats,
                       spill, need, ctx->d_status, o.sched, o.sched_n, o.cost);
    BVHT_HIP(hipGetLastError());
    return trace_order_end(ctx, blas, d_rays, n, o);
}

// Per-256-ray-block exclusive offsets of the hits (in ctx scratch) and the
// total hit count in *d_count; the caller's write kernel packs the hits.
int32_t hit_block_offsets(bvht_ctx ctx, const bvht_hit* d_hits, uint32_t n, uint32_t prim_limit, uint32_t inst_limit,
                          const void* rows, const void* blases, uint32_t* d_count, uint32_t** offsets)
{
    uint32_t nb = blocks_for(n, 256);
    size_t count_bytes = align256(sizeof(uint32_t) * std::max<uint32_t>(nb, 1));
    int32_t rc = ensure_scratch(ctx, count_bytes + sizeof(uint32_t) * blocks_for(nb, kScanTile));
    if (rc)
        return rc;
    uint32_t* bc = static_cast<uint32_t*>(ctx->scratch);
    *offsets = bc;
    if (nb == 0)
        return fill_u32(ctx, d_count, 0, 1);
    hipLaunchKernelGGL(k_hit_counts, dim3(nb), dim3(256), 0, ctx->stream, d_hits, n, prim_limit, inst_limit,
                       static_cast<const ObjRows*>(rows), static_cast<const BlasRef*>(blases), bc);
    launch_scan(ctx, bc, nb, reinterpret_cast<uint32_t*>(static_cast<char*>(ctx->scratch) + count_bytes), d_count);
    BVHT_HIP(hipGetLastError());
    return BVHT_OK;
}

// Scratch layout of the camera pass: [staged rays (32 B) and keys (4 B) per
// pixel] (staged form only), per-(bin, wave) counts, scan tile sums.
// Returns the bytes needed.
static size_t camera_pass_layout(uint32_t n, bool staged, size_t* count_at, size_t* sums_at)
{
    size_t entries = (size_t)kDirBins * ((n + 63) / 64);
    *count_at = staged ? align256((sizeof(float4) * 2 + sizeof(uint32_t)) * (size_t)n) : 0;
    *sums_at = align256(*count_at + sizeof(uint32_t) * entries);
    return *sums_at + sizeof(uint32_t) * blocks_for(entries, kScanTile);
}

// Camera rays + trace + direction-grouped bounce rays on a scratch area of
// camera_pass_layout's size (n > 0): the fused kernel, the scan of the
// bin-major wave counts, then the scatter of the staged rays or the regroup.
static int32_t enqueue_camera_pass(bvht_ctx ctx, bvht_blas blas, const bvht_camera* cam, uint32_t width,
                                   uint32_t height, int32_t order, uint64_t seed, bvht_hit* d_hits, bvht_ray* d_out,
                                   uint32_t* d_src, uint32_t* d_count, int wg, int lds, bool staged, char* base,
                                   uint32_t* spill, uint32_t need, const uint32_t* sched, const uint32_t* sched_n,
                                   uint32_t* cost)
{
    camera_bounce_fn fn = pick_camera_bounce(wg, lds, blas->packed_stack, staged);
    if (!fn)
        return set_error(BVHT_EINVAL, "no trace kernel for workgroup %d, LDS depth %d", wg, lds);
    uint32_t n = width * height;
    uint32_t n_waves = (n + 63) / 64;
    size_t count_at, sums_at;
    camera_pass_layout(n, staged, &count_at, &sums_at);
    size_t entries = (size_t)kDirBins * n_waves;
    float4* stg = staged ? reinterpret_cast<float4*>(base) : nullptr;
    uint32_t* key = staged ? reinterpret_cast<uint32_t*>(base + sizeof(float4) * 2 * (size_t)n) : nullptr;
    uint32_t* counts = reinterpret_cast<uint32_t*>(base + count_at);
    uint32_t* sums = reinterpret_cast<uint32_t*>(base + sums_at);
    hipLaunchKernelGGL(fn, dim3(blocks_for(n, wg)), dim3(wg), 0, ctx->stream,
                       reinterpret_cast<const float4*>(blas->d_nodes), blas->d_tris, blas->n_tris, *cam, width, height,
                       order, seed, d_hits, stg, key, counts, n_waves, spill, need, ctx->d_status, sched, sched_n,
                       cost);
    launch_scan(ctx, counts, (uint32_t)entries, sums, d_count);
    if (staged)
        hipLaunchKernelGGL(k_bounce_scatter, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, stg, key, n,
                           n_waves, counts, d_out, d_src);
    else
        hipLaunchKernelGGL(k_camera_bounce_regroup, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream,
                           blas->d_tris, blas->n_tris, *cam, width, height, order, seed, d_hits, n_waves, counts,
                           d_out, d_src);
    BVHT_HIP(hipGetLastError());
    return BVHT_OK;
}

// The camera pass with its staging buffers in the context scratch.
int32_t launch_trace_camera_bounce(bvht_ctx ctx, bvht_blas blas, const bvht_camera* cam, uint32_t width,
                                   uint32_t height, int32_t order, uint64_t seed, bvht_hit* d_hits, bvht_ray* d_out,
                                   uint32_t* d_src, uint32_t* d_count)
{
    uint32_t n = width * height;
    if (n == 0)
        return fill_u32(ctx, d_count, 0, 1);
    uint32_t need = stack_cap(ctx, blas->depth > 0 ? blas->depth - 1 : 0);
    int wg, lds;
    trace_config(ctx, need, &wg, &lds);
    uint32_t* spill = nullptr;
    if (need > (uint32_t)lds) {
        int32_t rc = ensure_spill(ctx, &ctx->spill, &ctx->spill_words, (size_t)(need - lds) * n);
        if (rc)
            return rc;
        spill = ctx->spill;
    }
    size_t count_at, sums_at;
    int32_t rc = ensure_scratch(ctx, camera_pass_layout(n, true, &count_at, &sums_at));
    if (rc)
        return rc;
    // longest-first wave order from the last pass over the same tree, camera
    // and frame; every pass records its waves' durations for the next
    uint32_t groups = (n + 63) / 64;
    if (groups > ctx->cam_cap) {
        if (ctx->live_graphs > 0 || ctx->capturing)
            return refuse_move(ctx, "the camera-pass wave schedule", sizeof(uint32_t) * 2 * (size_t)groups);
        BVHT_HIP(hipStreamSynchronize(ctx->stream));
        (void)hipFree(ctx->cam_cost);
        (void)hipFree(ctx->cam_sched);
        ctx->cam_cost = ctx->cam_sched = nullptr;
        ctx->cam_cap = 0;
        ctx->cam_valid = false;
        if (hipMalloc(&ctx->cam_cost, sizeof(uint32_t) * groups) != hipSuccess ||
            hipMalloc(&ctx->cam_sched, sizeof(uint32_t) * groups) != hipSuccess ||
            (!ctx->cam_bins && hipMalloc(&ctx->cam_bins, sizeof(uint32_t) * kSchedBinsWords) != hipSuccess))
            return set_error(BVHT_ENOMEM, "camera-pass wave schedule for %u groups", groups);
        ctx->cam_cap = groups;
    }
    bool same = ctx->cam_valid && ctx->cam_blas == blas && ctx->cam_w == width && ctx->cam_h == height &&
                ctx->cam_order == order && std::memcmp(&ctx->cam_key, cam, sizeof *cam) == 0;
    rc = enqueue_camera_pass(ctx, blas, cam, width, height, order, seed, d_hits, d_out, d_src, d_count, wg, lds,
                             true, static_cast<char*>(ctx->scratch), spill, need, same ? ctx->cam_sched : nullptr,
                             sched_units(ctx->cam_bins), ctx->cam_cost);
    if (rc)
        return rc;
    // the order changes little between passes over the same frame: rebuilt
    // after the first pass and then every kCamResched passes (four small
    // launches each time)
    constexpr uint32_t kCamResched = 8;
    ctx->cam_passes = same ? ctx->cam_passes + 1 : 0;
    if (ctx->cam_passes % kCamResched == 0) {
        rc = launch_longest_first(ctx, ctx->cam_cost, groups, ctx->cam_bins, ctx->cam_sched);
        if (rc)
            return rc;
    }
    ctx->cam_valid = true;
    ctx->cam_blas = blas;
    ctx->cam_w = width;
    ctx->cam_h = height;
    ctx->cam_order = order;
    ctx->cam_key = *cam;
    return BVHT_OK;
}

// Bounce rays grouped by direction bin, ray order inside a bin: counts per
// (bin, block), one scan (bin-major), write.  rows/blases select the TLAS form.
int32_t launch_bounce_sorted(bvht_ctx ctx, const float* d_tris, uint32_t ntris, const void* rows, const void* blases,
                             uint32_t n_obj, const bvht_ray* d_rays, const bvht_hit* d_hits, uint32_t n,
                             uint64_t seed, bvht_ray* d_out, uint32_t* d_src, uint32_t* d_count)
{
    if (n == 0)
        return fill_u32(ctx, d_count, 0, 1);
    uint32_t nb = blocks_for(n, 256);
    size_t entries = (size_t)kDirBins * nb;
    size_t sums_at = align256(sizeof(uint32_t) * entries);
    int32_t rc = ensure_scratch(ctx, sums_at + sizeof(uint32_t) * blocks_for(entries, kScanTile));
    if (rc)
        return rc;
    uint32_t* counts = static_cast<uint32_t*>(ctx->scratch);
    uint32_t* sums = reinterpret_cast<uint32_t*>(static_cast<char*>(ctx->scratch) + sums_at);
    BounceSrc src{d_tris, ntris, static_cast<const ObjRows*>(rows), static_cast<const BlasRef*>(blases), n_obj};
    hipLaunchKernelGGL(k_bounce_bin_counts, dim3(nb), dim3(256), 0, ctx->stream, src, d_rays, d_hits, n, seed, nb,
                       counts);
    launch_scan(ctx, counts, (uint32_t)entries, sums, d_count);
    hipLaunchKernelGGL(k_bounce_bin_write, dim3(nb), dim3(256), 0, ctx->stream, src, d_rays, d_hits, n, seed, nb,
                       counts, d_out, d_src);
    BVHT_HIP(hipGetLastError());
    return BVHT_OK;
}

int32_t launch_bounce(bvht_ctx ctx, const float* d_tris, uint32_t ntris, const bvht_ray* d_rays,
                      const bvht_hit* d_hits, uint32_t n, uint64_t seed, bvht_ray* d_out, uint32_t* d_count)
{
    uint32_t* bc = nullptr;
    int32_t rc = hit_block_offsets(ctx, d_hits, n, ntris, 0xFFFFFFFFu, nullptr, nullptr, d_count, &bc);
    if (rc || n == 0)
        return rc;
    uint32_t nb = blocks_for(n, 256);
    hipLaunchKernelGGL(k_bounce_write, dim3(nb), dim3(256), 0, ctx->stream, d_tris, ntris, d_rays, d_hits, n, bc, seed,
                       d_out);
    BVHT_HIP(hipGetLastError());
    return BVHT_OK;
}

}  // namespace bvht

// ------------------------------------------------------------- C ABI ----

extern "C" {

const char* bvht_last_error(void) { return bvht::g_error.c_str(); }

int32_t bvht_create(int32_t device, bvht_ctx* out)
{
    BVHT_REQUIRE(out, "out is NULL");
    int count = 0;
    BVHT_HIP(hipGetDeviceCount(&count));
    BVHT_REQUIRE(device >= 0 && device < count, "device %d out of range (%d devices)", device, count);
    BVHT_HIP(hipSetDevice(device));
    bvht_ctx c = new (std::nothrow) bvht_ctx_s();
    if (!c)
        return set_error(BVHT_ENOMEM, "host allocation failed");
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return set_error(BVHT_EDEVICE, "hipStreamCreate failed");
    }
    if (hipHostMalloc(reinterpret_cast<void**>(&c->h_status), 64, hipHostMallocMapped) != hipSuccess ||
        hipHostGetDevicePointer(reinterpret_cast<void**>(&c->d_status), c->h_status, 0) != hipSuccess) {
        if (c->h_status)
            (void)hipHostFree(c->h_status);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return set_error(BVHT_ENOMEM, "pinned status word allocation failed");
    }
    *c->h_status = 0;
    int32_t rc = bvht::noalloc_scratch(device, &c->noalloc_scratch);
    if (rc) {
        (void)hipHostFree(c->h_status);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return rc;
    }
    *out = c;
    return BVHT_OK;
}

int32_t bvht_destroy(bvht_ctx ctx)
{
    if (!ctx)
        return BVHT_OK;
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->scratch)
        (void)hipFree(ctx->scratch);
    if (ctx->spill)
        (void)hipFree(ctx->spill);
    (void)hipFree(ctx->cam_cost);
    (void)hipFree(ctx->cam_sched);
    (void)hipFree(ctx->cam_bins);
    (void)hipFree(ctx->tls_cost);
    (void)hipFree(ctx->tls_sched);
    (void)hipFree(ctx->tls_bins);
    if (ctx->h_status)
        (void)hipHostFree(ctx->h_status);
    if (ctx->h_stage)
        (void)hipHostFree(ctx->h_stage);
    for (auto& e : ctx->blas_pool)
        (void)hipFree(e.first);
    ctx->blas_pool.clear();
    for (hipEvent_t e : ctx->stage_ev)
        if (e)
            (void)hipEventDestroy(e);
    if (ctx->copy_ev)
        (void)hipEventDestroy(ctx->copy_ev);
    bvht::stop_copy_pool(ctx->copy_pool);
    release_build_ws(ctx);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return BVHT_OK;
}

int32_t bvht_sync(bvht_ctx ctx)
{
    BVHT_REQUIRE(ctx, "ctx is NULL");
    BVHT_HIP(hipStreamSynchronize(ctx->stream));
    return take_status(ctx);
}

int32_t bvht_set_stack_limit(bvht_ctx ctx, uint32_t max_entries)
{
    BVHT_REQUIRE(ctx, "ctx is NULL");
    ctx->stack_limit = max_entries;
    return BVHT_OK;
}

int32_t bvht_set_wave_schedule(bvht_ctx ctx, int32_t enable)
{
    BVHT_REQUIRE(ctx, "ctx is NULL");
    ctx->tls_on = enable != 0;
    ctx->tls_valid = false;  // a new order starts from the next call's durations
    return BVHT_OK;
}

void* bvht_stream(bvht_ctx ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int32_t bvht_device_alloc(bvht_ctx ctx, size_t bytes, void** out)
{
    BVHT_REQUIRE(ctx && out, "NULL argument");
    if (hipMalloc(out, bytes ? bytes : 1) != hipSuccess)
        return set_error(BVHT_ENOMEM, "hipMalloc(%zu) failed", bytes);
    return BVHT_OK;
}

int32_t bvht_device_free(bvht_ctx ctx, void* ptr)
{
    BVHT_REQUIRE(ctx, "ctx is NULL");
    if (ptr) {
        BVHT_HIP(hipStreamSynchronize(ctx->stream));
        BVHT_HIP(hipFree(ptr));
    }
    return BVHT_OK;
}

int32_t bvht_copy_to_device(bvht_ctx ctx, void* dst, const void* src, size_t bytes)
{
    BVHT_REQUIRE(ctx && (bytes == 0 || (dst && src)), "NULL argument");
    int32_t rc = bvht::host_to_device(ctx, dst, src, bytes);
    if (rc)
        return rc;
    BVHT_HIP(hipStreamSynchronize(ctx->stream));
    return BVHT_OK;
}

int32_t bvht_copy_to_host(bvht_ctx ctx, void* dst, const void* src, size_t bytes)
{
    BVHT_REQUIRE(ctx && (bytes == 0 || (dst && src)), "NULL argument");
    int32_t rc = bvht::device_to_host(ctx, dst, src, bytes);
    if (rc)
        return rc;
    return take_status(ctx);
}

int32_t bvht_event_create(bvht_ctx ctx, void** out)
{
    BVHT_REQUIRE(ctx && out, "NULL argument");
    hipEvent_t ev;
    BVHT_HIP(hipEventCreate(&ev));
    *out = (void*)ev;
    return BVHT_OK;
}

int32_t bvht_event_record(bvht_ctx ctx, void* ev)
{
    BVHT_REQUIRE(ctx && ev, "NULL argument");
    BVHT_HIP(hipEventRecord((hipEvent_t)ev, ctx->stream));
    return BVHT_OK;
}

int32_t bvht_event_elapsed(void* start, void* end, float* ms)
{
    BVHT_REQUIRE(start && end && ms, "NULL argument");
    BVHT_HIP(hipEventSynchronize((hipEvent_t)end));
    BVHT_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)end));
    return BVHT_OK;
}

int32_t bvht_event_destroy(void* ev)
{
    if (ev)
        BVHT_HIP(hipEventDestroy((hipEvent_t)ev));
    return BVHT_OK;
}

int32_t bvht_stream_copy(bvht_ctx ctx, void* d_dst, const void* d_src, size_t bytes)
{
    BVHT_REQUIRE(ctx && d_dst && d_src, "NULL argument");
    BVHT_REQUIRE(bytes % 16 == 0 && ((uintptr_t)d_dst % 16) == 0 && ((uintptr_t)d_src % 16) == 0,
                 "bvht_stream_copy needs 16-byte aligned pointers and a multiple of 16 bytes");
    size_t n = bytes / 16;
    if (n == 0)
        return BVHT_OK;
    uint32_t blocks = (uint32_t)std::min<size_t>((n + 255) / 256, 256 * 32);
    hipLaunchKernelGGL(k_stream_copy, dim3(blocks), dim3(256), 0, ctx->stream, static_cast<float4*>(d_dst),
                       static_cast<const float4*>(d_src), n);
    BVHT_HIP(hipGetLastError());
    return BVHT_OK;
}

int32_t bvht_blas_upload(bvht_ctx ctx, const bvht_node* nodes, uint32_t node_count, uint32_t nodes_used,
                         const float* tris, uint32_t n, bvht_blas* out)
{
    BVHT_REQUIRE(ctx && nodes && tris && out, "NULL argument");
    BVHT_REQUIRE(n > 0, "empty triangle array (the reference panics on an empty mesh)");
    int32_t rc = validate_nodes(nodes, node_count, nodes_used, n);
    if (rc)
        return rc;
    bvht_blas b = nullptr;
    rc = blas_alloc(ctx, n, node_count, &b);
    if (rc)
        return rc;
    b->nodes_used = nodes_used;
    rc = host_to_device(ctx, b->d_nodes, nodes, sizeof(bvht_node) * (size_t)node_count);
    if (!rc)
        rc = host_to_device(ctx, b->d_tris, tris, sizeof(float) * 9 * (size_t)n);
    if (rc) {
        bvht_blas_destroy(b);
        return rc;
    }
    rc = compute_blas_topology(b, nodes, 0);
    if (rc) {
        bvht_blas_destroy(b);
        return rc;
    }
    *out = b;
    return BVHT_OK;
}

int32_t bvht_blas_download(bvht_ctx ctx, bvht_blas blas, bvht_node* nodes, float* tris)
{
    BVHT_REQUIRE(ctx && blas, "NULL argument");
    int32_t rc = BVHT_OK;
    if (nodes)
        rc = bvht::device_to_host(ctx, nodes, blas->d_nodes, sizeof(bvht_node) * (size_t)blas->node_count);
    if (!rc && tris)
        rc = bvht::device_to_host(ctx, tris, blas->d_tris, sizeof(float) * 9 * (size_t)blas->n_tris);
    if (rc)
        return rc;
    BVHT_HIP(hipStreamSynchronize(ctx->stream));
    return BVHT_OK;
}

int32_t bvht_blas_info(bvht_blas blas, uint32_t* tri_count, uint32_t* node_count, uint32_t* nodes_used,
                       uint32_t* depth)
{
    BVHT_REQUIRE(blas, "blas is NULL");
    if (tri_count) *tri_count = blas->n_tris;
    if (node_count) *node_count = blas->node_count;
    if (nodes_used) *nodes_used = blas->nodes_used;
    if (depth) *depth = blas->depth;
    return BVHT_OK;
}

int32_t bvht_blas_set_vertices(bvht_ctx ctx, bvht_blas blas, const float* tris, uint32_t n)
{
    BVHT_REQUIRE(ctx && blas && tris, "NULL argument");
    BVHT_REQUIRE(n == blas->n_tris, "triangle count %u differs from the BLAS's %u", n, blas->n_tris);
    int32_t rc = bvht::host_to_device(ctx, blas->d_tris, tris, sizeof(float) * 9 * (size_t)n);
    if (rc)
        return rc;
    BVHT_HIP(hipStreamSynchronize(ctx->stream));
    return BVHT_OK;
}

int32_t bvht_blas_set_vertices_device(bvht_ctx ctx, bvht_blas blas, const float* d_tris, uint32_t n)
{
    BVHT_REQUIRE(ctx && blas && d_tris, "NULL argument");
    BVHT_REQUIRE(n == blas->n_tris, "triangle count %u differs from the BLAS's %u", n, blas->n_tris);
    BVHT_HIP(hipMemcpyAsync(blas->d_tris, d_tris, sizeof(float) * 9 * (size_t)n, hipMemcpyDeviceToDevice,
                            ctx->stream));
    return BVHT_OK;
}

int32_t bvht_blas_refit(bvht_ctx ctx, bvht_blas blas)
{
    BVHT_REQUIRE(ctx && blas, "NULL argument");
    return bvht::launch_refit(ctx, blas);
}

int32_t bvht_blas_bounds(bvht_ctx ctx, bvht_blas blas, float* out6)
{
    BVHT_REQUIRE(ctx && blas && out6, "NULL argument");
    bvht_node root;
    BVHT_HIP(hipMemcpyAsync(&root, blas->d_nodes, sizeof root, hipMemcpyDeviceToHost, ctx->stream));
    BVHT_HIP(hipStreamSynchronize(ctx->stream));
    std::memcpy(out6, root.bmin, 12);
    std::memcpy(out6 + 3, root.bmax, 12);
    return BVHT_OK;
}

int32_t bvht_blas_destroy(bvht_blas b)
{
    if (!b)
        return BVHT_OK;
    // traces of this tree may be queued on any context's stream; a buffer
    // goes back to the pool only once the device is idle (hipFree waits the
    // same way)
    (void)hipDeviceSynchronize();
    bvht_ctx c = b->ctx;
    size_t nc = b->node_count, nt = b->n_tris;
    blas_buffer_release(c, b->d_nodes, sizeof(bvht_node) * nc);
    blas_buffer_release(c, b->d_tris, sizeof(float) * 9 * nt);
    blas_buffer_release(c, b->d_perm, sizeof(uint32_t) * nt);
    blas_buffer_release(c, b->d_parent, sizeof(uint32_t) * nc);
    blas_buffer_release(c, b->d_level, sizeof(uint32_t) * nc);
    blas_buffer_release(c, b->d_level_order, sizeof(uint32_t) * nc);
    (void)hipFree(b->d_big_leaves);
    (void)hipFree(b->d_attr_normals);
    (void)hipFree(b->d_attr_tex);
    (void)hipFree(b->d_texture);
    (void)hipFree(b->d_attr_ref);
    delete b;
    return BVHT_OK;
}

int32_t bvht_trace_blas(bvht_ctx ctx, bvht_blas blas, const bvht_ray* d_rays, uint32_t n, bvht_hit* d_hits,
                        bvht_stats* d_stats)
{
    BVHT_REQUIRE(ctx && blas, "NULL argument");
    BVHT_REQUIRE(n == 0 || (d_rays && d_hits), "NULL ray or hit buffer");
    return bvht::launch_trace_blas(ctx, blas, d_rays, n, nullptr, d_hits, d_stats);
}

int32_t bvht_trace_blas_count(bvht_ctx ctx, bvht_blas blas, const bvht_ray* d_rays, const uint32_t* d_count,
                              uint32_t max_n, bvht_hit* d_hits, bvht_stats* d_stats)
{
    BVHT_REQUIRE(ctx && blas && d_count, "NULL argument");
    BVHT_REQUIRE(max_n == 0 || (d_rays && d_hits), "NULL ray or hit buffer");
    return bvht::launch_trace_blas(ctx, blas, d_rays, max_n, d_count, d_hits, d_stats);
}

int32_t bvht_set_trace_config(bvht_ctx ctx, int32_t wg, int32_t lds_depth)
{
    BVHT_REQUIRE(ctx, "ctx is NULL");
    if (wg == 0) {
        ctx->trace_wg = 0;
        ctx->trace_lds = 0;
        return BVHT_OK;
    }
    BVHT_REQUIRE(pick_trace_blas(wg, lds_depth, false, false) != nullptr,
                 "unsupported trace configuration: workgroup %d, LDS depth %d", wg, lds_depth);
    ctx->trace_wg = wg;
    ctx->trace_lds = lds_depth;
    return BVHT_OK;
}

int32_t bvht_get_trace_config(bvht_ctx ctx, int32_t* wg, int32_t* lds_depth)
{
    BVHT_REQUIRE(ctx && wg && lds_depth, "NULL argument");
    *wg = ctx->trace_wg;
    *lds_depth = ctx->trace_lds;
    return BVHT_OK;
}

int32_t bvht_raygen(bvht_ctx ctx, const bvht_camera* cam, uint32_t width, uint32_t height, int32_t order,
                    bvht_ray* d_rays)
{
    BVHT_REQUIRE(ctx && cam && d_rays, "NULL argument");
    BVHT_REQUIRE(width > 0 && height > 0 && (uint64_t)width * height < (1ull << 32), "bad image size %ux%u",
                 width, height);
    BVHT_REQUIRE(order == BVHT_ORDER_ROWS || (order == BVHT_ORDER_TILES && width % 8 == 0 && height % 8 == 0),
                 "order %d invalid for %ux%u (tile order needs multiples of 8)", order, width, height);
    uint32_t n = width * height;
    hipLaunchKernelGGL(k_raygen, dim3(bvht::blocks_for(n, 256)), dim3(256), 0, ctx->stream, *cam, width, height,
                       order, d_rays);
    BVHT_HIP(hipGetLastError());
    return BVHT_OK;
}

int32_t bvht_bounce(bvht_ctx ctx, bvht_blas blas, const bvht_ray* d_rays, const bvht_hit* d_hits, uint32_t n,
                    uint64_t seed, bvht_ray* d_out, uint32_t* d_count)
{
    BVHT_REQUIRE(ctx && blas && d_count, "NULL argument");
    BVHT_REQUIRE(n == 0 || (d_rays && d_hits && d_out), "NULL buffer");
    return bvht::launch_bounce(ctx, blas->d_tris, blas->n_tris, d_rays, d_hits, n, seed, d_out, d_count);
}

int32_t bvht_bounce_sorted(bvht_ctx ctx, bvht_blas blas, const bvht_ray* d_rays, const bvht_hit* d_hits, uint32_t n,
                           uint64_t seed, bvht_ray* d_out, uint32_t* d_src, uint32_t* d_count)
{
    BVHT_REQUIRE(ctx && blas && d_count, "NULL argument");
    BVHT_REQUIRE(n == 0 || (d_rays && d_hits && d_out), "NULL buffer");
    return bvht::launch_bounce_sorted(ctx, blas->d_tris, blas->n_tris, nullptr, nullptr, 0, d_rays, d_hits, n, seed,
                                      d_out, d_src, d_count);
}

int32_t bvht_trace_camera_bounce(bvht_ctx ctx, bvht_blas blas, const bvht_camera* cam, uint32_t width,
                                 uint32_t height, int32_t order, uint64_t seed, bvht_hit* d_hits, bvht_ray* d_out,
                                 uint32_t* d_src, uint32_t* d_count)
{
    BVHT_REQUIRE(ctx && blas && cam && d_count, "NULL argument");
    BVHT_REQUIRE(order == BVHT_ORDER_ROWS || (order == BVHT_ORDER_TILES && width % 8 == 0 && height % 8 == 0),
                 "order %d invalid for %ux%u (tile order needs multiples of 8)", order, width, height);
    BVHT_REQUIRE((uint64_t)width * height < (1ull << 31), "frame %ux%u too large", width, height);
    BVHT_REQUIRE(width * height == 0 || (d_hits && d_out), "NULL buffer");
    return bvht::launch_trace_camera_bounce(ctx, blas, cam, width, height, order, seed, d_hits, d_out, d_src, d_count);
}

int32_t bvht_trace_camera_bounce_noalloc(bvht_ctx ctx, bvht_blas blas, const bvht_camera* cam, uint32_t width,
                                         uint32_t height, int32_t order, uint64_t seed, bvht_hit* d_hits,
                                         bvht_ray* d_out, uint32_t* d_src, uint32_t* d_count)
{
    BVHT_REQUIRE(ctx && blas && cam && d_count, "NULL argument");
    BVHT_REQUIRE(order == BVHT_ORDER_ROWS || (order == BVHT_ORDER_TILES && width % 8 == 0 && height % 8 == 0),
                 "order %d invalid for %ux%u (tile order needs multiples of 8)", order, width, height);
    BVHT_REQUIRE((uint64_t)width * height <= kNoAllocMaxPixels,
                 "frame %ux%u exceeds the %u pixels of the allocation-free camera pass's fixed scratch", width,
                 height, kNoAllocMaxPixels);
    uint32_t n = width * height;
    BVHT_REQUIRE(n == 0 || (d_hits && d_out), "NULL buffer");
    uint32_t need = stack_cap(ctx, blas->depth > 0 ? blas->depth - 1 : 0);
    BVHT_REQUIRE(need <= 32,
                 "a tree of depth %u needs %u traversal stack entries; the allocation-free camera pass keeps at most "
                 "32 per ray, in LDS (bvht_set_stack_limit can cap it)", blas->depth, need);
    BVHT_REQUIRE(ctx->noalloc_scratch, "the fixed camera-pass scratch is not set up on this context");
    if (n == 0)
        return bvht::fill_u32(ctx, d_count, 0, 1);
    // the context's configuration if its LDS stack holds the tree, else 64
    // lanes with the smallest LDS stack that does: nothing spills
    int wg = 64, lds = need <= 8 ? 8 : (need <= 16 ? 16 : (need <= 24 ? 24 : 32));
    if (ctx->trace_wg > 0 && (uint32_t)ctx->trace_lds >= need) {
        wg = ctx->trace_wg;
        lds = ctx->trace_lds;
    }
    return bvht::enqueue_camera_pass(ctx, blas, cam, width, height, order, seed, d_hits, d_out, d_src, d_count, wg,
                                     lds, false, static_cast<char*>(ctx->noalloc_scratch), nullptr, need, nullptr,
                                     nullptr, nullptr);
}

int32_t bvht_depth(bvht_ctx ctx, const bvht_hit* d_hits, uint32_t n, float* d_depth)
{
    BVHT_REQUIRE(ctx, "ctx is NULL");
    if (n == 0)
        return BVHT_OK;
    BVHT_REQUIRE(d_hits && d_depth, "NULL buffer");
    hipLaunchKernelGGL(k_depth_acc, dim3(bvht::blocks_for(n, 256)), dim3(256), 0, ctx->stream, d_hits, n, d_depth);
    BVHT_HIP(hipGetLastError());
    return BVHT_OK;
}

}  // extern "C"
