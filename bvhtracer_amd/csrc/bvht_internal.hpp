// bvht_internal.hpp -- host-side internals shared by the library's .hip files.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/bvht.h"

namespace bvht {
struct CopyPool;
}

struct bvht_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
    // Reusable device scratch (grown on demand; freed with the context).
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    // Traversal stack spill of this context's trace launches (grown on
    // demand).  Per context, not per tree: a tree may be traced from several
    // contexts' streams at once.
    uint32_t* spill = nullptr;
    size_t spill_words = 0;
    // trace kernel configuration (0 = choose from the tree depth)
    int trace_wg = 0;
    int trace_lds = 0;
    // Upper bound on a trace's stack entries (0: the tree's depth bound).
    uint32_t stack_limit = 0;
    // Status word the traversal kernels set when a push is dropped at the
    // stack bound (kStatusStack).  Pinned host memory mapped into the
    // device, so a fence (bvht_sync, a read-back) reads it without a copy.
    uint32_t* h_status = nullptr;
    uint32_t* d_status = nullptr;
    // hipGraphs captured on this context and not yet destroyed, and whether
    // a capture is open: while either is non-zero, scratch and spill buffers
    // must keep their addresses (a graph holds them as kernel arguments).
    int live_graphs = 0;
    bool capturing = false;
    // Camera-pass wave schedule (bvht_trace_camera_bounce): the durations
    // of the last pass's 64-ray groups, and the longest-first order made
    // from them, used by the next pass with the same tree, camera and frame.
    uint32_t* cam_cost = nullptr;
    uint32_t* cam_sched = nullptr;
    uint32_t* cam_bins = nullptr;
    uint32_t cam_cap = 0;
    bool cam_valid = false;
    const void* cam_blas = nullptr;
    bvht_camera cam_key{};
    uint32_t cam_w = 0, cam_h = 0;
    int32_t cam_order = -1;
    uint32_t cam_passes = 0;  // passes since the key last changed
    // Device address, on this context's device, of the fixed scratch array
    // of bvht_trace_camera_bounce_noalloc (shared by every context there).
    void* noalloc_scratch = nullptr;
    // BLAS build workspace (bvht_build.hip), kept between builds.
    void* build_ws = nullptr;
    // Pinned staging for large copies between pageable host memory and the
    // device (two halves of stage_half bytes, allocated on first use), and
    // the event of the last transfer through each half.
    char* h_stage = nullptr;
    size_t stage_half = 0;
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    // Mark of the work a copy to the host beside later kernels waits for
    // (mark_for_copy / device_to_host_beside).
    hipEvent_t copy_ev = nullptr;
    // Helper threads of the staged copies (started with the staging, joined
    // by bvht_destroy).
    bvht::CopyPool* copy_pool = nullptr;
    // Device buffers of destroyed BLASes built or uploaded on this context,
    // kept for the next BLAS (a rebuild every frame allocates nothing), up
    // to kBlasPoolBytes in all; freed by bvht_destroy.
    std::vector<std::pair<void*, size_t>> blas_pool;
    size_t blas_pool_bytes = 0;
    // Longest-first wave order of bvht_trace_tlas (bvht_set_wave_schedule):
    // the last call's group durations and the order made from them, keyed by
    // TLAS, ray buffer and ray count.
    bool tls_on = false;
    uint32_t* tls_cost = nullptr;
    uint32_t* tls_sched = nullptr;
    uint32_t* tls_bins = nullptr;
    uint32_t tls_cap = 0;
    bool tls_valid = false;
    const void* tls_tlas = nullptr;
    const void* tls_rays = nullptr;
    uint32_t tls_n = 0;
    uint32_t tls_calls = 0;  // calls since the key last changed
};

// Bits of the context status word.
constexpr uint32_t kStatusStack = 1u;

// Per-model shading attributes as the attribute accumulators read them
// (device pointers; NULL where absent), triangles in build order.
struct AttrRef {
    const float* normals;     // 9 floats per triangle (three corner normals)
    const float* tex_coords;  // 6 floats per triangle
    const uint8_t* texture;   // RGB8, row-major, tex_w x tex_h
    uint32_t tex_w, tex_h;
    uint32_t ntris, pad;
};

struct bvht_blas_s {
    bvht_ctx ctx = nullptr;
    uint32_t n_tris = 0;
    uint32_t node_count = 0;  // length of the node array (2N for built BLASes)
    uint32_t nodes_used = 0;
    uint32_t depth = 0;       // number of levels (root alone = 1)
    bvht_node* d_nodes = nullptr;
    float* d_tris = nullptr;  // 9 floats per triangle, build order
    uint32_t* d_perm = nullptr;  // input index of the triangle at each build position (GPU builds only)
    uint32_t* d_parent = nullptr;       // parent of each used node (root: 0xFFFFFFFF)
    uint32_t* d_level = nullptr;        // depth of each used node (root 0; dummy node 1: 0xFFFFFFFF)
    uint32_t* d_level_order = nullptr;  // used nodes except the dummy, grouped by level
    std::vector<uint32_t> level_off;    // level L occupies [level_off[L], level_off[L+1]) of d_level_order
    uint32_t* d_big_leaves = nullptr;   // leaves with more than kBigLeaf triangles (refit a wave each)
    uint32_t n_big_leaves = 0;
    bool packed_stack = false;  // every node fits a packed stack entry (see bvht_trace.hpp)
    // shading attributes (bvht_blas_set_attributes)
    float* d_attr_normals = nullptr;
    float* d_attr_tex = nullptr;
    uint8_t* d_texture = nullptr;
    uint32_t tex_w = 0, tex_h = 0;
    uint64_t attr_version = 0;
    AttrRef* d_attr_ref = nullptr;      // this BLAS's AttrRef on the device
};

struct bvht_tlas_s {
    bvht_ctx ctx = nullptr;
    uint32_t n = 0;
    std::vector<bvht_blas> blases;
    void* d_blas_table = nullptr;   // per-BLAS {nodes, tris} device pointers
    bvht_instance* d_objects = nullptr;
    void* d_rows = nullptr;         // per-object 3x4 rows of M^-1 + BLAS id, as traced
    float* d_bounds = nullptr;      // n * 6 world bounds
    bvht_tnode* d_nodes = nullptr;  // 2n
    int32_t* d_list = nullptr;      // rebuild work list (n), reused as the depth-walk stack: 2 max(n, 4) entries
    uint32_t* d_meta = nullptr;     // {nodes_used, depth, status}
    uint32_t depth = 0;             // TLAS levels after the last rebuild
    uint32_t max_blas_depth = 0;
    AttrRef* d_attr_table = nullptr;        // per-BLAS attributes for the attribute accumulators
    std::vector<uint64_t> attr_versions;    // BLAS attribute versions the table was built from
    // physics binding (bvht_tlas_bind_world)
    bvht_world world = nullptr;
    uint32_t* d_object_body = nullptr;  // body of each object (BVHT_MISS: none)
    float* d_transform_init = nullptr;  // 16 per object, column-major
};

namespace bvht {

int32_t set_error(int32_t code, const char* fmt, ...);
int32_t ensure_scratch(bvht_ctx ctx, size_t bytes);
// Parents, levels (computed here unless known_depth > 0 says the build left
// them in d_level and the tree has that many levels), the level-grouped node
// order and the big-leaf list of a BLAS.
int32_t compute_blas_topology(bvht_blas b, const bvht_node* h_nodes, uint32_t known_depth);
// Copies between pageable host memory and the device on ctx's stream.
// host_to_device returns once src may be reused (the transfer completes in
// stream order); device_to_host returns once dst holds the data.  Large
// copies go through ctx's pinned staging, chunk by chunk, with the host's
// share of each chunk split over a few threads and overlapped with the DMA
// of the next chunk.
int32_t host_to_device(bvht_ctx ctx, void* dst, const void* src, size_t bytes);
// A device buffer of at least `bytes` for a BLAS: a pooled one of this
// context (at most a quarter larger) or a new hipMalloc; nullptr if neither.
void* blas_buffer(bvht_ctx ctx, size_t bytes);
// Back to the context's pool (or freed when the pool is full).  The caller
// has made sure no queued work on any stream still uses it.
void blas_buffer_release(bvht_ctx ctx, void* p, size_t bytes);
int32_t device_to_host(bvht_ctx ctx, void* dst, const void* src, size_t bytes);
// mark_for_copy marks the work queued on ctx's stream so far;
// device_to_host_beside then copies on the device's copy stream (one per
// device for the process) after that work only, so kernels queued on ctx's
// stream after the mark run during the copy (copies below the staging
// threshold run on ctx's stream after all queued work instead).  It returns
// once dst holds the data.
int32_t mark_for_copy(bvht_ctx ctx);
int32_t device_to_host_beside(bvht_ctx ctx, void* dst, const void* src, size_t bytes);
int32_t ensure_spill(bvht_ctx ctx, uint32_t** buf, size_t* words, size_t need);
// After a fence on ctx's stream: BVHT_ESTACK (and the status cleared) if a
// trace since the last check dropped a stack entry, else BVHT_OK.
int32_t take_status(bvht_ctx ctx);
// Stack entries a trace over a tree needing `need` may use: `need`, or the
// context's stack limit when one is set and smaller.
inline uint32_t stack_cap(bvht_ctx ctx, uint32_t need)
{
    return (ctx->stack_limit && ctx->stack_limit < need) ? ctx->stack_limit : need;
}
int32_t blas_alloc(bvht_ctx ctx, uint32_t n, uint32_t node_count, bvht_blas* out);
// Frees the context's BLAS build workspace (bvht_destroy).
void release_build_ws(bvht_ctx ctx);
int32_t fill_u32(bvht_ctx ctx, uint32_t* p, uint32_t v, size_t n);
// In-place exclusive scan of a[0, m) (m > 0, a 16-byte aligned) on ctx's
// stream; sums holds scan_sums_entries(m) entries; the total goes to *total
// unless it is NULL.
void launch_scan(bvht_ctx ctx, uint32_t* a, uint32_t m, uint32_t* sums, uint32_t* total);
uint32_t scan_sums_entries(uint32_t m);
// Longest first: d_sched[0, n) = the units ordered by descending d_cost (a
// permutation of 0..n-1); d_bins holds kSchedBinsWords words: 2 * kSchedBins
// counters, then the unit count the order in d_sched was made for.  A kernel
// takes d_sched only where that count equals its own: an order made for
// another count (an eager call of another size between the replays of a
// graph that captured a scheduled call) is not a permutation of its groups.
constexpr uint32_t kSchedBins = 512;
constexpr uint32_t kSchedBinsWords = 2 * kSchedBins + 1;
inline const uint32_t* sched_units(const uint32_t* d_bins) { return d_bins + 2 * kSchedBins; }
int32_t launch_longest_first(bvht_ctx ctx, const uint32_t* d_cost, uint32_t n, uint32_t* d_bins, uint32_t* d_sched);
// The opt-in wave order of bvht_trace_blas / bvht_trace_tlas
// (bvht_set_wave_schedule), keyed by tree, ray buffer and ray count.
// trace_order_begin gives the launch its order (NULL when none applies),
// the order's unit count and the duration buffer (both NULL when the order
// is off for this call); trace_order_end, after the launch, rebuilds the
// order when due.
struct TraceOrder {
    const uint32_t* sched = nullptr;
    const uint32_t* sched_n = nullptr;
    uint32_t* cost = nullptr;
    uint32_t groups = 0;
    bool on = false;
};
int32_t trace_order_begin(bvht_ctx ctx, const void* tree, const void* rays, uint32_t n, bool eligible,
                          TraceOrder* o);
int32_t trace_order_end(bvht_ctx ctx, const void* tree, const void* rays, uint32_t n, const TraceOrder& o);
void trace_config(bvht_ctx ctx, uint32_t need, int* wg, int* lds);
int32_t launch_bounce_sorted(bvht_ctx ctx, const float* d_tris, uint32_t ntris, const void* rows, const void* blases,
                             uint32_t n_obj, const bvht_ray* d_rays, const bvht_hit* d_hits, uint32_t n,
                             uint64_t seed, bvht_ray* d_out, uint32_t* d_src, uint32_t* d_count);
int32_t hit_block_offsets(bvht_ctx ctx, const bvht_hit* d_hits, uint32_t n, uint32_t prim_limit, uint32_t inst_limit,
                          const void* rows, const void* blases, uint32_t* d_count, uint32_t** offsets);
// World bounds and traced rows of every object from its current M, M^-1
// (SceneObject::set_transform on the device).
int32_t tlas_refresh(bvht_tlas t);
// World::run_physics on the device (bvht_physics.hip).
int32_t world_run_physics(bvht_world w, float duration);

}  // namespace bvht

#define BVHT_HIP(expr)                                                                   \
    do {                                                                                 \
        hipError_t _e = (expr);                                                          \
        if (_e != hipSuccess)                                                            \
            return bvht::set_error(BVHT_EDEVICE, "%s failed: %s", #expr, hipGetErrorString(_e)); \
    } while (0)

#define BVHT_REQUIRE(cond, ...)                              \
    do {                                                     \
        if (!(cond))                                         \
            return bvht::set_error(BVHT_EINVAL, __VA_ARGS__); \
    } while (0)
