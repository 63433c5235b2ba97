// bvht_frame.hip -- per-pair triangle tests and the device side of the
// PathTracer frame: accumulators writing the accumulation buffer at each
// ray's pixel, and the pixel shaders writing the RGBA8 frame buffer.
#include <algorithm>
#include <vector>

#include "bvht_internal.hpp"
#include "bvht_trace.hpp"

namespace bvht {
namespace {

inline uint32_t nblk(size_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

// Triangle::intersect / intersect_mut (geometry/triangle.rs:41-108) for
// pair i: t = min(ray.t, t_hit), u, v; prim = i.
__global__ void k_intersect_pairs(const float* __restrict__ tris, const bvht_ray* __restrict__ rays, uint32_t n,
                                  bvht_hit* __restrict__ hits)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    DRay ray = load_ray(rays, i);
    float t = 0.0f, u = 0.0f, v = 0.0f;
    bool hit = tri_test(tris + 9 * (size_t)i, ray, t, u, v);
    store_hit(hits, i, hit, t, u, v, i, 0);
}

// Pixel of ray j for the orders of bvht_raygen.
__device__ __forceinline__ void ray_pixel(uint32_t j, uint32_t width, int32_t order, uint32_t& x, uint32_t& y)
{
    if (order == BVHT_ORDER_TILES) {
        uint32_t tile = j >> 6, w = j & 63u, tiles_x = width >> 3;
        x = (tile % tiles_x) * 8 + (w & 7u);
        y = (tile / tiles_x) * 8 + (w >> 3);
    } else {
        x = j % width;
        y = j / width;
    }
}

// Accumulator::evaluate for one ray, stored at its pixel address
// y * width + x (PathTracer::evaluate, renderer.rs:365-367):
//   DepthAccumulator (renderer.rs:183-194): t splatted, f32::MAX on a miss;
//   IntersectionAccumulator (:144-152): hit_value / miss_value;
//   UvMappingAccumulator (:222-235): (u, v, 1 - (u + v)), 0 on a miss.
__global__ void k_accumulate(const bvht_hit* __restrict__ hits, uint32_t width, uint32_t n, int32_t order,
                             int32_t kind, bvht_accum_params p, float* __restrict__ accum)
{
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n)
        return;
    uint32_t x, y;
    ray_pixel(j, width, order, x, y);
    bvht_hit h = hits[j];
    bool hit = h.prim != BVHT_MISS;
    float r0, r1, r2;
    if (kind == BVHT_ACC_DEPTH) {
        r0 = r1 = r2 = hit ? h.t : kFltMax;
    } else if (kind == BVHT_ACC_INTERSECTION) {
        const float* s = hit ? p.hit_value : p.miss_value;
        r0 = s[0];
        r1 = s[1];
        r2 = s[2];
    } else {
        r0 = hit ? h.u : 0.0f;
        r1 = hit ? h.v : 0.0f;
        r2 = hit ? 1.0f - (h.u + h.v) : 0.0f;
    }
    float* o = accum + 3 * ((size_t)y * width + x);
    o[0] = r0;
    o[1] = r1;
    o[2] = r2;
}

// Rust's saturating float -> int casts: truncation toward zero, NaN -> 0,
// out-of-range values clamped.
__device__ __forceinline__ uint32_t sat_u8(float f)
{
    if (!(f > 0.0f))
        return 0u;  // negative, zero and NaN
    if (f >= 255.0f)
        return 255u;
    return (uint32_t)f;
}
__device__ __forceinline__ int32_t sat_i32(float f)
{
    if (f != f)
        return 0;
    if (f >= 2147483648.0f)
        return 2147483647;
    if (f <= -2147483648.0f)
        return (int32_t)0x80000000u;
    return (int32_t)f;
}

// PixelShader::evaluate for every pixel:
//   RadianceToRgbShader (renderer.rs:125-132): min(255, (255 r) as u8) per
//     channel, alpha 255;
//   IntersectionShader (:166-173): hit colour iff the radiance is non-zero;
//   DepthMappingShader (:207-222): c = (255 - ((t - offset) * scale) as i32
//     as u32) * 0x010101 (u32 arithmetic wrapping, as in a release build;
//     a debug build panics where it would wrap), RGB = bytes 2, 1, 0 of c;
//     black for t = f32::MAX.
__global__ void k_shade(const float* __restrict__ accum, uint32_t n, int32_t kind, bvht_shader_params p,
                        uchar4* __restrict__ rgba)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    float r = accum[3 * (size_t)i], g = accum[3 * (size_t)i + 1], b = accum[3 * (size_t)i + 2];
    uchar4 out;
    if (kind == BVHT_SHADE_RADIANCE_TO_RGB) {
        out = make_uchar4((unsigned char)sat_u8(255.0f * r), (unsigned char)sat_u8(255.0f * g),
                          (unsigned char)sat_u8(255.0f * b), 255);
    } else if (kind == BVHT_SHADE_INTERSECTION) {
        bool nonzero = r != 0.0f || g != 0.0f || b != 0.0f;
        const uint8_t* c = nonzero ? p.hit_rgba : p.miss_rgba;
        out = make_uchar4(c[0], c[1], c[2], c[3]);
    } else {
        if (r < kFltMax) {
            uint32_t q = (uint32_t)sat_i32((r - p.offset) * p.scale);
            uint32_t c = (255u - q) * 0x010101u;
            out = make_uchar4((unsigned char)((c & 0x00FF0000u) >> 16), (unsigned char)((c & 0x0000FF00u) >> 8),
                              (unsigned char)(c & 0x000000FFu), 255);
        } else {
            out = make_uchar4(0, 0, 0, 255);
        }
    }
    rgba[i] = out;
}

// Rust's saturating `f32 as usize` (64-bit): truncation, NaN and negatives
// to 0, values from 2^64 up to usize::MAX.
__device__ __forceinline__ uint64_t sat_usize(float f)
{
    if (!(f > 0.0f))
        return 0ull;
    if (f >= 18446744073709551616.0f)
        return ~0ull;
    return (uint64_t)f;
}

// Attribute accumulators, value stored at the ray's pixel address:
//   NormalMappingAccumulator (renderer.rs:248-290): the hit triangle's three
//     corner normals interpolated as n0 (1-u-v) + n1 u + n2 v, carried to
//     world space by the object's M (transform_vector), normalised, then
//     (n + 1) * 0.5 per component;
//   TextureMaterialAccumulator (renderer.rs:293-334): the corner texture
//     coordinates interpolated the same way, the texel at
//     ((u W) as usize % W, (v H) as usize % H) (TextureMaterial::evaluate,
//     materials/material.rs:44-52), channels times 1/256.
// Misses give 0.  The primitive and object indices are the hit's true ones
// (the reference decodes them from the packed word, whose object bits are 0).
// objs == nullptr: a single-BLAS trace, object transform the identity.
template <int KIND>
__global__ void k_accumulate_attr(const bvht_hit* __restrict__ hits, uint32_t width, uint32_t n, int32_t order,
                                  const AttrRef* __restrict__ models, uint32_t n_models,
                                  const bvht_instance* __restrict__ objs, uint32_t n_objs, float* __restrict__ accum)
{
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n)
        return;
    uint32_t x, y;
    ray_pixel(j, width, order, x, y);
    bvht_hit h = hits[j];
    float r0 = 0.0f, r1 = 0.0f, r2 = 0.0f;
    uint32_t inst = objs ? h.inst : 0u;
    uint32_t model = objs ? (inst < n_objs ? objs[inst].blas : n_models) : 0u;
    if (h.prim != BVHT_MISS && model < n_models) {
        AttrRef a = models[model];
        if (h.prim < a.ntris) {
            float w = (1.0f - h.u) - h.v;
            if (KIND == BVHT_ACC_NORMAL) {
                const float* c = a.normals + 9 * (size_t)h.prim;
                V3 m = add3(add3(scl3(mk3(c[0], c[1], c[2]), w), scl3(mk3(c[3], c[4], c[5]), h.u)),
                            scl3(mk3(c[6], c[7], c[8]), h.v));
                const float ident[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
                V3 ws = mat_mul4(objs ? objs[inst].m : ident, m, 0.0f);
                float len = sqrtf(dot3(ws, ws));
                r0 = (ws.x / len + 1.0f) * 0.5f;
                r1 = (ws.y / len + 1.0f) * 0.5f;
                r2 = (ws.z / len + 1.0f) * 0.5f;
            } else {
                const float* c = a.tex_coords + 6 * (size_t)h.prim;
                float tu = ((c[0] * w) + (c[2] * h.u)) + (c[4] * h.v);
                float tv = ((c[1] * w) + (c[3] * h.u)) + (c[5] * h.v);
                uint64_t iu = sat_usize(tu * (float)a.tex_w) % a.tex_w;
                uint64_t iv = sat_usize(tv * (float)a.tex_h) % a.tex_h;
                const uint8_t* px = a.texture + 3 * (iv * a.tex_w + iu);
                const float s = 1.0f / 256.0f;
                r0 = (float)px[0] * s;
                r1 = (float)px[1] * s;
                r2 = (float)px[2] * s;
            }
        }
    }
    float* o = accum + 3 * ((size_t)y * width + x);
    o[0] = r0;
    o[1] = r1;
    o[2] = r2;
}

}  // namespace
}  // namespace bvht

using namespace bvht;

static int32_t check_frame(bvht_ctx ctx, uint32_t width, uint32_t height, int32_t order)
{
    BVHT_REQUIRE(ctx, "ctx is NULL");
    BVHT_REQUIRE(order == BVHT_ORDER_ROWS || order == BVHT_ORDER_TILES, "unknown ray order %d", order);
    BVHT_REQUIRE((uint64_t)width * height <= 0xFFFFFF00u, "frame too large");
    if (order == BVHT_ORDER_TILES)
        BVHT_REQUIRE(width % 8 == 0 && height % 8 == 0, "tile order needs W and H multiples of 8");
    return BVHT_OK;
}

static int32_t attr_missing(const AttrRef& a, int32_t kind)
{
    if (kind == BVHT_ACC_NORMAL && !a.normals)
        return set_error(BVHT_EINVAL, "a model has no normals (bvht_blas_set_attributes)");
    if (kind == BVHT_ACC_TEXTURE && (!a.tex_coords || !a.texture || !a.tex_w || !a.tex_h))
        return set_error(BVHT_EINVAL, "a model has no texture coordinates or texture (bvht_blas_set_attributes)");
    return BVHT_OK;
}

static AttrRef attr_of(bvht_blas b)
{
    return AttrRef{b->d_attr_normals, b->d_attr_tex, b->d_texture, b->tex_w, b->tex_h, b->n_tris, 0u};
}

static int32_t launch_attr(bvht_ctx ctx, const bvht_hit* d_hits, uint32_t width, uint32_t height, int32_t order,
                           int32_t kind, const AttrRef* models, uint32_t n_models, const bvht_instance* objs,
                           uint32_t n_objs, float* d_accum)
{
    uint32_t n = width * height;
    if (n == 0)
        return BVHT_OK;
    BVHT_REQUIRE(d_hits && d_accum, "NULL device buffer");
    if (kind == BVHT_ACC_NORMAL)
        hipLaunchKernelGGL(k_accumulate_attr<BVHT_ACC_NORMAL>, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, d_hits,
                           width, n, order, models, n_models, objs, n_objs, d_accum);
    else
        hipLaunchKernelGGL(k_accumulate_attr<BVHT_ACC_TEXTURE>, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream,
                           d_hits, width, n, order, models, n_models, objs, n_objs, d_accum);
    BVHT_HIP(hipGetLastError());
    return BVHT_OK;
}

extern "C" int32_t bvht_blas_set_attributes(bvht_ctx ctx, bvht_blas blas, const float* normals,
                                            const float* tex_coords, const uint8_t* texture, uint32_t tex_width,
                                            uint32_t tex_height)
{
    BVHT_REQUIRE(ctx && blas && blas->ctx == ctx, "NULL argument or BLAS of another context");
    BVHT_REQUIRE(!texture || (tex_width > 0 && tex_height > 0), "a texture needs a non-zero size");
    BVHT_REQUIRE((uint64_t)tex_width * tex_height <= (1ull << 31), "texture too large");
    hipStream_t st = ctx->stream;
    BVHT_HIP(hipStreamSynchronize(st));
    auto replace = [&](void** dst, const void* src, size_t bytes) -> int32_t {
        (void)hipFree(*dst);
        *dst = nullptr;
        if (!src)
            return BVHT_OK;
        if (hipMalloc(dst, bytes) != hipSuccess)
            return set_error(BVHT_ENOMEM, "device allocation of %zu attribute bytes failed", bytes);
        BVHT_HIP(hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, st));
        return BVHT_OK;
    };
    size_t n = blas->n_tris;
    int32_t rc = replace(reinterpret_cast<void**>(&blas->d_attr_normals), normals, sizeof(float) * 9 * n);
    if (!rc)
        rc = replace(reinterpret_cast<void**>(&blas->d_attr_tex), tex_coords, sizeof(float) * 6 * n);
    if (!rc)
        rc = replace(reinterpret_cast<void**>(&blas->d_texture), texture, 3 * (size_t)tex_width * tex_height);
    if (rc)
        return rc;
    blas->tex_w = texture ? tex_width : 0;
    blas->tex_h = texture ? tex_height : 0;
    if (!blas->d_attr_ref && hipMalloc(&blas->d_attr_ref, sizeof(AttrRef)) != hipSuccess)
        return set_error(BVHT_ENOMEM, "device allocation failed");
    AttrRef a = attr_of(blas);
    BVHT_HIP(hipMemcpyAsync(blas->d_attr_ref, &a, sizeof a, hipMemcpyHostToDevice, st));
    BVHT_HIP(hipStreamSynchronize(st));
    ++blas->attr_version;
    return BVHT_OK;
}

extern "C" int32_t bvht_accumulate_blas(bvht_ctx ctx, bvht_blas blas, const bvht_hit* d_hits, uint32_t width,
                                        uint32_t height, int32_t order, int32_t kind, float* d_accum)
{
    int32_t rc = check_frame(ctx, width, height, order);
    if (rc)
        return rc;
    BVHT_REQUIRE(blas && blas->ctx == ctx, "NULL BLAS or BLAS of another context");
    BVHT_REQUIRE(kind == BVHT_ACC_NORMAL || kind == BVHT_ACC_TEXTURE, "unknown attribute accumulator %d", kind);
    rc = attr_missing(attr_of(blas), kind);
    if (rc)
        return rc;
    return launch_attr(ctx, d_hits, width, height, order, kind, blas->d_attr_ref, 1u, nullptr, 0u, d_accum);
}

extern "C" int32_t bvht_accumulate_tlas(bvht_ctx ctx, bvht_tlas tlas, const bvht_hit* d_hits, uint32_t width,
                                        uint32_t height, int32_t order, int32_t kind, float* d_accum)
{
    int32_t rc = check_frame(ctx, width, height, order);
    if (rc)
        return rc;
    BVHT_REQUIRE(tlas && tlas->ctx == ctx, "NULL TLAS or TLAS of another context");
    BVHT_REQUIRE(kind == BVHT_ACC_NORMAL || kind == BVHT_ACC_TEXTURE, "unknown attribute accumulator %d", kind);
    size_t nb = tlas->blases.size();
    std::vector<AttrRef> table(nb);
    bool stale = !tlas->d_attr_table || tlas->attr_versions.size() != nb;
    for (size_t k = 0; k < nb; ++k) {
        table[k] = attr_of(tlas->blases[k]);
        rc = attr_missing(table[k], kind);
        if (rc)
            return rc;
        stale = stale || tlas->attr_versions[k] != tlas->blases[k]->attr_version;
    }
    if (stale) {
        if (!tlas->d_attr_table && hipMalloc(&tlas->d_attr_table, sizeof(AttrRef) * std::max<size_t>(nb, 1)) !=
                                       hipSuccess)
            return set_error(BVHT_ENOMEM, "device allocation failed");
        BVHT_HIP(hipMemcpyAsync(tlas->d_attr_table, table.data(), sizeof(AttrRef) * nb, hipMemcpyHostToDevice,
                                ctx->stream));
        tlas->attr_versions.assign(nb, 0);
        for (size_t k = 0; k < nb; ++k)
            tlas->attr_versions[k] = tlas->blases[k]->attr_version;
    }
    return launch_attr(ctx, d_hits, width, height, order, kind, tlas->d_attr_table, (uint32_t)nb, tlas->d_objects,
                       tlas->n, d_accum);
}

extern "C" int32_t bvht_intersect_pairs(bvht_ctx ctx, const float* d_tris, const bvht_ray* d_rays, uint32_t n,
                                        bvht_hit* d_hits)
{
    BVHT_REQUIRE(ctx, "ctx is NULL");
    if (n == 0)
        return BVHT_OK;
    BVHT_REQUIRE(d_tris && d_rays && d_hits, "NULL device buffer");
    BVHT_REQUIRE(n <= 0xFFFFFF00u, "too many pairs (%u)", n);
    hipLaunchKernelGGL(k_intersect_pairs, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, d_tris, d_rays, n, d_hits);
    BVHT_HIP(hipGetLastError());
    return BVHT_OK;
}

extern "C" int32_t bvht_accumulate(bvht_ctx ctx, const bvht_hit* d_hits, uint32_t width, uint32_t height,
                                   int32_t order, int32_t kind, const bvht_accum_params* params, float* d_accum)
{
    BVHT_REQUIRE(ctx, "ctx is NULL");
    BVHT_REQUIRE(order == BVHT_ORDER_ROWS || order == BVHT_ORDER_TILES, "unknown ray order %d", order);
    BVHT_REQUIRE(kind == BVHT_ACC_DEPTH || kind == BVHT_ACC_INTERSECTION || kind == BVHT_ACC_UV,
                 "unknown accumulator %d", kind);
    BVHT_REQUIRE(kind != BVHT_ACC_INTERSECTION || params, "the intersection accumulator needs its values");
    BVHT_REQUIRE((uint64_t)width * height <= 0xFFFFFF00u, "frame too large");
    if (order == BVHT_ORDER_TILES)
        BVHT_REQUIRE(width % 8 == 0 && height % 8 == 0, "tile order needs W and H multiples of 8");
    uint32_t n = width * height;
    if (n == 0)
        return BVHT_OK;
    BVHT_REQUIRE(d_hits && d_accum, "NULL device buffer");
    bvht_accum_params p{};
    if (params)
        p = *params;
    hipLaunchKernelGGL(k_accumulate, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, d_hits, width, n, order, kind, p,
                       d_accum);
    BVHT_HIP(hipGetLastError());
    return BVHT_OK;
}

extern "C" int32_t bvht_shade(bvht_ctx ctx, const float* d_accum, uint32_t n_pixels, int32_t kind,
                              const bvht_shader_params* params, uint8_t* d_rgba)
{
    BVHT_REQUIRE(ctx, "ctx is NULL");
    BVHT_REQUIRE(kind == BVHT_SHADE_RADIANCE_TO_RGB || kind == BVHT_SHADE_INTERSECTION ||
                     kind == BVHT_SHADE_DEPTH_MAPPING,
                 "unknown pixel shader %d", kind);
    BVHT_REQUIRE(kind == BVHT_SHADE_RADIANCE_TO_RGB || params, "this pixel shader needs its parameters");
    BVHT_REQUIRE(n_pixels <= 0xFFFFFF00u, "frame too large");
    if (n_pixels == 0)
        return BVHT_OK;
    BVHT_REQUIRE(d_accum && d_rgba, "NULL device buffer");
    bvht_shader_params p{};
    if (params)
        p = *params;
    hipLaunchKernelGGL(k_shade, dim3(nblk(n_pixels, 256)), dim3(256), 0, ctx->stream, d_accum, n_pixels, kind, p,
                       reinterpret_cast<uchar4*>(d_rgba));
    BVHT_HIP(hipGetLastError());
    return BVHT_OK;
}
