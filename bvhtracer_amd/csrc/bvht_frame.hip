// bvht_frame.hip -- per-pair triangle tests and the device side of the
// PathTracer frame: accumulators writing the accumulation buffer at each
// ray's pixel, and the pixel shaders writing the RGBA8 frame buffer.
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

}  // namespace
}  // namespace bvht

using namespace bvht;

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
