/*
 * bvhtracer.hpp -- C++17 host API over the C ABI (bvht.h), mirroring the
 * reference crate's Rust API for the traversal path (bvhtracer/src, paths
 * relative to the reference repository root): the same type and method
 * names, argument meaning and results, so a caller -- or a test written like
 * the reference's own tests -- reads the same.  Header-only; link
 * libbvht_hip.so.
 *
 * Every query runs on the GPU through libbvht_hip.so; these classes hold
 * host bookkeeping (meshes, transforms, object lists) and device handles.
 * There is no CPU path.
 *
 * Mapping notes
 *   - Rust `T::new(..)` constructors are C++ constructors; `Option<T>` is
 *     std::optional<T>; `Rc<RefCell<Model>>` is std::shared_ptr<Model>.
 *   - Where the reference panics (empty mesh, bvh.rs:530-535; empty TLAS,
 *     tlas.rs:248) or the library reports an error, bvhtracer::Error is
 *     thrown with the library's message and status code.
 *   - Covered: geometry and query types, BvhBuilder/Bvh, Mesh/Model/
 *     ModelInstance, SceneObject/Scene (TLAS), Camera/PathTracer/Renderer
 *     with the depth, intersection and uv accumulators and the pixel
 *     shaders and the normal-mapping and texture accumulators, the
 *     rigid-body World, and this library's own batched trace (Batch) and
 *     forward/adjoint pair (DiffMesh).
 *   - Vector arithmetic is binary32, left to right (dot = x*x' + y*y' +
 *     z*z', normalize = v / sqrt(dot(v, v))), the convention the GPU path
 *     and its tests use.  Compile with -ffp-contract=off.
 */
#ifndef BVHTRACER_HPP
#define BVHTRACER_HPP

#include <array>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "bvht.h"

namespace bvhtracer {

/* ---- errors ------------------------------------------------------------ */

class Error : public std::runtime_error {
public:
    Error(int32_t code, const std::string& what) : std::runtime_error(what), code_(code) {}
    int32_t code() const { return code_; }

private:
    int32_t code_;
};

inline void check(int32_t rc)
{
    if (rc != BVHT_OK)
        throw Error(rc, bvht_last_error());
}

/* ---- device context and buffers ---------------------------------------- */

/* One HIP device and stream (bvht_ctx). */
class Context {
public:
    explicit Context(int32_t device = 0) { check(bvht_create(device, &ctx_)); }
    ~Context()
    {
        if (ctx_)
            bvht_destroy(ctx_);
    }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    bvht_ctx handle() const { return ctx_; }
    void sync() const { check(bvht_sync(ctx_)); }
    /* Trace kernel configuration (workgroup x LDS stack entries; wg = 0
     * restores the depth-based default), bvht_set_trace_config. */
    void set_trace_config(int32_t wg, int32_t lds_depth) const { check(bvht_set_trace_config(ctx_, wg, lds_depth)); }
    /* Bound on the traversal stack (0: the tree depth), bvht_set_stack_limit. */
    void set_stack_limit(uint32_t max_entries) const { check(bvht_set_stack_limit(ctx_, max_entries)); }
    /* Longest-first wave order for repeated traces of the same device rays,
     * bvht_set_wave_schedule; hits are the same either way. */
    void set_wave_schedule(bool enable) const { check(bvht_set_wave_schedule(ctx_, enable ? 1 : 0)); }

private:
    bvht_ctx ctx_ = nullptr;
};

/* The context the value types below use when none is given (device 0),
 * created on first use.  It is never destroyed: the HIP runtime may be torn
 * down before static destructors run, so the process exit releases it. */
inline std::shared_ptr<Context> default_context()
{
    static auto* ctx = new std::shared_ptr<Context>(std::make_shared<Context>(0));
    return *ctx;
}

/* Device array of n T (bvht_device_alloc), freed with the object. */
template <class T>
class DeviceBuffer {
public:
    DeviceBuffer(std::shared_ptr<Context> ctx, size_t n) : ctx_(std::move(ctx)), n_(n)
    {
        check(bvht_device_alloc(ctx_->handle(), sizeof(T) * (n ? n : 1), &ptr_));
    }
    DeviceBuffer(std::shared_ptr<Context> ctx, const std::vector<T>& host) : DeviceBuffer(std::move(ctx), host.size())
    {
        upload(host);
    }
    ~DeviceBuffer()
    {
        if (ptr_)
            bvht_device_free(ctx_->handle(), ptr_);
    }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    T* data() const { return static_cast<T*>(ptr_); }
    size_t size() const { return n_; }
    void upload(const std::vector<T>& host)
    {
        if (host.size() > n_)
            throw Error(BVHT_EINVAL, "upload larger than the device buffer");
        if (!host.empty())
            check(bvht_copy_to_device(ctx_->handle(), ptr_, host.data(), sizeof(T) * host.size()));
    }
    std::vector<T> download(size_t count) const
    {
        std::vector<T> out(count);
        if (count)
            check(bvht_copy_to_host(ctx_->handle(), out.data(), ptr_, sizeof(T) * count));
        return out;
    }

private:
    std::shared_ptr<Context> ctx_;
    size_t n_;
    void* ptr_ = nullptr;
};

/* ---- vectors and matrices (cglinalg's Vector3 / Matrix4x4 subset) ------ */

struct Vector3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;

    constexpr Vector3() = default;
    constexpr Vector3(float x_, float y_, float z_) : x(x_), y(y_), z(z_) {}
    static constexpr Vector3 zero() { return Vector3(); }

    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    Vector3 operator+(const Vector3& o) const { return {x + o.x, y + o.y, z + o.z}; }
    Vector3 operator-(const Vector3& o) const { return {x - o.x, y - o.y, z - o.z}; }
    Vector3 operator*(float s) const { return {x * s, y * s, z * s}; }
    Vector3 operator/(float s) const { return {x / s, y / s, z / s}; }
    Vector3& operator+=(const Vector3& o) { return *this = *this + o; }
    Vector3& operator-=(const Vector3& o) { return *this = *this - o; }
    bool operator==(const Vector3& o) const { return x == o.x && y == o.y && z == o.z; }
    bool operator!=(const Vector3& o) const { return !(*this == o); }

    float dot(const Vector3& o) const { return x * o.x + y * o.y + z * o.z; }
    Vector3 cross(const Vector3& o) const { return {y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x}; }
    float magnitude() const { return std::sqrt(dot(*this)); }
    Vector3 normalize() const
    {
        float n = magnitude();
        return {x / n, y / n, z / n};
    }
};
inline Vector3 operator*(float s, const Vector3& v) { return v * s; }

/* Column-major 4x4 (m[col * 4 + row]), the layout of bvht_instance. */
struct Matrix4x4 {
    std::array<float, 16> m{};

    static Matrix4x4 identity()
    {
        Matrix4x4 r;
        r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0f;
        return r;
    }
    static Matrix4x4 from_affine_translation(const Vector3& t)
    {
        Matrix4x4 r = identity();
        r.m[12] = t.x, r.m[13] = t.y, r.m[14] = t.z;
        return r;
    }
    static Matrix4x4 from_affine_scale(float s)
    {
        Matrix4x4 r = identity();
        r.m[0] = r.m[5] = r.m[10] = s;
        return r;
    }
    float at(int row, int col) const { return m[col * 4 + row]; }
    /* Inverse in binary64, rounded to binary32 (the cached M^-1 of a scene
     * object; transform_component.rs:7-27).  Throws on a singular matrix. */
    Matrix4x4 inverse() const
    {
        double a[4][8];
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 8; ++c)
                a[r][c] = c < 4 ? (double)at(r, c) : (c - 4 == r ? 1.0 : 0.0);
        for (int c = 0; c < 4; ++c) {
            int p = c;
            for (int r = c + 1; r < 4; ++r)
                if (std::fabs(a[r][c]) > std::fabs(a[p][c]))
                    p = r;
            if (a[p][c] == 0.0)
                throw Error(BVHT_EINVAL, "singular transform");
            for (int k = 0; k < 8; ++k)
                std::swap(a[c][k], a[p][k]);
            double inv = 1.0 / a[c][c];
            for (int k = 0; k < 8; ++k)
                a[c][k] *= inv;
            for (int r = 0; r < 4; ++r)
                if (r != c) {
                    double f = a[r][c];
                    for (int k = 0; k < 8; ++k)
                        a[r][k] -= f * a[c][k];
                }
        }
        Matrix4x4 out;
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c)
                out.m[c * 4 + r] = (float)a[r][c + 4];
        return out;
    }
};

/* ---- query types (query/ray.rs, query/intersection.rs) ------------------ */

struct Ray {
    Vector3 origin;
    Vector3 direction;
    float t = FLT_MAX;

    Ray() = default;
    Ray(const Vector3& origin_, const Vector3& direction_, float t_) : origin(origin_), direction(direction_), t(t_) {}
    static Ray from_origin_dir(const Vector3& origin, const Vector3& direction)
    {
        return Ray(origin, direction, FLT_MAX);
    }
    Vector3 interpolate(float s) const { return origin + direction * s; }
    bool operator==(const Ray& o) const { return origin == o.origin && direction == o.direction && t == o.t; }

    bvht_ray record() const
    {
        bvht_ray r{};
        r.o[0] = origin.x, r.o[1] = origin.y, r.o[2] = origin.z;
        r.d[0] = direction.x, r.d[1] = direction.y, r.d[2] = direction.z;
        r.tmax = t;
        return r;
    }
};

struct SurfaceInteraction {
    float t = 0.0f, u = 0.0f, v = 0.0f;

    SurfaceInteraction() = default;
    SurfaceInteraction(float t_, float u_, float v_) : t(t_), u(u_), v(v_) {}
    bool operator==(const SurfaceInteraction& o) const { return t == o.t && u == o.u && v == o.v; }
};

/* Instance and primitive index packed into one word: 12 instance bits, 20
 * primitive bits (intersection.rs:33-67). */
class InstancePrimitiveIndex {
public:
    constexpr InstancePrimitiveIndex() = default;
    constexpr InstancePrimitiveIndex(uint32_t instance_index, uint32_t primitive_index)
        : data_(((instance_index & 0x00000FFFu) << 20) | (primitive_index & 0x000FFFFFu))
    {
    }
    static constexpr InstancePrimitiveIndex from_primitive(uint32_t primitive_index)
    {
        return InstancePrimitiveIndex(0, primitive_index);
    }
    static constexpr InstancePrimitiveIndex from_word(uint32_t word)
    {
        InstancePrimitiveIndex r;
        r.data_ = word;
        return r;
    }
    constexpr uint32_t instance_index() const { return (data_ & 0xFFF00000u) >> 20; }
    constexpr uint32_t primitive_index() const { return data_ & 0x000FFFFFu; }
    constexpr bool operator==(const InstancePrimitiveIndex& o) const { return data_ == o.data_; }

private:
    uint32_t data_ = 0;
};

struct Intersection {
    Ray ray;
    SurfaceInteraction interaction;
    InstancePrimitiveIndex instance_primitive;
    /* The hit's true primitive and object indices (the packed word above
     * keeps the reference's instance bits, always 0). */
    uint32_t primitive = BVHT_MISS;
    uint32_t object = BVHT_MISS;

    /* The query ray with t = the hit's t, as the reference's intersect_mut
     * leaves it; nullopt on a miss. */
    static std::optional<Intersection> from_hit(const Ray& query, const bvht_hit& h)
    {
        if (h.prim == BVHT_MISS)
            return std::nullopt;
        Intersection r;
        r.ray = Ray(query.origin, query.direction, h.t);
        r.interaction = SurfaceInteraction(h.t, h.u, h.v);
        r.instance_primitive = InstancePrimitiveIndex::from_word(h.packed);
        r.primitive = h.prim;
        r.object = h.inst;
        return r;
    }
};

/* ---- geometry (geometry/aabb.rs, geometry/triangle.rs) ------------------ */

struct Aabb {
    Vector3 bounds_min;
    Vector3 bounds_max;

    Aabb() = default;
    Aabb(const Vector3& bmin, const Vector3& bmax) : bounds_min(bmin), bounds_max(bmax) {}
    Vector3 extent() const { return bounds_max - bounds_min; }
    bool operator==(const Aabb& o) const { return bounds_min == o.bounds_min && bounds_max == o.bounds_max; }
};

struct Triangle;
std::vector<bvht_hit> intersect_pairs(const std::vector<Triangle>& tris, const std::vector<Ray>& rays,
                                      const std::shared_ptr<Context>& ctx = default_context());

struct Triangle {
    std::array<Vector3, 3> vertices;

    Triangle() = default;
    Triangle(const Vector3& v0, const Vector3& v1, const Vector3& v2) : vertices{v0, v1, v2} {}
    Vector3 centroid() const
    {
        float one_third = 1.0f / 3.0f;
        return (vertices[0] + vertices[1] + vertices[2]) * one_third;
    }
    bool operator==(const Triangle& o) const { return vertices == o.vertices; }

    /* Triangle::intersect (triangle.rs:41-72): t = min(ray.t, t), u, v. */
    std::optional<SurfaceInteraction> intersect(const Ray& ray) const
    {
        bvht_hit h = intersect_pairs({*this}, {ray})[0];
        if (h.prim == BVHT_MISS)
            return std::nullopt;
        return SurfaceInteraction(h.t, h.u, h.v);
    }
    /* Triangle::intersect_mut (triangle.rs:74-108). */
    bool intersect_mut(Intersection& intersection) const
    {
        auto s = intersect(intersection.ray);
        if (!s)
            return false;
        intersection.ray.t = s->t;
        intersection.interaction = *s;
        return true;
    }
};
static_assert(sizeof(Triangle) == 36, "Triangle must be nine packed floats");

/* Triangle::intersect for n independent (triangle, ray) pairs in one GPU
 * call (bvht_intersect_pairs). */
inline std::vector<bvht_hit> intersect_pairs(const std::vector<Triangle>& tris, const std::vector<Ray>& rays,
                                             const std::shared_ptr<Context>& ctx)
{
    if (tris.size() != rays.size())
        throw Error(BVHT_EINVAL, "intersect_pairs: one triangle per ray");
    if (tris.empty())
        return {};
    std::vector<bvht_ray> recs;
    recs.reserve(rays.size());
    for (const Ray& r : rays)
        recs.push_back(r.record());
    DeviceBuffer<Triangle> d_tris(ctx, tris);
    DeviceBuffer<bvht_ray> d_rays(ctx, recs);
    DeviceBuffer<bvht_hit> d_hits(ctx, tris.size());
    check(bvht_intersect_pairs(ctx->handle(), reinterpret_cast<const float*>(d_tris.data()), d_rays.data(),
                               (uint32_t)tris.size(), d_hits.data()));
    return d_hits.download(tris.size());
}

/* ---- BVH (model/bvh.rs) ------------------------------------------------- */

class Bvh {
public:
    Bvh(std::shared_ptr<Context> ctx, bvht_blas blas) : ctx_(std::move(ctx)), blas_(blas) {}
    ~Bvh()
    {
        if (blas_)
            bvht_blas_destroy(blas_);
    }
    Bvh(const Bvh&) = delete;
    Bvh& operator=(const Bvh&) = delete;
    Bvh(Bvh&& o) noexcept : ctx_(std::move(o.ctx_)), blas_(std::exchange(o.blas_, nullptr)) {}

    bvht_blas handle() const { return blas_; }
    const std::shared_ptr<Context>& context() const { return ctx_; }

    /* Bvh::intersect (bvh.rs:307-309).  `mesh` is the build-ordered slice the
     * BVH was built over (kept for the reference's signature; the GPU holds
     * its own copy). */
    std::optional<Intersection> intersect(const std::vector<Triangle>& /*mesh*/, const Ray& ray) const
    {
        return Intersection::from_hit(ray, intersect_batch({ray})[0]);
    }
    /* Bvh::intersect for a batch of rays in one launch. */
    std::vector<bvht_hit> intersect_batch(const std::vector<Ray>& rays) const
    {
        if (rays.empty())
            return {};
        std::vector<bvht_ray> recs;
        recs.reserve(rays.size());
        for (const Ray& r : rays)
            recs.push_back(r.record());
        DeviceBuffer<bvht_ray> d_rays(ctx_, recs);
        DeviceBuffer<bvht_hit> d_hits(ctx_, rays.size());
        check(bvht_trace_blas(ctx_->handle(), blas_, d_rays.data(), (uint32_t)rays.size(), d_hits.data(), nullptr));
        return d_hits.download(rays.size());
    }
    /* Bvh::refit (bvh.rs:469-493) against the mesh's current vertices. */
    void refit(const std::vector<Triangle>& mesh)
    {
        check(bvht_blas_set_vertices(ctx_->handle(), blas_, reinterpret_cast<const float*>(mesh.data()),
                                     (uint32_t)mesh.size()));
        check(bvht_blas_refit(ctx_->handle(), blas_));
    }
    /* Bvh::bounds (bvh.rs:499-501). */
    Aabb bounds() const
    {
        float b[6];
        check(bvht_blas_bounds(ctx_->handle(), blas_, b));
        return Aabb(Vector3(b[0], b[1], b[2]), Vector3(b[3], b[4], b[5]));
    }
    uint32_t nodes_used() const { return info()[2]; }
    uint32_t depth() const { return info()[3]; }
    /* The node array (bvh.rs:228-233) as the device holds it. */
    std::vector<bvht_node> nodes() const
    {
        std::vector<bvht_node> out(info()[1]);
        check(bvht_blas_download(ctx_->handle(), blas_, out.data(), nullptr));
        return out;
    }
    /* Bvh's derived PartialEq: the same node array and node count. */
    bool operator==(const Bvh& o) const
    {
        auto a = nodes(), b = o.nodes();
        return nodes_used() == o.nodes_used() && a.size() == b.size() &&
               std::memcmp(a.data(), b.data(), a.size() * sizeof(bvht_node)) == 0;
    }

private:
    std::array<uint32_t, 4> info() const
    {
        std::array<uint32_t, 4> v{};
        check(bvht_blas_info(blas_, &v[0], &v[1], &v[2], &v[3]));
        return v;
    }

    std::shared_ptr<Context> ctx_;
    bvht_blas blas_ = nullptr;
};

/* BvhBuilder (bvh.rs:504-541).  build_for permutes `mesh` in place into the
 * build order, as the reference reorders its slice.  mode BVHT_BUILD_SAH with
 * a leaf cap is this library's binned-SAH variant. */
class BvhBuilder {
public:
    explicit BvhBuilder(std::shared_ptr<Context> ctx = default_context(), int32_t mode = BVHT_BUILD_REFERENCE,
                        uint32_t max_leaf = 0)
        : ctx_(std::move(ctx)), mode_(mode), max_leaf_(max_leaf)
    {
    }
    Bvh build_for(std::vector<Triangle>& mesh) const
    {
        bvht_blas b = nullptr;
        check(bvht_blas_build(ctx_->handle(), reinterpret_cast<float*>(mesh.data()), (uint32_t)mesh.size(), mode_,
                              max_leaf_, &b));
        return Bvh(ctx_, b);
    }

private:
    std::shared_ptr<Context> ctx_;
    int32_t mode_;
    uint32_t max_leaf_;
};

/* ---- mesh and model (mesh/mesh.rs, model/model.rs) ---------------------- */

struct TextureCoordinates {
    std::array<std::array<float, 2>, 3> uv{};
};
struct Normals {
    std::array<Vector3, 3> normals{};
};

class Mesh {
public:
    const std::vector<Triangle>& primitives() const { return primitives_; }
    std::vector<Triangle>& primitives_mut() { return primitives_; }
    const std::vector<TextureCoordinates>& tex_coords() const { return tex_coords_; }
    const std::vector<Normals>& normals() const { return normals_; }
    size_t len() const { return primitives_.size(); }
    /* Reorder the per-triangle attributes like the triangles: new[k] =
     * old[perm[k]] (the BVH build's permutation; the reference permutes only
     * the triangles, bvh.rs:420-429, leaving attributes misaligned). */
    void permute(const std::vector<uint32_t>& perm)
    {
        std::vector<TextureCoordinates> t(perm.size());
        std::vector<Normals> nr(perm.size());
        for (size_t k = 0; k < perm.size(); ++k)
            t[k] = tex_coords_.at(perm[k]), nr[k] = normals_.at(perm[k]);
        tex_coords_.swap(t);
        normals_.swap(nr);
    }

private:
    friend class MeshBuilder;
    std::vector<Triangle> primitives_;
    std::vector<TextureCoordinates> tex_coords_;
    std::vector<Normals> normals_;
};

class MeshBuilder {
public:
    MeshBuilder& with_primitive(const Triangle& primitive, const TextureCoordinates& tex_coords,
                                const Normals& normals)
    {
        mesh_.primitives_.push_back(primitive);
        mesh_.tex_coords_.push_back(tex_coords);
        mesh_.normals_.push_back(normals);
        return *this;
    }
    Mesh build() const { return mesh_; }

private:
    Mesh mesh_;
};

/* Model (model.rs:55-110): a mesh (triangles in build order) and its BVH on
 * the GPU. */
class Model {
public:
    Model(Mesh mesh, Bvh bvh) : mesh_(std::move(mesh)), bvh_(std::move(bvh)) {}
    std::optional<Intersection> intersect(const Ray& ray) const { return bvh_.intersect(mesh_.primitives(), ray); }
    /* Model::refit (model.rs:69-71) against the current primitives. */
    void refit() { bvh_.refit(mesh_.primitives()); }
    Aabb bounds() const { return bvh_.bounds(); }
    const Mesh& mesh() const { return mesh_; }
    const std::vector<Triangle>& primitives() const { return mesh_.primitives(); }
    std::vector<Triangle>& primitives_mut() { return mesh_.primitives_mut(); }
    size_t len() const { return mesh_.len(); }
    size_t len_primitives() const { return mesh_.len(); }
    const Bvh& bvh() const { return bvh_; }

private:
    Mesh mesh_;
    Bvh bvh_;
};

/* ModelInstance (model.rs:15-53): shared ownership of one Model. */
class ModelInstance {
public:
    explicit ModelInstance(std::shared_ptr<Model> model) : model_(std::move(model)) {}
    std::optional<Intersection> intersect(const Ray& ray) const { return model_->intersect(ray); }
    void refit() { model_->refit(); }
    Aabb bounds() const { return model_->bounds(); }
    std::shared_ptr<Model> model() const { return model_; }
    size_t len() const { return model_->len(); }
    size_t len_primitives() const { return model_->len_primitives(); }

private:
    std::shared_ptr<Model> model_;
};

/* ModelBuilder (model.rs:113-150): build() builds the BVH on the GPU over
 * the mesh's primitives, which end up in build order. */
class ModelBuilder {
public:
    explicit ModelBuilder(std::shared_ptr<Context> ctx = default_context()) : ctx_(std::move(ctx)) {}
    ModelBuilder& with_mesh(Mesh mesh)
    {
        mesh_ = std::move(mesh);
        return *this;
    }
    ModelBuilder& with_build(int32_t mode, uint32_t max_leaf = 0)
    {
        mode_ = mode, max_leaf_ = max_leaf;
        return *this;
    }
    /* ModelBuilder::with_texture (model.rs:134-139): an RGB8 image,
     * row-major, width x height. */
    ModelBuilder& with_texture(uint32_t width, uint32_t height, std::vector<uint8_t> rgb)
    {
        if (rgb.size() != (size_t)3 * width * height)
            throw Error(BVHT_EINVAL, "with_texture: rgb must hold 3 * width * height bytes");
        tex_w_ = width, tex_h_ = height, texture_ = std::move(rgb);
        return *this;
    }
    /* Builds the BVH on the GPU (the mesh's triangles end up in build order,
     * its attributes permuted with them) and uploads the normals, texture
     * coordinates and texture for the attribute accumulators. */
    ModelInstance build()
    {
        Bvh bvh = BvhBuilder(ctx_, mode_, max_leaf_).build_for(mesh_.primitives_mut());
        size_t n = mesh_.len();
        std::vector<uint32_t> perm(n);
        if (mode_ == BVHT_BUILD_REFERENCE || mode_ == BVHT_BUILD_SAH)
            check(bvht_blas_permutation(ctx_->handle(), bvh.handle(), perm.data()));
        mesh_.permute(perm);
        std::vector<float> nrm(9 * n), tc(6 * n);
        for (size_t k = 0; k < n; ++k)
            for (int c = 0; c < 3; ++c) {
                const Vector3& v = mesh_.normals()[k].normals[c];
                nrm[9 * k + 3 * c] = v.x, nrm[9 * k + 3 * c + 1] = v.y, nrm[9 * k + 3 * c + 2] = v.z;
                tc[6 * k + 2 * c] = mesh_.tex_coords()[k].uv[c][0];
                tc[6 * k + 2 * c + 1] = mesh_.tex_coords()[k].uv[c][1];
            }
        check(bvht_blas_set_attributes(ctx_->handle(), bvh.handle(), nrm.data(), tc.data(),
                                       texture_.empty() ? nullptr : texture_.data(), tex_w_, tex_h_));
        return ModelInstance(std::make_shared<Model>(std::move(mesh_), std::move(bvh)));
    }

private:
    std::shared_ptr<Context> ctx_;
    std::vector<uint8_t> texture_;
    uint32_t tex_w_ = 0, tex_h_ = 0;
    Mesh mesh_;
    int32_t mode_ = BVHT_BUILD_REFERENCE;
    uint32_t max_leaf_ = 0;
};

/* ---- scene (scene/scene_object.rs, scene/tlas.rs, scene/scene.rs) ------- */

class Tlas;

/* SceneObject (scene_object.rs:11-89): a model instance placed by M, with
 * its cached inverse.  intersect() and bounds() run on the GPU through a
 * one-object TLAS made on first use (and remade after set_transform). */
class SceneObject {
public:
    SceneObject(ModelInstance model, const Matrix4x4& transform)
        : model_(std::move(model)), m_(transform), minv_(transform.inverse())
    {
    }
    const ModelInstance& model() const { return model_; }
    const Matrix4x4& transform() const { return m_; }
    const Matrix4x4& transform_inv() const { return minv_; }
    /* SceneObject::set_transform (scene_object.rs:60-75); a scene sees it at
     * its next rebuild. */
    void set_transform(const Matrix4x4& transform)
    {
        m_ = transform;
        minv_ = transform.inverse();
        alone_.reset();
    }
    /* SceneObject::intersect (scene_object.rs:77-89): the world-space ray
     * carried into model space by M^-1, traced against the model. */
    std::optional<Intersection> intersect(const Ray& ray) const;
    /* SceneObject::bounds (scene_object.rs:77-79): world-space bounds of the
     * model's bounds under M. */
    Aabb bounds() const;

private:
    const Tlas& alone() const;
    ModelInstance model_;
    Matrix4x4 m_, minv_;
    mutable std::shared_ptr<Tlas> alone_;
};

class SceneObjectBuilder {
public:
    explicit SceneObjectBuilder(ModelInstance model) : model_(std::move(model)) {}
    SceneObjectBuilder& with_transform(const Matrix4x4& transform)
    {
        m_ = transform;
        return *this;
    }
    SceneObject build() const { return SceneObject(model_, m_); }

private:
    ModelInstance model_;
    Matrix4x4 m_ = Matrix4x4::identity();
};

/* Tlas (scene/tlas.rs): the top-level BVH over scene objects, on the GPU.
 * Objects made from one ModelInstance share one BLAS; rebuild() takes the
 * same objects' current transforms (the set of models is fixed at build). */
class Tlas {
public:
    Tlas(std::shared_ptr<Context> ctx, const std::vector<SceneObject>& objects) : ctx_(std::move(ctx))
    {
        std::vector<bvht_blas> blases;
        std::vector<bvht_instance> recs = records(objects, &blases);
        check(bvht_tlas_create(ctx_->handle(), blases.data(), (uint32_t)blases.size(), recs.data(),
                               (uint32_t)recs.size(), &tlas_));
        check(bvht_tlas_rebuild(ctx_->handle(), tlas_));
    }
    ~Tlas()
    {
        if (tlas_)
            bvht_tlas_destroy(tlas_);
    }
    Tlas(const Tlas&) = delete;
    Tlas& operator=(const Tlas&) = delete;
    bvht_tlas handle() const { return tlas_; }

    /* Tlas::intersect (tlas.rs:123-177).  `objects` are the ones the TLAS
     * was built or rebuilt from (the GPU holds their transforms). */
    std::optional<Intersection> intersect(const std::vector<SceneObject>& /*objects*/, const Ray& ray) const
    {
        return Intersection::from_hit(ray, intersect_batch({ray})[0]);
    }
    std::vector<bvht_hit> intersect_batch(const std::vector<Ray>& rays) const
    {
        if (rays.empty())
            return {};
        std::vector<bvht_ray> recs;
        recs.reserve(rays.size());
        for (const Ray& r : rays)
            recs.push_back(r.record());
        DeviceBuffer<bvht_ray> d_rays(ctx_, recs);
        DeviceBuffer<bvht_hit> d_hits(ctx_, rays.size());
        check(bvht_trace_tlas(ctx_->handle(), tlas_, d_rays.data(), (uint32_t)rays.size(), d_hits.data(), nullptr));
        return d_hits.download(rays.size());
    }
    /* Tlas::rebuild (tlas.rs:204-250): the objects' transforms and world
     * bounds from the models' current bounds, then the tree. */
    void rebuild(const std::vector<SceneObject>& objects)
    {
        std::vector<bvht_instance> recs = records(objects, nullptr);
        check(bvht_tlas_set_objects(ctx_->handle(), tlas_, recs.data(), 0, (uint32_t)recs.size()));
        check(bvht_tlas_rebuild(ctx_->handle(), tlas_));
    }
    /* The node array (tlas.rs:41-46) and its used count. */
    std::vector<bvht_tnode> nodes(uint32_t* nodes_used = nullptr) const
    {
        uint32_t n = 0;
        std::vector<bvht_tnode> out(2 * count_);
        check(bvht_tlas_download(ctx_->handle(), tlas_, out.data(), nodes_used ? nodes_used : &n));
        return out;
    }

private:
    std::vector<bvht_instance> records(const std::vector<SceneObject>& objects, std::vector<bvht_blas>* blases)
    {
        count_ = (uint32_t)objects.size();
        std::vector<bvht_instance> recs(objects.size());
        for (size_t i = 0; i < objects.size(); ++i) {
            const Model* m = objects[i].model().model().get();
            uint32_t k = 0;
            while (k < models_.size() && models_[k] != m)
                ++k;
            if (k == models_.size()) {
                if (!blases)
                    throw Error(BVHT_EINVAL, "rebuild: an object names a model the TLAS was not built with");
                models_.push_back(m);
                blases->push_back(m->bvh().handle());
            }
            std::memcpy(recs[i].m, objects[i].transform().m.data(), sizeof(recs[i].m));
            std::memcpy(recs[i].minv, objects[i].transform_inv().m.data(), sizeof(recs[i].minv));
            recs[i].blas = k;
        }
        return recs;
    }

    std::shared_ptr<Context> ctx_;
    bvht_tlas tlas_ = nullptr;
    std::vector<const Model*> models_;
    uint32_t count_ = 0;
};

/* TlasBuilder (tlas.rs:258-279). */
class TlasBuilder {
public:
    explicit TlasBuilder(std::shared_ptr<Context> ctx = default_context()) : ctx_(std::move(ctx)) {}
    std::unique_ptr<Tlas> build_for(const std::vector<SceneObject>& objects) const
    {
        return std::make_unique<Tlas>(ctx_, objects);
    }

private:
    std::shared_ptr<Context> ctx_;
};

inline const Tlas& SceneObject::alone() const
{
    if (!alone_)
        alone_ = std::make_shared<Tlas>(model_.model()->bvh().context(), std::vector<SceneObject>{*this});
    return *alone_;
}
inline std::optional<Intersection> SceneObject::intersect(const Ray& ray) const
{
    return alone().intersect({}, ray);
}
inline Aabb SceneObject::bounds() const
{
    uint32_t used = 0;
    bvht_tnode root = alone().nodes(&used)[0];
    return Aabb(Vector3(root.bmin[0], root.bmin[1], root.bmin[2]), Vector3(root.bmax[0], root.bmax[1], root.bmax[2]));
}

/* Scene (scene.rs:7-58): scene objects and their TLAS on the GPU. */
class Scene {
public:
    Scene(std::shared_ptr<Context> ctx, std::vector<SceneObject> objects)
        : objects_(std::move(objects)), tlas_(std::make_unique<Tlas>(std::move(ctx), objects_))
    {
    }
    /* Scene::intersect (scene.rs:32-34) for a world-space ray. */
    std::optional<Intersection> intersect(const Ray& ray) const { return tlas_->intersect(objects_, ray); }
    std::vector<bvht_hit> intersect_batch(const std::vector<Ray>& rays) const { return tlas_->intersect_batch(rays); }
    /* Scene::rebuild (scene.rs:36-38). */
    void rebuild() { tlas_->rebuild(objects_); }
    bvht_tlas handle() const { return tlas_->handle(); }
    const Tlas& tlas() const { return *tlas_; }
    SceneObject& get_mut_unchecked(size_t index) { return objects_[index]; }
    const SceneObject& get_unchecked(size_t index) const { return objects_[index]; }
    size_t len() const { return objects_.size(); }

private:
    std::vector<SceneObject> objects_;
    std::unique_ptr<Tlas> tlas_;
};

/* SceneBuilder (scene.rs:61-96). */
class SceneBuilder {
public:
    explicit SceneBuilder(std::shared_ptr<Context> ctx = default_context()) : ctx_(std::move(ctx)) {}
    SceneBuilder& with_object(SceneObject object)
    {
        objects_.push_back(std::move(object));
        return *this;
    }
    std::unique_ptr<Scene> build() { return std::make_unique<Scene>(ctx_, std::move(objects_)); }

private:
    std::shared_ptr<Context> ctx_;
    std::vector<SceneObject> objects_;
};

/* ---- batched scenes (no reference counterpart) -------------------------- */

/* Many scenes updated and traced with one launch per stage (bvht_batch_*):
 * refit of every distinct BLAS, rebuild of every TLAS, and one persistent
 * trace of all scenes' rays.  The scenes must live on ctx and outlive the
 * batch. */
class Batch {
public:
    Batch(std::shared_ptr<Context> ctx, const std::vector<const Scene*>& scenes) : ctx_(std::move(ctx))
    {
        std::vector<bvht_tlas> handles;
        for (const Scene* sc : scenes)
            handles.push_back(sc->handle());
        check(bvht_batch_create(ctx_->handle(), handles.data(), (uint32_t)handles.size(), &batch_));
        n_scenes_ = (uint32_t)handles.size();
    }
    ~Batch()
    {
        if (batch_)
            bvht_batch_destroy(batch_);
    }
    Batch(const Batch&) = delete;
    Batch& operator=(const Batch&) = delete;

    void refit() { check(bvht_batch_refit(ctx_->handle(), batch_)); }
    void rebuild() { check(bvht_batch_rebuild(ctx_->handle(), batch_)); }
    /* Scene s owns rays [offsets[s], offsets[s+1]) of the ray buffer. */
    void set_rays(const std::vector<uint32_t>& offsets, uint32_t rays_per_item = 128)
    {
        if (offsets.size() != n_scenes_ + 1)
            throw Error(BVHT_EINVAL, "set_rays: one offset per scene plus the end");
        check(bvht_batch_set_rays(ctx_->handle(), batch_, offsets.data(), rays_per_item));
        n_rays_ = offsets.back();
    }
    void config(int32_t wg, int32_t lds_depth, int32_t rays_per_item)
    {
        check(bvht_batch_config(ctx_->handle(), batch_, wg, lds_depth, rays_per_item));
    }
    /* Every scene's rays in one persistent launch (device buffers). */
    void trace(const DeviceBuffer<bvht_ray>& rays, DeviceBuffer<bvht_hit>& hits)
    {
        check(bvht_batch_trace(ctx_->handle(), batch_, rays.data(), hits.data()));
    }
    /* Host convenience: upload, trace, download. */
    std::vector<bvht_hit> trace(const std::vector<Ray>& rays)
    {
        if (rays.size() != n_rays_)
            throw Error(BVHT_EINVAL, "trace: the ray count differs from set_rays");
        std::vector<bvht_ray> recs;
        recs.reserve(rays.size());
        for (const Ray& r : rays)
            recs.push_back(r.record());
        DeviceBuffer<bvht_ray> d_rays(ctx_, recs);
        DeviceBuffer<bvht_hit> d_hits(ctx_, rays.size());
        trace(d_rays, d_hits);
        return d_hits.download(rays.size());
    }

private:
    std::shared_ptr<Context> ctx_;
    bvht_batch batch_ = nullptr;
    uint32_t n_scenes_ = 0, n_rays_ = 0;
};

/* ---- differentiable depth/shading (no reference counterpart) ------------ */

/* An indexed triangle mesh (vertices float[3 nv], faces uint32[3 nf]) bound
 * to the BLAS built from its face soup (bvht_mesh_*): forward trace + shade,
 * the adjoint of the depth/shading L2 loss, a gradient step and the refit. */
class DiffMesh {
public:
    DiffMesh(const Bvh& bvh, const std::vector<float>& vertices, const std::vector<uint32_t>& faces)
        : ctx_(bvh.context())
    {
        check(bvht_mesh_create(ctx_->handle(), bvh.handle(), vertices.data(), (uint32_t)(vertices.size() / 3),
                               faces.data(), (uint32_t)(faces.size() / 3), nullptr, &mesh_));
        n_vertices_ = (uint32_t)(vertices.size() / 3);
    }
    ~DiffMesh()
    {
        if (mesh_)
            bvht_mesh_destroy(mesh_);
    }
    DiffMesh(const DiffMesh&) = delete;
    DiffMesh& operator=(const DiffMesh&) = delete;
    uint32_t n_vertices() const { return n_vertices_; }

    /* hits[n] as a one-object TLAS trace; image[2n] = {depth, shade}. */
    void forward(const Matrix4x4& minv, const Vector3& light, const DeviceBuffer<bvht_ray>& rays, uint32_t n,
                 DeviceBuffer<bvht_hit>& hits, DeviceBuffer<float>& image) const
    {
        float l[3] = {light.x, light.y, light.z};
        check(bvht_trace_forward(ctx_->handle(), mesh_, minv.m.data(), l, rays.data(), n, hits.data(), image.data()));
    }
    /* dLoss/dvertex (float[3 nv]), dLoss/dM^-1 (float[16]) and the loss
     * (one double) against target[2n]. */
    void adjoint(const Matrix4x4& minv, const Vector3& light, const DeviceBuffer<bvht_ray>& rays, uint32_t n,
                 const DeviceBuffer<bvht_hit>& hits, const DeviceBuffer<float>& target, float w_depth, float w_shade,
                 bool accumulate, DeviceBuffer<float>& grad_vertices, DeviceBuffer<float>& grad_minv,
                 DeviceBuffer<double>& loss) const
    {
        float l[3] = {light.x, light.y, light.z};
        check(bvht_trace_adjoint(ctx_->handle(), mesh_, minv.m.data(), l, rays.data(), n, hits.data(), target.data(),
                                 w_depth, w_shade, accumulate ? 1 : 0, grad_vertices.data(), grad_minv.data(),
                                 loss.data()));
    }
    /* v -= lr g_v (each step clipped to max_step when max_step > 0). */
    void sgd(const DeviceBuffer<float>& grad_vertices, float lr, float max_step, int32_t flags = 0)
    {
        check(bvht_mesh_sgd(ctx_->handle(), mesh_, grad_vertices.data(), lr, max_step, flags));
    }
    /* Gather the vertices into the BLAS, refit it, and recompute normals. */
    void update(bool normals = true) { check(bvht_mesh_update(ctx_->handle(), mesh_, normals ? 1 : 0)); }

private:
    std::shared_ptr<Context> ctx_;
    bvht_mesh mesh_ = nullptr;
    uint32_t n_vertices_ = 0;
};

/* ---- camera and renderer (camera/camera.rs, renderer.rs) --------------- */

/* Pinhole camera as the frame path uses it: eye-space viewport corners at the
 * near plane and the eye-to-world matrix (Camera::get_ray_world,
 * camera.rs:994-1010). */
class Camera {
public:
    Camera(const Vector3& top_left, const Vector3& top_right, const Vector3& bottom_left, const Matrix4x4& view_inv)
    {
        set3(rec_.tl, top_left);
        set3(rec_.tr, top_right);
        set3(rec_.bl, bottom_left);
        std::memcpy(rec_.view_inv, view_inv.m.data(), sizeof(rec_.view_inv));
    }
    /* Camera::new(&BoxSpec, &CameraAttitudeSpec) (camera.rs): the frustum's
     * near-plane corners, and [right | up | -forward | eye] as eye-to-world. */
    static Camera from_box_spec(float left, float right, float bottom, float top, float near_, const Vector3& eye,
                                const Vector3& forward, const Vector3& right_axis, const Vector3& up_axis)
    {
        Matrix4x4 m = Matrix4x4::identity();
        const Vector3 cols[4] = {right_axis, up_axis, Vector3() - forward, eye};
        for (int c = 0; c < 4; ++c)
            m.m[c * 4 + 0] = cols[c].x, m.m[c * 4 + 1] = cols[c].y, m.m[c * 4 + 2] = cols[c].z;
        return Camera(Vector3(left, top, -near_), Vector3(right, top, -near_), Vector3(left, bottom, -near_), m);
    }
    const bvht_camera& record() const { return rec_; }

private:
    static void set3(float* d, const Vector3& v) { d[0] = v.x, d[1] = v.y, d[2] = v.z; }
    bvht_camera rec_{};
};

/* Accumulators (renderer.rs:134-334): what PathTracer stores per pixel. */
struct Accumulator {
    int32_t kind;
    bvht_accum_params params{};
};
struct DepthAccumulator : Accumulator {
    DepthAccumulator() : Accumulator{BVHT_ACC_DEPTH} {}
};
struct IntersectionAccumulator : Accumulator {
    IntersectionAccumulator(const Vector3& hit_value, const Vector3& miss_value) : Accumulator{BVHT_ACC_INTERSECTION}
    {
        params.hit_value[0] = hit_value.x, params.hit_value[1] = hit_value.y, params.hit_value[2] = hit_value.z;
        params.miss_value[0] = miss_value.x, params.miss_value[1] = miss_value.y, params.miss_value[2] = miss_value.z;
    }
};
struct UvMappingAccumulator : Accumulator {
    UvMappingAccumulator() : Accumulator{BVHT_ACC_UV} {}
};
/* renderer.rs:248-290: the hit triangle's interpolated corner normal carried
 * to world space (needs the models' normals). */
struct NormalMappingAccumulator : Accumulator {
    NormalMappingAccumulator() : Accumulator{BVHT_ACC_NORMAL} {}
};
/* renderer.rs:293-334: the texel at the interpolated texture coordinates
 * (needs the models' texture coordinates and texture). */
struct TextureMaterialAccumulator : Accumulator {
    TextureMaterialAccumulator() : Accumulator{BVHT_ACC_TEXTURE} {}
};
inline bool is_attribute_accumulator(int32_t kind) { return kind == BVHT_ACC_NORMAL || kind == BVHT_ACC_TEXTURE; }

/* Pixel shaders (renderer.rs:113-222): accumulation buffer -> RGBA8. */
struct PixelShader {
    int32_t kind;
    bvht_shader_params params{};
};
struct RadianceToRgbShader : PixelShader {
    RadianceToRgbShader() : PixelShader{BVHT_SHADE_RADIANCE_TO_RGB} {}
};
struct IntersectionShader : PixelShader {
    IntersectionShader(std::array<uint8_t, 4> hit_rgba, std::array<uint8_t, 4> miss_rgba)
        : PixelShader{BVHT_SHADE_INTERSECTION}
    {
        std::memcpy(params.hit_rgba, hit_rgba.data(), 4);
        std::memcpy(params.miss_rgba, miss_rgba.data(), 4);
    }
};
struct DepthMappingShader : PixelShader {
    DepthMappingShader(float scale, float offset) : PixelShader{BVHT_SHADE_DEPTH_MAPPING}
    {
        params.scale = scale, params.offset = offset;
    }
};

/* RendererState (renderer.rs:77-101): the accumulator, the pixel shader, and
 * the frame's device buffers (rays, hits, accumulation buffer float[3 W H],
 * RGBA8 frame buffer, row-major). */
class RendererState {
public:
    RendererState(Accumulator accumulator, PixelShader shader, uint32_t width, uint32_t height,
                  std::shared_ptr<Context> ctx = default_context())
        : accumulator(accumulator), pixel_shader(shader), width_(width), height_(height), ctx_(ctx),
          rays_(ctx, (size_t)width * height), hits_(ctx, (size_t)width * height),
          accum_(ctx, (size_t)3 * width * height), frame_(ctx, (size_t)4 * width * height)
    {
        if (width % 8 || height % 8)
            throw Error(BVHT_EINVAL, "the tiled frame needs width and height multiples of 8");
        std::vector<uint8_t> black((size_t)4 * width * height, 0);
        for (size_t i = 3; i < black.size(); i += 4)
            black[i] = 255;  // FrameBuffer::from_fill(.., black)
        frame_.upload(black);
    }
    uint32_t width() const { return width_; }
    uint32_t height() const { return height_; }
    std::vector<float> accumulation_buffer() const { return accum_.download(accum_.size()); }
    std::vector<uint8_t> frame_buffer() const { return frame_.download(frame_.size()); }

    Accumulator accumulator;
    PixelShader pixel_shader;

private:
    friend class PathTracer;
    uint32_t width_, height_;
    std::shared_ptr<Context> ctx_;
    DeviceBuffer<bvht_ray> rays_;
    DeviceBuffer<bvht_hit> hits_;
    DeviceBuffer<float> accum_;
    DeviceBuffer<uint8_t> frame_;
};

/* PathTracer::evaluate (renderer.rs:345-384): one camera ray per pixel in
 * 8x8 tiles (u = x / W, v = y / H), the accumulator at each pixel, then the
 * pixel shader over the frame -- all on the GPU.  Returns the rays traced. */
class PathTracer {
public:
    size_t evaluate(RendererState& st, const Scene& scene, const Camera& camera) const
    {
        bvht_ctx c = st.ctx_->handle();
        bvht_tlas t = scene_handle(scene);
        return run(
            st, camera,
            [&](const bvht_ray* d_rays, uint32_t n, bvht_hit* d_hits) {
                check(bvht_trace_tlas(c, t, d_rays, n, d_hits, nullptr));
            },
            [&](const bvht_hit* d_hits, float* d_accum) {
                check(bvht_accumulate_tlas(c, t, d_hits, st.width_, st.height_, BVHT_ORDER_TILES, st.accumulator.kind,
                                           d_accum));
            });
    }
    size_t evaluate(RendererState& st, const ModelInstance& model, const Camera& camera) const
    {
        bvht_ctx c = st.ctx_->handle();
        bvht_blas b = model.model()->bvh().handle();
        return run(
            st, camera,
            [&](const bvht_ray* d_rays, uint32_t n, bvht_hit* d_hits) {
                check(bvht_trace_blas(c, b, d_rays, n, d_hits, nullptr));
            },
            [&](const bvht_hit* d_hits, float* d_accum) {
                check(bvht_accumulate_blas(c, b, d_hits, st.width_, st.height_, BVHT_ORDER_TILES, st.accumulator.kind,
                                           d_accum));
            });
    }

private:
    static bvht_tlas scene_handle(const Scene& scene);
    template <class Trace, class AccumulateAttributes>
    size_t run(RendererState& st, const Camera& camera, Trace trace, AccumulateAttributes accumulate_attributes) const
    {
        bvht_ctx c = st.ctx_->handle();
        uint32_t n = st.width_ * st.height_;
        check(bvht_raygen(c, &camera.record(), st.width_, st.height_, BVHT_ORDER_TILES, st.rays_.data()));
        trace(st.rays_.data(), n, st.hits_.data());
        if (is_attribute_accumulator(st.accumulator.kind))
            accumulate_attributes(st.hits_.data(), st.accum_.data());
        else
            check(bvht_accumulate(c, st.hits_.data(), st.width_, st.height_, BVHT_ORDER_TILES, st.accumulator.kind,
                                  &st.accumulator.params, st.accum_.data()));
        check(bvht_shade(c, st.accum_.data(), n, st.pixel_shader.kind, &st.pixel_shader.params, st.frame_.data()));
        return n;
    }
};

/* An attribute accumulator (normal mapping or texture material) over hits of
 * the scene given in ray order (order BVHT_ORDER_ROWS: hit i is pixel i of a
 * width x height image); returns the accumulation buffer float[3 W H]. */
inline std::vector<float> accumulate_attributes(const Scene& scene, const std::vector<bvht_hit>& hits,
                                                uint32_t width, uint32_t height, const Accumulator& acc,
                                                std::shared_ptr<Context> ctx = default_context())
{
    if (!is_attribute_accumulator(acc.kind) || hits.size() != (size_t)width * height)
        throw Error(BVHT_EINVAL, "accumulate_attributes: an attribute accumulator and width * height hits");
    DeviceBuffer<bvht_hit> d_hits(ctx, hits);
    DeviceBuffer<float> d_acc(ctx, (size_t)3 * width * height);
    check(bvht_accumulate_tlas(ctx->handle(), scene.handle(), d_hits.data(), width, height, BVHT_ORDER_ROWS, acc.kind,
                               d_acc.data()));
    return d_acc.download(d_acc.size());
}

/* Renderer (renderer.rs:386-398). */
class Renderer {
public:
    explicit Renderer(PathTracer integrator) : integrator_(integrator) {}
    template <class SceneT>
    size_t render(RendererState& st, const SceneT& scene, const Camera& camera) const
    {
        return integrator_.evaluate(st, scene, camera);
    }

private:
    PathTracer integrator_;
};

inline bvht_tlas PathTracer::scene_handle(const Scene& scene) { return scene.handle(); }

/* ---- rigid-body physics (physics/rigid_body.rs, force_generator.rs,
 *      world.rs) -------------------------------------------------------- */

/* RigidBody (rigid_body.rs:66-463): the reference's record field for field
 * (bvht_body), with its setters doing their binary32 arithmetic in the
 * reference's order.  RigidBody() is RigidBody::default() (:88-112). */
class RigidBody {
public:
    RigidBody()
    {
        std::memset(&rec_, 0, sizeof(rec_));
        rec_.inverse_mass = FLT_MAX;
        for (int k = 0; k < 3; ++k)
            rec_.inverse_inertia_tensor[4 * k] = rec_.inverse_inertia_tensor_world[4 * k] = 1.0f;
        rec_.linear_drag = rec_.angular_drag = 1.0f;
        rec_.orientation[3] = 1.0f;  // Quaternion::unit_z, (s, x, y, z) = (0, 0, 0, 1)
        rec_.can_sleep = 1;
        for (int k = 0; k < 4; ++k)
            rec_.transform[5 * k] = 1.0f;
    }
    explicit RigidBody(const bvht_body& rec) : rec_(rec) {}
    const bvht_body& record() const { return rec_; }

    void set_mass(float mass)  // :205-212
    {
        if (mass == 0.0f)
            throw Error(BVHT_EINVAL, "mass must be non-zero");
        rec_.inverse_mass = 1.0f / mass;
    }
    void set_inverse_mass(float inverse_mass) { rec_.inverse_mass = inverse_mass; }
    float get_inverse_mass() const { return rec_.inverse_mass; }
    bool has_finite_mass() const { return rec_.inverse_mass >= 0.0f; }
    void set_drag(float linear_drag, float angular_drag) { rec_.linear_drag = linear_drag, rec_.angular_drag = angular_drag; }
    void set_position(const Vector3& p) { set3(rec_.position, p); }
    Vector3 get_position() const { return get3(rec_.position); }
    /* Stores orientation.normalize(): q / sqrt(s*s + ((x*x + y*y) + z*z)). */
    void set_orientation(float s, float x, float y, float z)
    {
        float mag = std::sqrt(s * s + ((x * x + y * y) + z * z));
        rec_.orientation[0] = s / mag, rec_.orientation[1] = x / mag;
        rec_.orientation[2] = y / mag, rec_.orientation[3] = z / mag;
    }
    void set_velocity(const Vector3& v) { set3(rec_.velocity, v); }
    Vector3 get_velocity() const { return get3(rec_.velocity); }
    void set_rotation(const Vector3& r) { set3(rec_.rotation, r); }
    Vector3 get_rotation() const { return get3(rec_.rotation); }
    void set_acceleration(const Vector3& a) { set3(rec_.acceleration, a); }
    bool get_awake() const { return rec_.is_awake != 0; }
    void set_awake(bool awake)  // :386-398
    {
        if (awake) {
            rec_.is_awake = 1;
            rec_.motion = 0.3f * 2.0f;
        } else {
            rec_.is_awake = 0;
            set3(rec_.velocity, Vector3());
            set3(rec_.rotation, Vector3());
        }
    }
    void set_can_sleep(bool can_sleep)  // :404-410
    {
        rec_.can_sleep = can_sleep ? 1 : 0;
        if (!can_sleep && !get_awake())
            set_awake(true);
    }
    void apply_force(const Vector3& f)  // :428-431
    {
        set3(rec_.force_accumulator, get3(rec_.force_accumulator) + f);
        rec_.is_awake = 1;
    }
    /* Body to world, column-major. */
    Matrix4x4 get_transform() const
    {
        Matrix4x4 m;
        std::memcpy(m.m.data(), rec_.transform, sizeof(rec_.transform));
        return m;
    }

private:
    static void set3(float* d, const Vector3& v) { d[0] = v.x, d[1] = v.y, d[2] = v.z; }
    static Vector3 get3(const float* d) { return Vector3(d[0], d[1], d[2]); }
    bvht_body rec_;
};

/* Force generators (force_generator.rs): records of the force registry. */
struct ForceGenerator {
    bvht_force rec{};
};
struct NullForce : ForceGenerator {  // :9-31
    NullForce() { rec.kind = BVHT_FORCE_NULL; }
};
struct Gravity : ForceGenerator {  // :34-56
    explicit Gravity(const Vector3& g)
    {
        rec.kind = BVHT_FORCE_GRAVITY;
        rec.v0[0] = g.x, rec.v0[1] = g.y, rec.v0[2] = g.z;
    }
};

class World;

/* The shared handle of a registered body (force_registry.rs:15-28). */
class RigidBodyInstance {
public:
    RigidBodyInstance(World* world, uint32_t index) : world_(world), index_(index) {}
    uint32_t index() const { return index_; }
    RigidBody borrow() const;
    /* RefCell::borrow_mut: edit a copy of the body, written back after fn. */
    template <class Fn>
    void borrow_mut(Fn fn) const;

private:
    World* world_;
    uint32_t index_;
};

struct Spring : ForceGenerator {  // :58-121: between the registered body and `other`
    Spring(const RigidBodyInstance& other, const Vector3& connection_point_body,
           const Vector3& other_connection_point_body, float spring_constant, float rest_length)
    {
        rec.kind = BVHT_FORCE_SPRING;
        rec.other = other.index();
        rec.v0[0] = connection_point_body.x, rec.v0[1] = connection_point_body.y, rec.v0[2] = connection_point_body.z;
        rec.v1[0] = other_connection_point_body.x, rec.v1[1] = other_connection_point_body.y;
        rec.v1[2] = other_connection_point_body.z;
        rec.spring_constant = spring_constant, rec.rest_length = rest_length;
    }
};

/* World (world.rs:14-61).  Bodies and forces are registered on the host;
 * the first start_frame / run_physics moves them to the device, where the
 * registry is applied and the bodies integrated. */
class World {
public:
    explicit World(std::shared_ptr<Context> ctx = default_context()) : ctx_(std::move(ctx)) {}
    ~World()
    {
        if (world_)
            bvht_world_destroy(ctx_->handle(), world_);
    }
    World(const World&) = delete;
    World& operator=(const World&) = delete;

    RigidBodyInstance register_body(const RigidBody& body)
    {
        if (world_)
            throw Error(BVHT_EINVAL, "bodies are registered before the world first runs");
        bodies_.push_back(body.record());
        return RigidBodyInstance(this, (uint32_t)bodies_.size() - 1);
    }
    void register_force_generator(const RigidBodyInstance& body, const ForceGenerator& generator)
    {
        if (world_)
            throw Error(BVHT_EINVAL, "force generators are registered before the world first runs");
        bvht_force f = generator.rec;
        f.body = body.index();
        forces_.push_back(f);
    }
    /* World::start_frame (world.rs:44-49). */
    void start_frame() { check(bvht_world_start_frame(ctx_->handle(), device())); }
    /* World::run_physics (world.rs:57-60). */
    void run_physics(float duration) { check(bvht_world_run_physics(ctx_->handle(), device(), duration)); }
    size_t len() const { return bodies_.size(); }

    RigidBody body(uint32_t i) const
    {
        if (!world_)
            return RigidBody(bodies_.at(i));
        bvht_body rec;
        check(bvht_world_get_bodies(ctx_->handle(), world_, &rec, i, 1));
        return RigidBody(rec);
    }
    void set_body(uint32_t i, const RigidBody& body)
    {
        if (!world_) {
            bodies_.at(i) = body.record();
            return;
        }
        check(bvht_world_set_bodies(ctx_->handle(), world_, &body.record(), i, 1));
    }

private:
    bvht_world device()
    {
        if (!world_)
            check(bvht_world_create(ctx_->handle(), bodies_.data(), (uint32_t)bodies_.size(),
                                    forces_.empty() ? nullptr : forces_.data(), (uint32_t)forces_.size(), &world_));
        return world_;
    }

    std::shared_ptr<Context> ctx_;
    std::vector<bvht_body> bodies_;
    std::vector<bvht_force> forces_;
    bvht_world world_ = nullptr;
};

inline RigidBody RigidBodyInstance::borrow() const { return world_->body(index_); }
template <class Fn>
void RigidBodyInstance::borrow_mut(Fn fn) const
{
    RigidBody b = world_->body(index_);
    fn(b);
    world_->set_body(index_, b);
}

}  // namespace bvhtracer

#endif  // BVHTRACER_HPP
