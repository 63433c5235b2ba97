// The reference crate's own tests for the traversal path, written against the
// C++ host API (include/bvhtracer.hpp) the way the reference writes them in
// Rust, and run on the GPU.  Each test names the reference test it restates
// (bvhtracer/tests/*.rs, relative to the reference repository root).
// Prints "PASS <name>" / "FAIL <name>: <why>" per test; exit status = number
// of failures.
#include <cmath>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

#include <fstream>

#include "bvhtracer.hpp"
#include "bvhtracer_mesh.hpp"

using namespace bvhtracer;

static std::string g_golden = "tests/golden";  // the reference's assets (argv[1] overrides)

namespace {

struct Case {
    const char* name;
    std::function<void()> fn;
};
std::vector<Case>& cases()
{
    static std::vector<Case> c;
    return c;
}
struct Register {
    Register(const char* name, std::function<void()> fn) { cases().push_back({name, std::move(fn)}); }
};
struct Failure {
    std::string what;
};
#define TEST(name)                                 \
    static void name();                            \
    static Register reg_##name(#name, name);       \
    static void name()
#define CHECK(cond)                                                                      \
    do {                                                                                 \
        if (!(cond))                                                                     \
            throw Failure{std::string(__FILE__ ":") + std::to_string(__LINE__) + ": " #cond}; \
    } while (0)

const float kSqrt3 = std::sqrt(3.0f);

// bvhtracer/tests/test_triangle_intersection.rs:11-17 (and the meshes built from it)
Triangle unit_triangle()
{
    return Triangle(Vector3(0.0f, 1.0f / 2.0f, 0.0f), Vector3(-1.0f / kSqrt3, -1.0f / 2.0f, 0.0f),
                    Vector3(1.0f / kSqrt3, -1.0f / 2.0f, 0.0f));
}

Triangle displaced(const Triangle& t, const Vector3& d)
{
    return Triangle(t.vertices[0] + d, t.vertices[1] + d, t.vertices[2] + d);
}

ModelInstance model_of(const std::vector<Triangle>& tris)
{
    MeshBuilder builder;
    for (const Triangle& t : tris)
        builder.with_primitive(t, TextureCoordinates{}, Normals{});
    return ModelBuilder().with_mesh(builder.build()).build();
}

Ray ray_towards(const Vector3& origin, const Vector3& target, float t = FLT_MAX)
{
    return Ray(origin, (target - origin).normalize(), t);
}

}  // namespace

// ---- test_triangle_intersection.rs ----------------------------------------

TEST(test_triangle_intersection_hits)  // :19-29
{
    CHECK(unit_triangle().intersect(ray_towards({0, 0, 5}, Vector3::zero())).has_value());
}

TEST(test_triangle_intersection_vertex_hits)  // :31-65
{
    Triangle tri = unit_triangle();
    for (const Vector3& v : tri.vertices)
        CHECK(tri.intersect(ray_towards({0, 0, 5}, v)).has_value());
}

TEST(test_triangle_intersection_center)  // :67-80
{
    Ray ray = ray_towards({0, 0, 5}, Vector3::zero());
    float t = unit_triangle().intersect(ray)->t;
    CHECK(Ray(ray.origin, ray.direction, t) == Ray(ray.origin, ray.direction, 5.0f));
}

TEST(test_triangle_intersection_vertex1)  // :82-95
{
    Triangle tri = unit_triangle();
    CHECK(tri.intersect(ray_towards({0, 0, 5}, tri.vertices[0]))->t == std::sqrt(101.0f / 4.0f));
}

TEST(test_triangle_intersection_vertex2_and_3)  // :97-125
{
    Triangle tri = unit_triangle();
    CHECK(tri.intersect(ray_towards({0, 0, 5}, tri.vertices[1]))->t == std::sqrt(307.0f / 12.0f));
    CHECK(tri.intersect(ray_towards({0, 0, 5}, tri.vertices[2]))->t == std::sqrt(307.0f / 12.0f));
}

TEST(test_triangle_intersection_vertex_misses)  // :127-167
{
    Triangle tri = unit_triangle();
    CHECK(!tri.intersect(ray_towards({0, 0, 5}, tri.vertices[0] + Vector3(0, 0.5f, 0))).has_value());
    CHECK(!tri.intersect(ray_towards({0, 0, 5}, tri.vertices[1] + Vector3(0, -0.5f, 0))).has_value());
    CHECK(!tri.intersect(ray_towards({0, 0, 5}, tri.vertices[2] + Vector3(0, -0.5f, 0))).has_value());
}

TEST(test_triangle_closest_t)  // :169-207
{
    for (float ray_t : {FLT_MAX, 100.0f, 0.01f}) {
        auto r = unit_triangle().intersect(ray_towards({0, 0, 5}, Vector3::zero(), ray_t));
        CHECK(r.has_value() && r->t <= ray_t);
    }
}

TEST(test_triangle_intersect_mut_updates_ray)  // triangle.rs:74-108
{
    Intersection it;
    it.ray = ray_towards({0, 0, 5}, Vector3::zero());
    CHECK(unit_triangle().intersect_mut(it));
    CHECK(it.ray.t == 5.0f && it.interaction.t == 5.0f);
}

// ---- test_bvh_one_triangle.rs ----------------------------------------------

TEST(test_one_triangle_intersection_hits)  // :31-41
{
    CHECK(model_of({unit_triangle()}).intersect(ray_towards({0, 0, 5}, Vector3::zero())).has_value());
}

TEST(test_one_triangle_intersection_vertex_hits)  // :43-79
{
    ModelInstance scene = model_of({unit_triangle()});
    for (const Vector3& v : unit_triangle().vertices)
        CHECK(scene.intersect(ray_towards({0, 0, 5}, v)).has_value());
}

TEST(test_one_triangle_exact_t)  // :81-139
{
    ModelInstance scene = model_of({unit_triangle()});
    Triangle tri = unit_triangle();
    CHECK(scene.intersect(ray_towards({0, 0, 5}, Vector3::zero()))->interaction.t == 5.0f);
    CHECK(scene.intersect(ray_towards({0, 0, 5}, tri.vertices[0]))->interaction.t == std::sqrt(101.0f / 4.0f));
    CHECK(scene.intersect(ray_towards({0, 0, 5}, tri.vertices[1]))->interaction.t == std::sqrt(307.0f / 12.0f));
}

// ---- test_bvh_two_triangles.rs / test_bvh_three_triangles.rs ----------------

TEST(test_two_triangles_intersection_hits_and_miss)  // two_triangles.rs:16-82
{
    Vector3 d(5, 0, 0);
    ModelInstance scene = model_of({displaced(unit_triangle(), Vector3() - d), displaced(unit_triangle(), d)});
    for (const Triangle& t : scene.model()->primitives())
        CHECK(scene.intersect(ray_towards({0, 0, 5}, t.centroid())).has_value());
    CHECK(!scene.intersect(ray_towards({0, 0, 5}, Vector3::zero())).has_value());
}

TEST(test_three_triangles_centroids_hit)  // three_triangles.rs:16-60
{
    Vector3 d(0, 10, 0);
    ModelInstance scene = model_of({unit_triangle(), displaced(unit_triangle(), Vector3() - d),
                                    displaced(unit_triangle(), d)});
    for (const Triangle& t : scene.model()->primitives()) {
        Ray ray = ray_towards({0, 0, 10}, t.centroid());
        CHECK(t.intersect(ray).has_value());
        CHECK(scene.intersect(ray).has_value());
    }
}

// ---- test_bvh_closest_intersection.rs ----------------------------------------

static ModelInstance stack_scene()  // :25-43: 100 triangles stacked along -z
{
    std::vector<Triangle> tris;
    for (int i = 0; i < 100; ++i)
        tris.push_back(displaced(unit_triangle(), Vector3() - Vector3(0, 0, (float)i)));
    return model_of(tris);
}

TEST(test_scene_stack_xy)  // :45-56
{
    ModelInstance scene = stack_scene();
    Triangle top = unit_triangle();
    for (const Triangle& t : scene.model()->primitives())
        for (int k = 0; k < 3; ++k)
            CHECK(t.vertices[k].x == top.vertices[k].x && t.vertices[k].y == top.vertices[k].y);
}

TEST(test_intersection_should_return_closest_point_triangle)  // :58-77
{
    ModelInstance scene = stack_scene();
    Triangle closest = unit_triangle();
    Ray ray = ray_towards({0, 0, 5}, closest.centroid());
    auto result = scene.intersect(ray);
    auto expected = closest.intersect(ray);
    CHECK(result.has_value() && expected.has_value());
    CHECK(result->interaction.t == expected->t);
}

TEST(test_intersection_all_primitives_along_ray_should_hit)  // :79-93
{
    ModelInstance scene = stack_scene();
    Vector3 c = unit_triangle().centroid();
    Ray ray = ray_towards({c.x, c.y, 5}, c);
    const auto& prims = scene.model()->primitives();
    auto hits = intersect_pairs(prims, std::vector<Ray>(prims.size(), ray));
    for (const bvht_hit& h : hits)
        CHECK(h.prim != BVHT_MISS);
}

TEST(test_context_options_keep_hits)  // no reference counterpart: trace configuration, wave order
{
    ModelInstance scene = stack_scene();
    const auto& prims = scene.model()->primitives();
    std::vector<Ray> rays;
    for (int k = 0; k < 300; ++k)
        rays.push_back(ray_towards({0.01f * (float)(k % 17) - 0.08f, 0.01f * (float)(k / 17) - 0.08f, 5},
                                   unit_triangle().centroid()));
    const auto& bvh = scene.model()->bvh();
    auto want = bvh.intersect_batch(rays);
    const auto& ctx = bvh.context();
    ctx->set_trace_config(128, 8);
    ctx->set_wave_schedule(true);
    for (int rep = 0; rep < 3; ++rep) {
        auto got = bvh.intersect_batch(rays);
        for (size_t i = 0; i < rays.size(); ++i)
            CHECK(got[i].prim == want[i].prim && got[i].t == want[i].t);
    }
    ctx->set_wave_schedule(false);
    ctx->set_trace_config(0, 0);
    CHECK(prims.size() > 0);
}

TEST(test_intersection_closest_point_should_have_lowest_t)  // :95-119
{
    ModelInstance scene = stack_scene();
    Triangle closest = unit_triangle();
    Vector3 c = closest.centroid();
    Ray ray = ray_towards({c.x, c.y, 5}, c);
    SurfaceInteraction expected = *closest.intersect(ray);
    const auto& prims = scene.model()->primitives();
    auto hits = intersect_pairs(prims, std::vector<Ray>(prims.size(), ray));
    SurfaceInteraction result(FLT_MAX, FLT_MAX, FLT_MAX);
    for (const bvht_hit& h : hits) {
        SurfaceInteraction s(h.t, h.u, h.v);
        if (!(s.t < result.t))
            continue;
        result = s;
    }
    CHECK(result == expected);
    CHECK(scene.intersect(ray)->interaction == expected);
}

// ---- test_bvh_refit.rs -------------------------------------------------------

static std::vector<Triangle> refit_mesh()  // :11-27
{
    Vector3 dx(5, 0, 0), dy(0, 5, 0);
    std::vector<Triangle> mesh;
    for (int i = -100, j = -100; i < 100; ++i, ++j)
        mesh.push_back(displaced(unit_triangle(), dx * (float)i + dy * (float)j));
    return mesh;
}

static void animate(std::vector<Triangle>& mesh)  // :29-40
{
    for (Triangle& t : mesh) {
        t.vertices[0] += Vector3(0.3f, 0, 0);
        t.vertices[1] += Vector3(0, 0.3f, 0);
        t.vertices[2] += Vector3(0, 0, 0.3f);
    }
}

TEST(test_bvh_refit_unchanged_on_same_mesh)  // :42-49
{
    std::vector<Triangle> mesh = refit_mesh();
    std::vector<Triangle> copy = mesh;
    Bvh result = BvhBuilder().build_for(mesh);
    Bvh expected = BvhBuilder().build_for(copy);  // the reference clones the built Bvh
    CHECK(mesh == copy);
    result.refit(mesh);
    CHECK(result == expected);
}

TEST(test_bvh_refit_idemotent)  // :51-61
{
    std::vector<Triangle> mesh = refit_mesh();
    Bvh bvh = BvhBuilder().build_for(mesh);
    animate(mesh);
    bvh.refit(mesh);
    std::vector<bvht_node> expected = bvh.nodes();
    bvh.refit(mesh);
    std::vector<bvht_node> result = bvh.nodes();
    CHECK(result.size() == expected.size() &&
          std::memcmp(result.data(), expected.data(), result.size() * sizeof(bvht_node)) == 0);
    // refit moved the root bounds by the animation's displacement
    Aabb b = bvh.bounds();
    CHECK(b.bounds_max.z == 0.3f);
}

// ---- the reference's panics become errors -----------------------------------

TEST(test_empty_mesh_is_an_error)  // bvh.rs:530-535 panics on an empty mesh
{
    std::vector<Triangle> empty;
    bool threw = false;
    try {
        BvhBuilder().build_for(empty);
    } catch (const Error& e) {
        threw = e.code() == BVHT_EINVAL;
    }
    CHECK(threw);
}

// ---- scene: two instances of one model (scene.rs, tlas.rs) ------------------

TEST(test_scene_two_instances_nearest_and_object_index)
{
    ModelInstance model = model_of({unit_triangle()});
    SceneBuilder builder;
    builder.with_object(SceneObjectBuilder(model).with_transform(Matrix4x4::from_affine_translation({0, 0, -2})).build());
    builder.with_object(SceneObjectBuilder(model).with_transform(Matrix4x4::from_affine_translation({0, 0, 1})).build());
    auto scene = builder.build();
    // translations are exact, so the object-space rays are (0, 0, 4) and
    // (0, 0, 7) looking down -z, and the scene's t is the triangle's own
    Vector3 down(0, 0, -1);
    auto hit = scene->intersect(Ray::from_origin_dir({0, 0, 5}, down));
    CHECK(hit.has_value());
    CHECK(hit->object == 1 && hit->interaction.t == unit_triangle().intersect(Ray::from_origin_dir({0, 0, 4}, down))->t);
    // move the near object out of the way: the far one is hit after a rebuild
    scene->get_mut_unchecked(1).set_transform(Matrix4x4::from_affine_translation({10, 0, 1}));
    scene->rebuild();
    hit = scene->intersect(Ray::from_origin_dir({0, 0, 5}, down));
    CHECK(hit.has_value() && hit->object == 0 &&
          hit->interaction.t == unit_triangle().intersect(Ray::from_origin_dir({0, 0, 7}, down))->t);
}

TEST(test_scene_object_intersect_and_bounds)  // scene_object.rs:77-89, tlas.rs:258-279
{
    ModelInstance model = model_of({unit_triangle()});
    SceneObject near = SceneObjectBuilder(model).with_transform(Matrix4x4::from_affine_translation({0, 0, 1})).build();
    SceneObject far = SceneObjectBuilder(model).with_transform(Matrix4x4::from_affine_translation({0, 0, -2})).build();
    Vector3 down(0, 0, -1);
    // each object alone: its own hit, in world space
    CHECK(near.intersect(Ray::from_origin_dir({0, 0, 5}, down))->interaction.t ==
          unit_triangle().intersect(Ray::from_origin_dir({0, 0, 4}, down))->t);
    CHECK(far.intersect(Ray::from_origin_dir({0, 0, 5}, down))->interaction.t ==
          unit_triangle().intersect(Ray::from_origin_dir({0, 0, 7}, down))->t);
    // world bounds = the model's bounds translated
    Aabb mb = model.bounds(), b = near.bounds();
    CHECK(b.bounds_min == mb.bounds_min + Vector3(0, 0, 1) && b.bounds_max == mb.bounds_max + Vector3(0, 0, 1));
    near.set_transform(Matrix4x4::from_affine_translation({0, 0, 3}));
    CHECK(near.bounds().bounds_min.z == 3.0f);
    // TlasBuilder over both; the nearest of the two is hit
    std::vector<SceneObject> objects{near, far};
    auto tlas = TlasBuilder().build_for(objects);
    uint32_t used = 0;
    tlas->nodes(&used);
    CHECK(used == 4);  // two leaves, their merged node, and the root that copies it (tlas.rs:204-250)
    auto hit = tlas->intersect(objects, Ray::from_origin_dir({0, 0, 5}, down));
    CHECK(hit.has_value() && hit->object == 0);
}

// ---- this library's batched trace and forward/adjoint pair -----------------

static std::vector<Ray> fan_rays(size_t n, const Vector3& origin, float spread, uint32_t seed)
{
    std::vector<Ray> rays;
    uint32_t s = seed;
    auto next = [&]() {
        s = s * 1664525u + 1013904223u;
        return (float)(s >> 8) * (1.0f / 16777216.0f);
    };
    for (size_t i = 0; i < n; ++i) {
        Vector3 target((next() - 0.5f) * spread, (next() - 0.5f) * spread, 0.0f);
        rays.push_back(Ray::from_origin_dir(origin, (target - origin).normalize()));
    }
    return rays;
}

TEST(test_batch_trace_equals_each_scene)
{
    std::vector<Triangle> grid;
    for (int i = -3; i <= 3; ++i)
        for (int j = -3; j <= 3; ++j)
            grid.push_back(displaced(unit_triangle(), Vector3((float)i, (float)j, 0.1f * (float)(i * j))));
    ModelInstance a = model_of(grid), b = model_of({unit_triangle()});
    SceneBuilder sa, sb;
    sa.with_object(SceneObjectBuilder(a).with_transform(Matrix4x4::from_affine_translation({0, 0, -1})).build());
    sb.with_object(SceneObjectBuilder(b).with_transform(Matrix4x4::from_affine_scale(2.0f)).build());
    sb.with_object(SceneObjectBuilder(a).with_transform(Matrix4x4::from_affine_translation({0.5f, 0, -3})).build());
    auto scene_a = sa.build(), scene_b = sb.build();
    std::vector<Ray> ra = fan_rays(300, {0, 0, 8}, 9.0f, 1), rb = fan_rays(500, {0, 0, 8}, 9.0f, 2);
    std::vector<Ray> all = ra;
    all.insert(all.end(), rb.begin(), rb.end());
    Batch batch(default_context(), {scene_a.get(), scene_b.get()});
    batch.set_rays({0, 300, 800}, 64);
    std::vector<bvht_hit> got = batch.trace(all);
    std::vector<bvht_hit> want = scene_a->intersect_batch(ra), wb = scene_b->intersect_batch(rb);
    want.insert(want.end(), wb.begin(), wb.end());
    CHECK(got.size() == want.size() && std::memcmp(got.data(), want.data(), got.size() * sizeof(bvht_hit)) == 0);
    size_t hits = 0;
    for (const bvht_hit& h : got)
        hits += h.prim != BVHT_MISS;
    CHECK(hits > 100);
}

TEST(test_diff_mesh_forward_matches_scene_and_adjoint_loss)
{
    // an octahedron of radius 2, faces in order
    std::vector<float> v = {2, 0, 0, -2, 0, 0, 0, 2, 0, 0, -2, 0, 0, 0, 2, 0, 0, -2};
    std::vector<uint32_t> f = {0, 2, 4, 2, 1, 4, 1, 3, 4, 3, 0, 4, 2, 0, 5, 1, 2, 5, 3, 1, 5, 0, 3, 5};
    std::vector<Triangle> soup;
    for (size_t k = 0; k < f.size(); k += 3)
        soup.emplace_back(Vector3(v[3 * f[k]], v[3 * f[k] + 1], v[3 * f[k] + 2]),
                          Vector3(v[3 * f[k + 1]], v[3 * f[k + 1] + 1], v[3 * f[k + 1] + 2]),
                          Vector3(v[3 * f[k + 2]], v[3 * f[k + 2] + 1], v[3 * f[k + 2] + 2]));
    std::vector<Triangle> soup2 = soup;
    Bvh bvh = BvhBuilder().build_for(soup);
    DiffMesh mesh(bvh, v, f);
    std::vector<Ray> rays = fan_rays(256, {0.3f, 0.2f, 9}, 5.0f, 3);
    std::vector<bvht_ray> recs;
    for (const Ray& r : rays)
        recs.push_back(r.record());
    auto ctx = default_context();
    uint32_t n = (uint32_t)rays.size();
    DeviceBuffer<bvht_ray> d_rays(ctx, recs);
    DeviceBuffer<bvht_hit> d_hits(ctx, n);
    DeviceBuffer<float> d_img(ctx, 2 * n), d_tgt(ctx, 2 * n), g_v(ctx, 3 * mesh.n_vertices()), g_m(ctx, 16);
    DeviceBuffer<double> d_loss(ctx, 1);
    Matrix4x4 eye = Matrix4x4::identity();
    mesh.forward(eye, {0, 0, 1}, d_rays, n, d_hits, d_img);
    std::vector<bvht_hit> hits = d_hits.download(n);
    // the forward's hits are a one-object scene trace of the same soup
    SceneBuilder sb;
    sb.with_object(SceneObjectBuilder(model_of(soup2)).build());
    std::vector<bvht_hit> want = sb.build()->intersect_batch(rays);
    CHECK(std::memcmp(hits.data(), want.data(), n * sizeof(bvht_hit)) == 0);
    // target = image: the adjoint re-solves t in binary64, so what remains is
    // the binary32 rounding of the image (|t - t32| below half an ulp of 8)
    std::vector<float> img = d_img.download(2 * n);
    d_tgt.upload(img);
    mesh.adjoint(eye, {0, 0, 1}, d_rays, n, d_hits, d_tgt, 1.0f, 1.0f, false, g_v, g_m, d_loss);
    CHECK(d_loss.download(1)[0] < 1e-9);
    for (float g : g_v.download(3 * mesh.n_vertices()))
        CHECK(std::fabs(g) < 1e-3f);
    // target depth + 0.5 on every hit pixel: loss = sum of the squared depth gaps
    std::vector<float> tgt = img;
    double expected = 0.0;
    size_t hit = 0;
    for (uint32_t i = 0; i < n; ++i)
        if (hits[i].prim != BVHT_MISS) {
            tgt[2 * i] = img[2 * i] + 0.5f;
            double d = (double)img[2 * i] - (double)tgt[2 * i];
            expected += d * d;
            ++hit;
        }
    CHECK(hit > 50);
    d_tgt.upload(tgt);
    mesh.adjoint(eye, {0, 0, 1}, d_rays, n, d_hits, d_tgt, 1.0f, 0.0f, false, g_v, g_m, d_loss);
    double loss = d_loss.download(1)[0];
    CHECK(std::fabs(loss - expected) <= 1e-4 * expected);
}

// ---- test_scene_cube.rs ------------------------------------------------------

TEST(test_cube_scene_normal_mapping_known_answer)  // :139-175, 230-265
{
    std::ifstream in(g_golden + "/cube.obj");
    CHECK(in.good());
    ModelInstance model = ModelBuilder().with_mesh(ObjMeshDecoder(in).read_mesh()).build();
    Matrix4x4 t = Matrix4x4::from_affine_scale(2.0f);
    t.m[12] = t.m[13] = t.m[14] = -1.0f;  // scaled by 2 and moved by (-1, -1, -1)
    SceneBuilder builder;
    builder.with_object(SceneObjectBuilder(model).with_transform(t).build());
    auto scene = builder.build();
    Vector3 o(0, 4, 0);
    std::vector<Ray> rays = {ray_towards(o, {0.5f, 1, -0.5f}), ray_towards(o, {-0.5f, 1, 0.5f})};
    std::vector<bvht_hit> hits = scene->intersect_batch(rays);
    CHECK(hits[0].prim == 6 && hits[1].prim == 9);  // :124-135, 215-226
    // the top face's world normal is +y at every corner: (0.5, 1, 0.5) for both
    // (with the attributes permuted like the triangles; the reference reads
    // the unpermuted normal array there)
    std::vector<float> acc = accumulate_attributes(*scene, hits, 2, 1, NormalMappingAccumulator());
    CHECK(acc == std::vector<float>({0.5f, 1.0f, 0.5f, 0.5f, 1.0f, 0.5f}));
}

// ---- test_scene_quad.rs (texture and physics of the scene left out) --------

static std::unique_ptr<Scene> quad_scene()  // :52-131
{
    MeshBuilder mesh;
    mesh.with_primitive(Triangle({-1, -1, 0}, {1, 1, 0}, {-1, 1, 0}), {}, {});
    mesh.with_primitive(Triangle({-1, -1, 0}, {1, -1, 0}, {1, 1, 0}), {}, {});
    ModelInstance model = ModelBuilder().with_mesh(mesh.build()).build();
    SceneBuilder builder;
    builder.with_object(SceneObjectBuilder(model).build());
    return builder.build();
}

static Camera quad_camera()  // BoxSpec(-1, 1, -1, 1, 1, 100), eye (0, 0, 2) looking down -z
{
    return Camera::from_box_spec(-1, 1, -1, 1, 1, {0, 0, 2}, {0, 0, -1}, {1, 0, 0}, {0, 1, 0});
}

struct QuadUv {
    float tl_u, tl_v, tr_u, bl_v;
};
static QuadUv quad_test_case()  // :133-198 in binary32
{
    Vector3 o(0, 0, 2);
    float tc = std::sqrt(6.0f) / 2.0f;
    auto plane = [&](const Vector3& corner) { return Ray::from_origin_dir(o, (corner - o).normalize()).interpolate(tc); };
    Vector3 tl(-1, 1, 0), tr(1, 1, 0), bl(-1, -1, 0);
    Vector3 ptl = plane(tl), ptr = plane(tr), pbl = plane(bl);
    float du = (ptr.x - ptl.x) / 2.0f, dv = (ptl.y - pbl.y) / 2.0f;
    QuadUv q;
    q.tl_u = (ptl.x - tl.x) / 2.0f;
    q.tl_v = (tl.y - ptl.y) / 2.0f;
    q.tr_u = q.tl_u + du;
    q.bl_v = q.tl_v + dv;
    return q;
}

TEST(test_quad_depth_centroid)  // :200-209, through the GPU frame: pixel (320, 320) has u = v = 0.5
{
    auto scene = quad_scene();
    RendererState state(DepthAccumulator(), RadianceToRgbShader(), 640, 640);
    CHECK(Renderer(PathTracer()).render(state, *scene, quad_camera()) == 640u * 640u);
    std::vector<float> acc = state.accumulation_buffer();
    size_t px = 320 * 640 + 320;
    CHECK(acc[3 * px] == 2.0f && acc[3 * px + 1] == 2.0f && acc[3 * px + 2] == 2.0f);
}

TEST(test_scene_intersection_entire_viewport)  // :354-378
{
    auto scene = quad_scene();
    QuadUv q = quad_test_case();
    RendererState state(IntersectionAccumulator({1, 1, 1}, {0, 0, 0}),
                        IntersectionShader({255, 255, 255, 255}, {0, 0, 0, 255}), 640, 640);
    PathTracer().evaluate(state, *scene, quad_camera());
    std::vector<uint8_t> frame = state.frame_buffer();
    int on_edge = 0, edge_mismatch = 0;
    for (int y = 0; y < 640; ++y)
        for (int x = 0; x < 640; ++x) {
            float u = (float)x / 640.0f, v = (float)y / 640.0f;
            bool in_u = u >= q.tl_u && u <= q.tr_u, in_v = v >= q.tl_v && v <= q.bl_v;
            bool expected = in_u && in_v;
            bool result = frame[4 * (y * 640 + x)] == 255;
            bool edge = ((u == q.tl_u || u == q.tr_u) && in_v) || ((v == q.tl_v || v == q.bl_v) && in_u);
            if (edge) {
                ++on_edge;
                edge_mismatch += result != expected;
            } else {
                CHECK(result == expected);
            }
        }
    // the rays on the quad's four edge lines land exactly on a triangle edge:
    // their hit bits follow cglinalg's unpinned rounding, and the GPU gives
    // the oracle's answer there (16 of the 1280 differ from the solid angle)
    CHECK(on_edge == 1280 && edge_mismatch == 16);
}

// ---- physics (physics/world.rs, rigid_body.rs, force_generator.rs) ---------

TEST(test_world_free_fall_is_exact_with_power_of_two_steps)
{
    World world;
    RigidBody body;
    body.set_mass(1.0f);
    body.set_can_sleep(false);  // wakes it
    RigidBodyInstance b = world.register_body(body);
    world.register_force_generator(b, Gravity({0, -8, 0}));
    for (int n = 1; n <= 10; ++n) {
        world.run_physics(0.5f);
        RigidBody now = b.borrow();
        CHECK(now.get_velocity() == Vector3(0, -4.0f * n, 0));
        CHECK(now.get_position() == Vector3(0, -(float)(n * (n + 1)), 0));
    }
    // Quaternion::unit_z is a half turn about z; the translation is the position
    Matrix4x4 m = b.borrow().get_transform();
    CHECK(m.at(0, 0) == -1.0f && m.at(1, 1) == -1.0f && m.at(2, 2) == 1.0f);
    CHECK(m.at(0, 3) == 0.0f && m.at(1, 3) == -110.0f && m.at(2, 3) == 0.0f);
}

TEST(test_world_default_mass_gravity_wakes_and_asleep_bodies_stay)
{
    World world;
    RigidBodyInstance a = world.register_body(RigidBody());
    RigidBody still;
    still.set_velocity({1, 2, 3});
    RigidBodyInstance c = world.register_body(still);
    world.register_force_generator(a, Gravity({0, -9.81f, 0}));
    world.register_force_generator(c, NullForce());
    world.run_physics(0.25f);
    CHECK(a.borrow().get_awake());
    CHECK(!c.borrow().get_awake());
    CHECK(c.borrow().get_velocity() == Vector3(1, 2, 3) && c.borrow().get_position() == Vector3());
    // borrow_mut writes the body back: a woken body then moves
    c.borrow_mut([](RigidBody& body) { body.set_can_sleep(false); });
    world.run_physics(0.25f);
    CHECK(c.borrow().get_position() != Vector3());
}

TEST(test_world_spring_pulls_two_bodies_together)
{
    World world;
    RigidBody left, right;
    left.set_mass(1.0f), right.set_mass(1.0f);
    left.set_orientation(1, 0, 0, 0), right.set_orientation(1, 0, 0, 0);
    left.set_can_sleep(false), right.set_can_sleep(false);
    right.set_position({4, 0, 0});
    RigidBodyInstance l = world.register_body(left), r = world.register_body(right);
    world.register_force_generator(l, Spring(r, {0, 0, 0}, {0, 0, 0}, 2.0f, 1.0f));
    world.register_force_generator(r, Spring(l, {0, 0, 0}, {0, 0, 0}, 2.0f, 1.0f));
    world.start_frame();
    world.run_physics(1.0f / 64.0f);
    float vl = l.borrow().get_velocity().x, vr = r.borrow().get_velocity().x;
    // force 2 * (4 - 1) = 6 toward each other; one step of dt = 1/64 with drag 1
    CHECK(vl > 0.0f && vr < 0.0f && vl == -vr);
    CHECK(std::fabs(vl - 6.0f / 64.0f) < 1e-6f);
}

int main(int argc, char** argv)
{
    if (argc > 1)
        g_golden = argv[1];
    int failed = 0;
    for (const Case& c : cases()) {
        try {
            c.fn();
            std::printf("PASS %s\n", c.name);
        } catch (const Failure& f) {
            std::printf("FAIL %s: %s\n", c.name, f.what.c_str());
            ++failed;
        } catch (const std::exception& e) {
            std::printf("FAIL %s: exception %s\n", c.name, e.what());
            ++failed;
        }
        std::fflush(stdout);
    }
    std::printf("%zu tests, %d failed\n", cases().size(), failed);
    return failed;
}
