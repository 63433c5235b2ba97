/*
 * bvht_oracle.c -- CPU oracle (test infrastructure only; see bvht_oracle.h).
 *
 * Plain-C restatement of the reference crate's traversal hot path, in the
 * reference's operation order, for use as the checker of the HIP library and
 * as the CPU baseline timed beside it.  Build with -ffp-contract=off.
 */
#include "bvht_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    float x, y, z;
} v3;

typedef struct {
    v3 o, d, r;
    float t;
} ray_t;

static inline v3 mk(float x, float y, float z) { v3 v = {x, y, z}; return v; }
static inline v3 ld3(const float* p) { return mk(p[0], p[1], p[2]); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 scl(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross(v3 a, v3 b)
{
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline float comp(v3 a, int axis) { return axis == 0 ? a.x : (axis == 1 ? a.y : a.z); }
static inline v3 vmin(v3 a, v3 b) { return mk(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
static inline v3 vmax(v3 a, v3 b) { return mk(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); }

/* Ray::new: recip = 1/d per component, IEEE division (query/ray.rs:23-31). */
static inline ray_t ray_new(v3 o, v3 d, float t)
{
    ray_t r;
    r.o = o;
    r.d = d;
    r.r = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    r.t = t;
    return r;
}

static inline ray_t ray_from_rec(const orc_ray* in)
{
    return ray_new(ld3(in->o), ld3(in->d), in->tmax);
}

void orc_ray_recip(const float d[3], float out[3])
{
    ray_t r = ray_new(mk(0, 0, 0), ld3(d), FLT_MAX);
    out[0] = r.r.x;
    out[1] = r.r.y;
    out[2] = r.r.z;
}

/* cglinalg normalize: v / |v|, |v| = sqrt(dot(v, v)). */
static inline v3 normalize(v3 v)
{
    float n = sqrtf(dot(v, v));
    return mk(v.x / n, v.y / n, v.z / n);
}

void orc_normalize(const float in[3], float out[3])
{
    v3 r = normalize(ld3(in));
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}

/* Aabb::intersect, slab test (geometry/aabb.rs:106-125). */
static inline int aabb_hit(v3 bmin, v3 bmax, const ray_t* ray, float* tmin_out)
{
    float tx1 = (bmin.x - ray->o.x) * ray->r.x;
    float tx2 = (bmax.x - ray->o.x) * ray->r.x;
    float tmin = fminf(tx1, tx2);
    float tmax = fmaxf(tx1, tx2);
    float ty1 = (bmin.y - ray->o.y) * ray->r.y;
    float ty2 = (bmax.y - ray->o.y) * ray->r.y;
    tmin = fmaxf(tmin, fminf(ty1, ty2));
    tmax = fminf(tmax, fmaxf(ty1, ty2));
    float tz1 = (bmin.z - ray->o.z) * ray->r.z;
    float tz2 = (bmax.z - ray->o.z) * ray->r.z;
    tmin = fmaxf(tmin, fminf(tz1, tz2));
    tmax = fminf(tmax, fmaxf(tz1, tz2));
    if ((tmax >= tmin) && (tmin < ray->t) && (tmax > 0.0f)) {
        *tmin_out = tmin;
        return 1;
    }
    return 0;
}

int orc_aabb_intersect(const float bmin[3], const float bmax[3], const orc_ray* ray, float* tmin)
{
    ray_t r = ray_from_rec(ray);
    return aabb_hit(ld3(bmin), ld3(bmax), &r, tmin);
}

/* Triangle::intersect, Moeller-Trumbore with the absolute 1e-4 thresholds
 * (geometry/triangle.rs:41-72). Returns t = min(ray.t, t). */
static inline int tri_hit(const float* tri, const ray_t* ray, float* t_out, float* u_out, float* v_out)
{
    v3 v0 = ld3(tri), v1 = ld3(tri + 3), v2 = ld3(tri + 6);
    v3 edge1 = sub(v1, v0);
    v3 edge2 = sub(v2, v0);
    v3 normal = cross(ray->d, edge2);
    float area = dot(edge1, normal);
    const float threshold = 0.0001f;
    if (fabsf(area) < threshold)
        return 0;
    float f = 1.0f / area;
    v3 s = sub(ray->o, v0);
    float u = f * dot(s, normal);
    if (u < 0.0f || u > 1.0f)
        return 0;
    v3 q = cross(s, edge1);
    float v = f * dot(ray->d, q);
    if (v < 0.0f || u + v > 1.0f)
        return 0;
    float t = f * dot(edge2, q);
    if (t > threshold) {
        *t_out = fminf(ray->t, t);
        *u_out = u;
        *v_out = v;
        return 1;
    }
    return 0;
}

int orc_triangle_intersect(const float tri[9], const orc_ray* ray, float* t, float* u, float* v)
{
    ray_t r = ray_from_rec(ray);
    return tri_hit(tri, &r, t, u, v);
}

/* Triangle::centroid: (v0 + v1 + v2) * (1/3) (geometry/triangle.rs:33-39). */
static inline v3 centroid(const float* tri)
{
    const float one = 1.0f;
    const float one_third = one / (one + one + one);
    return scl(add(add(ld3(tri), ld3(tri + 3)), ld3(tri + 6)), one_third);
}

void orc_triangle_centroid(const float tri[9], float out[3])
{
    v3 c = centroid(tri);
    out[0] = c.x;
    out[1] = c.y;
    out[2] = c.z;
}

/* InstancePrimitiveIndex::new (query/intersection.rs:41-47). */
uint32_t orc_pack_index(uint32_t inst, uint32_t prim)
{
    return ((inst & 0x00000FFFu) << 20) | (prim & 0x000FFFFFu);
}

/* ------------------------------------------------------------------ */
/* BLAS build (model/bvh.rs:317-541)                                    */
/* ------------------------------------------------------------------ */

typedef struct {
    v3 bmin, bmax;
} box_t;

/* Aabb::area without the factor 2 (geometry/aabb.rs:48-52). */
static inline float box_area(box_t b)
{
    v3 e = sub(b.bmax, b.bmin);
    return e.x * e.y + e.y * e.z + e.z * e.x;
}

/* Aabb::grow / grow_aabb (geometry/aabb.rs:36-46). */
static inline void box_grow(box_t* b, v3 p)
{
    b->bmin = vmin(b->bmin, p);
    b->bmax = vmax(b->bmax, p);
}

static inline void box_grow_box(box_t* b, const box_t* o)
{
    if (o->bmin.x != FLT_MAX) {
        box_grow(b, o->bmin);
        box_grow(b, o->bmax);
    }
}

static inline box_t box_empty(void)
{
    box_t b = {{FLT_MAX, FLT_MAX, FLT_MAX}, {-FLT_MAX, -FLT_MAX, -FLT_MAX}};
    return b;
}

static inline box_t box_zero(void)
{
    box_t b = {{0, 0, 0}, {0, 0, 0}};
    return b;
}

static inline box_t node_box(const orc_node* n)
{
    box_t b = {ld3(n->bmin), ld3(n->bmax)};
    return b;
}

static inline void set_node_box(orc_node* n, box_t b)
{
    n->bmin[0] = b.bmin.x; n->bmin[1] = b.bmin.y; n->bmin[2] = b.bmin.z;
    n->bmax[0] = b.bmax.x; n->bmax[1] = b.bmax.y; n->bmax[2] = b.bmax.z;
}

/* Bvh::update_node_bounds: min/max over the leaf's vertices starting from
 * (+MAX, -MAX) (model/bvh.rs:317-330). */
static void update_node_bounds(orc_node* node, const float* tris)
{
    box_t b = box_empty();
    for (uint32_t k = 0; k < node->count; ++k) {
        const float* t = tris + 9u * (size_t)(node->left_first + k);
        b.bmin = vmin(b.bmin, ld3(t));
        b.bmin = vmin(b.bmin, ld3(t + 3));
        b.bmin = vmin(b.bmin, ld3(t + 6));
        b.bmax = vmax(b.bmax, ld3(t));
        b.bmax = vmax(b.bmax, ld3(t + 3));
        b.bmax = vmax(b.bmax, ld3(t + 6));
    }
    set_node_box(node, b);
}

#define BIN_COUNT 8

/* `f32 as usize` saturating cast followed by usize::min(BIN_COUNT-1, _)
 * (model/bvh.rs:352-353): NaN and negatives give 0. */
static inline uint32_t bin_index(float p)
{
    if (!(p > 0.0f))
        return 0;
    if (p >= (float)(BIN_COUNT - 1))
        return BIN_COUNT - 1;
    return (uint32_t)p;
}

/* Bvh::find_best_split_plane (model/bvh.rs:333-394).  In faithful mode the
 * quirks of the reference are kept: the centroid range starts at
 * (1e30, 1e-30) and every bin and sweep box starts as the zero box, so it
 * contains the origin.  Non-faithful mode starts them empty and skips planes
 * with an empty side. */
static void find_best_split_plane(const orc_node* node, const float* tris, int faithful,
                                  int* best_axis_out, float* best_pos_out, float* best_cost_out)
{
    int best_axis = -1;
    float best_position = 0.0f;
    float best_cost = FLT_MAX;
    for (int axis = 0; axis < 3; ++axis) {
        float bounds_min = faithful ? 1e30f : FLT_MAX;
        float bounds_max = faithful ? 1e-30f : -FLT_MAX;
        for (uint32_t k = 0; k < node->count; ++k) {
            float c = comp(centroid(tris + 9u * (size_t)(node->left_first + k)), axis);
            bounds_min = fminf(bounds_min, c);
            bounds_max = fmaxf(bounds_max, c);
        }
        if (bounds_min == bounds_max)
            continue;

        box_t bin_box[BIN_COUNT];
        uint32_t bin_count[BIN_COUNT];
        for (int b = 0; b < BIN_COUNT; ++b) {
            bin_box[b] = faithful ? box_zero() : box_empty();
            bin_count[b] = 0;
        }
        float bin_scale = (float)BIN_COUNT / (bounds_max - bounds_min);
        for (uint32_t k = 0; k < node->count; ++k) {
            const float* t = tris + 9u * (size_t)(node->left_first + k);
            uint32_t bi = bin_index((comp(centroid(t), axis) - bounds_min) * bin_scale);
            bin_count[bi] += 1;
            box_grow(&bin_box[bi], ld3(t));
            box_grow(&bin_box[bi], ld3(t + 3));
            box_grow(&bin_box[bi], ld3(t + 6));
        }

        float left_area[BIN_COUNT - 1], right_area[BIN_COUNT - 1];
        uint32_t left_count[BIN_COUNT - 1], right_count[BIN_COUNT - 1];
        box_t left_box = faithful ? box_zero() : box_empty();
        box_t right_box = faithful ? box_zero() : box_empty();
        uint32_t left_sum = 0, right_sum = 0;
        for (int i = 0; i < BIN_COUNT - 1; ++i) {
            left_sum += bin_count[i];
            left_count[i] = left_sum;
            box_grow_box(&left_box, &bin_box[i]);
            left_area[i] = box_area(left_box);

            right_sum += bin_count[BIN_COUNT - 1 - i];
            right_count[BIN_COUNT - 2 - i] = right_sum;
            box_grow_box(&right_box, &bin_box[BIN_COUNT - 1 - i]);
            right_area[BIN_COUNT - 2 - i] = box_area(right_box);
        }

        float scale = (bounds_max - bounds_min) / (float)BIN_COUNT;
        for (int i = 0; i < BIN_COUNT - 1; ++i) {
            if (!faithful && (left_count[i] == 0 || right_count[i] == 0))
                continue;
            float plane_cost = (float)left_count[i] * left_area[i] + (float)right_count[i] * right_area[i];
            if (plane_cost < best_cost) {
                best_axis = axis;
                best_position = bounds_min + scale * (float)(i + 1);
                best_cost = plane_cost;
            }
        }
    }
    *best_axis_out = best_axis;
    *best_pos_out = best_position;
    *best_cost_out = best_cost;
}

/* In-place two-pointer partition (model/bvh.rs:420-429).  Signed 64-bit
 * cursors: the reference's u32 `j` underflows (a panic) when every primitive
 * of the root goes right; here that case yields the same swap sequence
 * without the panic. */
static uint32_t partition(float* tris, uint32_t first, uint32_t count, int axis, float split)
{
    int64_t i = first;
    int64_t j = (int64_t)first + count - 1;
    while (i <= j) {
        float* ti = tris + 9 * (size_t)i;
        if (comp(centroid(ti), axis) < split) {
            i += 1;
        } else {
            float* tj = tris + 9 * (size_t)j;
            float tmp[9];
            memcpy(tmp, ti, sizeof tmp);
            memcpy(ti, tj, sizeof tmp);
            memcpy(tj, tmp, sizeof tmp);
            j -= 1;
        }
    }
    return (uint32_t)i;
}

typedef struct {
    orc_node* nodes;
    uint32_t nodes_used;
    float* tris;
    int faithful;
    uint32_t max_leaf;
} builder_t;

/* Bvh::subdivide (model/bvh.rs:396-467), with the recursion replaced by an
 * explicit stack that visits nodes in the same depth-first order, so child
 * pairs are allocated in the same pre-order (left subtree first). */
static int subdivide(builder_t* b, uint32_t root)
{
    uint32_t cap = 64, sp = 0;
    uint32_t* stack = (uint32_t*)malloc(cap * sizeof(uint32_t));
    if (!stack)
        return ORC_ENOMEM;
    stack[sp++] = root;
    while (sp) {
        uint32_t ni = stack[--sp];
        orc_node* node = &b->nodes[ni];
        int axis;
        float pos, cost;
        find_best_split_plane(node, b->tris, b->faithful, &axis, &pos, &cost);
        float no_split_cost = (float)node->count * box_area(node_box(node));
        uint32_t first = node->left_first;
        uint32_t count = node->count;
        uint32_t split_i;
        int force = !b->faithful && b->max_leaf > 0 && count > b->max_leaf;
        if (!(cost >= no_split_cost) && axis >= 0) {
            split_i = partition(b->tris, first, count, axis, pos);
        } else if (force && axis >= 0) {
            split_i = partition(b->tris, first, count, axis, pos);
        } else if (force) {
            split_i = first + count / 2;
        } else {
            continue; /* leaf */
        }
        uint32_t left_count = split_i - first;
        if (left_count == 0 || left_count == count) {
            if (!force)
                continue;
            /* non-faithful median fallback */
            split_i = first + count / 2;
            left_count = count / 2;
        }
        uint32_t left = b->nodes_used++;
        uint32_t right = b->nodes_used++;
        b->nodes[left].left_first = first;
        b->nodes[left].count = left_count;
        b->nodes[right].left_first = split_i;
        b->nodes[right].count = count - left_count;
        node->left_first = left;
        node->count = 0;
        update_node_bounds(&b->nodes[left], b->tris);
        update_node_bounds(&b->nodes[right], b->tris);
        if (sp + 2 > cap) {
            cap *= 2;
            uint32_t* ns = (uint32_t*)realloc(stack, cap * sizeof(uint32_t));
            if (!ns) {
                free(stack);
                return ORC_ENOMEM;
            }
            stack = ns;
        }
        stack[sp++] = right;
        stack[sp++] = left;
    }
    free(stack);
    return ORC_OK;
}

/* BvhBuilder::build_for (model/bvh.rs:509-541): 2N default nodes, node 0 is
 * the root leaf over all primitives, node 1 stays an all-zero dummy,
 * nodes_used starts at 2.  The triangle array is permuted in place. */
int orc_bvh_build(float* tris, uint32_t n, int faithful, uint32_t max_leaf, orc_node* nodes,
                  uint32_t* nodes_used)
{
    if (n == 0 || !tris || !nodes || !nodes_used)
        return ORC_EINVAL;
    memset(nodes, 0, sizeof(orc_node) * 2u * (size_t)n);
    nodes[0].left_first = 0;
    nodes[0].count = n;
    update_node_bounds(&nodes[0], tris);
    builder_t b = {nodes, 2, tris, faithful, max_leaf};
    int rc = subdivide(&b, 0);
    *nodes_used = b.nodes_used;
    return rc;
}

/* Bvh::refit: nodes_used-1 down to 0, skipping 1 (model/bvh.rs:469-493). */
int orc_bvh_refit(orc_node* nodes, uint32_t nodes_used, const float* tris)
{
    if (!nodes || !tris)
        return ORC_EINVAL;
    for (int64_t i = (int64_t)nodes_used - 1; i >= 0; --i) {
        if (i == 1)
            continue;
        orc_node* n = &nodes[i];
        if (n->count > 0) {
            update_node_bounds(n, tris);
            continue;
        }
        const orc_node* l = &nodes[n->left_first];
        const orc_node* r = &nodes[n->left_first + 1];
        box_t b;
        b.bmin = vmin(ld3(l->bmin), ld3(r->bmin));
        b.bmax = vmax(ld3(l->bmax), ld3(r->bmax));
        set_node_box(n, b);
    }
    return ORC_OK;
}

static uint32_t depth_rec(const orc_node* nodes, uint32_t i)
{
    /* iterative DFS to avoid deep recursion */
    uint32_t cap = 64, sp = 0, best = 0;
    uint32_t* st = (uint32_t*)malloc(cap * 2 * sizeof(uint32_t));
    if (!st)
        return 0;
    st[0] = i;
    st[1] = 1;
    sp = 1;
    while (sp) {
        --sp;
        uint32_t ni = st[2 * sp], d = st[2 * sp + 1];
        if (d > best)
            best = d;
        if (nodes[ni].count == 0) {
            if (sp + 2 > cap) {
                cap *= 2;
                uint32_t* ns = (uint32_t*)realloc(st, cap * 2 * sizeof(uint32_t));
                if (!ns) {
                    free(st);
                    return 0;
                }
                st = ns;
            }
            st[2 * sp] = nodes[ni].left_first;
            st[2 * sp + 1] = d + 1;
            ++sp;
            st[2 * sp] = nodes[ni].left_first + 1;
            st[2 * sp + 1] = d + 1;
            ++sp;
        }
    }
    free(st);
    return best;
}

uint32_t orc_bvh_depth(const orc_node* nodes) { return depth_rec(nodes, 0); }

static inline void miss_hit(orc_hit* h)
{
    h->t = FLT_MAX;
    h->u = 0.0f;
    h->v = 0.0f;
    h->prim = 0xFFFFFFFFu;
    h->inst = 0xFFFFFFFFu;
    h->packed = 0xFFFFFFFFu;
}

typedef struct {
    float t, u, v;
    uint32_t prim;
    int found;
} blas_result;

/* Bvh::intersect_subtree (model/bvh.rs:242-305).  The root box is never
 * tested; both child boxes are tested against the current closest t; the
 * nearer is entered (ties go right); the farther is pushed only if hit;
 * popped nodes are not re-culled; leaf primitives are tested in array order
 * against the original ray and accepted iff t < closest (first found wins
 * among equal t). */
static blas_result bvh_traverse(const orc_node* nodes, const float* tris, const ray_t* ray,
                                uint32_t* stack, orc_stats* st)
{
    blas_result res = {0, 0, 0, 0, 0};
    const orc_node* cur = &nodes[0];
    ray_t closest = *ray;
    uint32_t sp = 0;
    float best_t = 0, best_u = 0, best_v = 0;
    uint32_t best_prim = 0;
    int have = 0;
    for (;;) {
        if (cur->count > 0) {
            uint32_t base = cur->left_first;
            for (uint32_t k = 0; k < cur->count; ++k) {
                uint32_t pi = base + k;
                float t, u, v;
                if (st)
                    st->tests++;
                if (tri_hit(tris + 9u * (size_t)pi, ray, &t, &u, &v)) {
                    if (t < closest.t) {
                        closest.t = t;
                        best_t = t;
                        best_u = u;
                        best_v = v;
                        best_prim = pi;
                        have = 1;
                    }
                }
            }
            if (sp)
                cur = &nodes[stack[--sp]];
            else
                break;
        } else {
            if (st)
                st->pairs++;
            uint32_t li = cur->left_first, ri = cur->left_first + 1;
            const orc_node* l = &nodes[li];
            const orc_node* r = &nodes[ri];
            float dl, dr;
            int hl = aabb_hit(ld3(l->bmin), ld3(l->bmax), &closest, &dl);
            int hr = aabb_hit(ld3(r->bmin), ld3(r->bmax), &closest, &dr);
            float kl = hl ? dl : FLT_MAX;
            float kr = hr ? dr : FLT_MAX;
            uint32_t near_i, far_i;
            int near_h, far_h;
            if (kl < kr) {
                near_i = li; near_h = hl; far_i = ri; far_h = hr;
            } else {
                near_i = ri; near_h = hr; far_i = li; far_h = hl;
            }
            if (near_h) {
                cur = &nodes[near_i];
                if (far_h)
                    stack[sp++] = far_i;
                continue;
            }
            if (sp)
                cur = &nodes[stack[--sp]];
            else
                break;
        }
    }
    if (closest.t < FLT_MAX && have) {
        res.t = best_t;
        res.u = best_u;
        res.v = best_v;
        res.prim = best_prim;
        res.found = 1;
    }
    return res;
}

int orc_bvh_intersect(const orc_node* nodes, const float* tris, const orc_ray* ray, orc_hit* hit,
                      orc_stats* st)
{
    uint32_t depth = orc_bvh_depth(nodes);
    uint32_t* stack = (uint32_t*)malloc((depth + 1) * sizeof(uint32_t));
    if (!stack)
        return ORC_ENOMEM;
    ray_t r = ray_from_rec(ray);
    if (st)
        memset(st, 0, sizeof *st);
    blas_result res = bvh_traverse(nodes, tris, &r, stack, st);
    free(stack);
    if (res.found) {
        hit->t = res.t;
        hit->u = res.u;
        hit->v = res.v;
        hit->prim = res.prim;
        hit->inst = 0;
        hit->packed = orc_pack_index(0, res.prim);
    } else {
        miss_hit(hit);
    }
    return res.found;
}

/* ------------------------------------------------------------------ */
/* Transforms, instances and TLAS                                       */
/* ------------------------------------------------------------------ */

/* Matrix4x4 * Vector4 then contract() (transform.rs:219-234). */
static inline v3 mat_mul4(const float* m, v3 p, float w)
{
    float r[3];
    for (int i = 0; i < 3; ++i)
        r[i] = m[0 * 4 + i] * p.x + m[1 * 4 + i] * p.y + m[2 * 4 + i] * p.z + m[3 * 4 + i] * w;
    return mk(r[0], r[1], r[2]);
}

void orc_transform_point(const float m[16], const float p[3], float out[3])
{
    v3 r = mat_mul4(m, ld3(p), 1.0f);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

void orc_transform_vector(const float m[16], const float v[3], float out[3])
{
    v3 r = mat_mul4(m, ld3(v), 0.0f);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

/* SceneObject::set_transform / SceneObjectBuilder::with_transform: world box
 * of the 8 transformed corners of the model's BVH root box
 * (scene/scene_object.rs:60-75, 109-126). */
void orc_instance_bounds(const float m[16], const float model_bmin[3], const float model_bmax[3],
                         float out_bmin[3], float out_bmax[3])
{
    box_t nb = box_empty();
    for (int i = 0; i < 8; ++i) {
        float x = (i & 1) ? model_bmax[0] : model_bmin[0];
        float y = (i & 2) ? model_bmax[1] : model_bmin[1];
        float z = (i & 4) ? model_bmax[2] : model_bmin[2];
        box_grow(&nb, mat_mul4(m, mk(x, y, z), 1.0f));
    }
    out_bmin[0] = nb.bmin.x; out_bmin[1] = nb.bmin.y; out_bmin[2] = nb.bmin.z;
    out_bmax[0] = nb.bmax.x; out_bmax[1] = nb.bmax.y; out_bmax[2] = nb.bmax.z;
}

/* Tlas::find_best_match (scene/tlas.rs:179-202). */
static int32_t find_best_match(const orc_tnode* nodes, const int32_t* list, int32_t n, int32_t a)
{
    float smallest = FLT_MAX;
    int32_t best_b = -1;
    for (int32_t b = 0; b < n; ++b) {
        if (b != a) {
            const orc_tnode* na = &nodes[list[a]];
            const orc_tnode* nb = &nodes[list[b]];
            v3 bmax = vmax(ld3(na->bmax), ld3(nb->bmax));
            v3 bmin = vmin(ld3(na->bmin), ld3(nb->bmin));
            v3 e = sub(bmax, bmin);
            float sa = e.x * e.y + e.y * e.z + e.z * e.x;
            if (sa < smallest) {
                smallest = sa;
                best_b = b;
            }
        }
    }
    return best_b;
}

/* TlasBuilder::build_for + Tlas::rebuild: leaves at 1..N in object order,
 * agglomerative clustering by mutual best match, root = copy of the last
 * merged node (scene/tlas.rs:204-279).  LeftRightIndex::new(a, b) stores
 * a + (b << 16) (tlas.rs:18-22). */
int orc_tlas_build(const float* bounds, uint32_t n, orc_tnode* nodes, uint32_t* nodes_used)
{
    if (n == 0 || n > 32767 || !bounds || !nodes || !nodes_used)
        return ORC_EINVAL;
    for (uint32_t i = 0; i < 2 * n; ++i) {
        nodes[i].bmin[0] = nodes[i].bmin[1] = nodes[i].bmin[2] = FLT_MAX;
        nodes[i].bmax[0] = nodes[i].bmax[1] = nodes[i].bmax[2] = -FLT_MAX;
        nodes[i].left_right = 0;
        nodes[i].blas = 0;
    }
    int32_t* list = (int32_t*)malloc(n * sizeof(int32_t));
    if (!list)
        return ORC_ENOMEM;
    int32_t count = (int32_t)n;
    int32_t used = 1;
    for (uint32_t i = 0; i < n; ++i) {
        list[i] = used;
        memcpy(nodes[used].bmin, bounds + 6 * i, 3 * sizeof(float));
        memcpy(nodes[used].bmax, bounds + 6 * i + 3, 3 * sizeof(float));
        nodes[used].blas = i;
        nodes[used].left_right = 0;
        used++;
    }
    int32_t a = 0;
    int32_t b = find_best_match(nodes, list, count, a);
    while (count > 1) {
        if (b < 0) {
            free(list);
            return ORC_EINVAL; /* every candidate area was >= FLT_MAX (reference panics) */
        }
        int32_t c = find_best_match(nodes, list, count, b);
        if (a == c) {
            int32_t ia = list[a], ib = list[b];
            orc_tnode na = nodes[ia], nb = nodes[ib];
            orc_tnode* nn = &nodes[used];
            nn->left_right = (uint32_t)ia + ((uint32_t)ib << 16);
            v3 mn = vmin(ld3(na.bmin), ld3(nb.bmin));
            v3 mx = vmax(ld3(na.bmax), ld3(nb.bmax));
            nn->bmin[0] = mn.x; nn->bmin[1] = mn.y; nn->bmin[2] = mn.z;
            nn->bmax[0] = mx.x; nn->bmax[1] = mx.y; nn->bmax[2] = mx.z;
            list[a] = used;
            used++;
            list[b] = list[count - 1];
            count--;
            b = find_best_match(nodes, list, count, a);
        } else {
            a = b;
            b = c;
        }
    }
    nodes[0] = nodes[list[a]];
    *nodes_used = (uint32_t)used;
    free(list);
    return ORC_OK;
}

/* Tlas::intersect (scene/tlas.rs:123-177) with SceneObject::intersect at the
 * leaves (scene_object.rs:77-89): the ray is carried into model space by the
 * cached inverse (point and vector), Ray::new recomputes the reciprocal with
 * the current closest t, and the BLAS result is accepted iff its t is below
 * the current closest.  LeftRightIndex::left() decodes the high half. */
static int tlas_traverse(const orc_tnode* tnodes, const orc_instance* inst,
                         const orc_node* const* blas_nodes, const float* const* blas_tris,
                         const ray_t* ray, uint32_t* tstack, uint32_t* bstack, orc_hit* hit,
                         orc_stats* st)
{
    const orc_tnode* cur = &tnodes[0];
    ray_t closest = *ray;
    uint32_t sp = 0;
    int have = 0;
    blas_result best = {0, 0, 0, 0, 0};
    uint32_t best_inst = 0;
    for (;;) {
        if (cur->left_right == 0) {
            const orc_instance* obj = &inst[cur->blas];
            if (st)
                st->inst_visits++;
            v3 om = mat_mul4(obj->minv, closest.o, 1.0f);
            v3 dm = mat_mul4(obj->minv, closest.d, 0.0f);
            ray_t rm = ray_new(om, dm, closest.t);
            blas_result r = bvh_traverse(blas_nodes[obj->blas], blas_tris[obj->blas], &rm, bstack, st);
            if (r.found && r.t < closest.t) {
                closest.t = r.t;
                best = r;
                best_inst = cur->blas;
                have = 1;
            }
            if (sp)
                cur = &tnodes[tstack[--sp]];
            else
                break;
        } else {
            if (st)
                st->tlas_pairs++;
            uint32_t li = (cur->left_right & 0xFFFF0000u) >> 16;
            uint32_t ri = cur->left_right & 0x0000FFFFu;
            const orc_tnode* l = &tnodes[li];
            const orc_tnode* r = &tnodes[ri];
            float dl, dr;
            int hl = aabb_hit(ld3(l->bmin), ld3(l->bmax), &closest, &dl);
            int hr = aabb_hit(ld3(r->bmin), ld3(r->bmax), &closest, &dr);
            float kl = hl ? dl : FLT_MAX;
            float kr = hr ? dr : FLT_MAX;
            uint32_t near_i, far_i;
            int near_h, far_h;
            if (kl < kr) {
                near_i = li; near_h = hl; far_i = ri; far_h = hr;
            } else {
                near_i = ri; near_h = hr; far_i = li; far_h = hl;
            }
            if (near_h) {
                cur = &tnodes[near_i];
                if (far_h)
                    tstack[sp++] = far_i;
                continue;
            }
            if (sp)
                cur = &tnodes[tstack[--sp]];
            else
                break;
        }
    }
    if (closest.t < FLT_MAX && have) {
        hit->t = best.t;
        hit->u = best.u;
        hit->v = best.v;
        hit->prim = best.prim;
        hit->inst = best_inst;
        hit->packed = orc_pack_index(0, best.prim);
        return 1;
    }
    miss_hit(hit);
    return 0;
}

static uint32_t tlas_depth(const orc_tnode* t)
{
    /* a TLAS over N <= 32767 objects has at most 65534 nodes */
    uint32_t cap = 2 * 65536 + 2;
    uint32_t* st = (uint32_t*)malloc(cap * sizeof(uint32_t));
    uint32_t sp = 1, best = 0;
    if (!st)
        return 0;
    st[0] = 0;
    st[1] = 1;
    while (sp) {
        --sp;
        uint32_t ni = st[2 * sp], d = st[2 * sp + 1];
        if (d > best)
            best = d;
        if (t[ni].left_right != 0 && 2 * (sp + 2) <= cap) {
            st[2 * sp] = (t[ni].left_right >> 16) & 0xFFFF;
            st[2 * sp + 1] = d + 1;
            ++sp;
            st[2 * sp] = t[ni].left_right & 0xFFFF;
            st[2 * sp + 1] = d + 1;
            ++sp;
        }
    }
    free(st);
    return best;
}

static uint32_t max_blas_depth(const orc_tnode* tnodes, const orc_instance* inst,
                               const orc_node* const* blas_nodes)
{
    /* scan the TLAS leaves reachable from the root; each holds an object id */
    uint32_t cap = 65536, sp = 1, best = 0;
    uint32_t* st = (uint32_t*)malloc(cap * sizeof(uint32_t));
    if (!st)
        return 0;
    st[0] = 0;
    while (sp) {
        uint32_t ni = st[--sp];
        if (tnodes[ni].left_right == 0) {
            uint32_t d = orc_bvh_depth(blas_nodes[inst[tnodes[ni].blas].blas]);
            if (d > best)
                best = d;
        } else if (sp + 2 <= cap) {
            st[sp++] = (tnodes[ni].left_right >> 16) & 0xFFFF;
            st[sp++] = tnodes[ni].left_right & 0xFFFF;
        }
    }
    free(st);
    return best;
}

int orc_tlas_intersect(const orc_tnode* tnodes, const orc_instance* inst,
                       const orc_node* const* blas_nodes, const float* const* blas_tris,
                       const orc_ray* ray, orc_hit* hit, orc_stats* st)
{
    uint32_t td = tlas_depth(tnodes) + 1, bd = max_blas_depth(tnodes, inst, blas_nodes) + 1;
    uint32_t* ts = (uint32_t*)malloc(td * sizeof(uint32_t));
    uint32_t* bs = (uint32_t*)malloc(bd * sizeof(uint32_t));
    if (!ts || !bs) {
        free(ts);
        free(bs);
        return ORC_ENOMEM;
    }
    if (st)
        memset(st, 0, sizeof *st);
    ray_t r = ray_from_rec(ray);
    int found = tlas_traverse(tnodes, inst, blas_nodes, blas_tris, &r, ts, bs, hit, st);
    free(ts);
    free(bs);
    return found;
}

/* ------------------------------------------------------------------ */
/* Batched traces (contiguous ray chunks per thread)                    */
/* ------------------------------------------------------------------ */

typedef struct {
    int kind; /* 0 blas, 1 tlas, 2 differentiable forward */
    const uint32_t* vid;
    const float* normals;
    const float* minv;
    const float* light;
    float* image;
    const orc_node* nodes;
    const float* tris;
    const orc_tnode* tnodes;
    const orc_instance* inst;
    const orc_node* const* blas_nodes;
    const float* const* blas_tris;
    const orc_ray* rays;
    orc_hit* hits;
    orc_stats* stats;
    uint64_t begin, end;
    uint32_t depth, tdepth;
    int rc;
} job_t;

static void diff_forward_one(const job_t* j, uint64_t i, uint32_t* stack);

static void* run_job(void* arg)
{
    job_t* j = (job_t*)arg;
    uint32_t* bs = (uint32_t*)malloc((j->depth + 1) * sizeof(uint32_t));
    uint32_t* ts = (uint32_t*)malloc((j->tdepth + 1) * sizeof(uint32_t));
    if (!bs || !ts) {
        free(bs);
        free(ts);
        j->rc = ORC_ENOMEM;
        return NULL;
    }
    for (uint64_t i = j->begin; i < j->end; ++i) {
        ray_t r = ray_from_rec(&j->rays[i]);
        orc_stats* st = j->stats ? &j->stats[i] : NULL;
        if (st)
            memset(st, 0, sizeof *st);
        if (j->kind == 2) {
            diff_forward_one(j, i, bs);
        } else if (j->kind == 0) {
            blas_result res = bvh_traverse(j->nodes, j->tris, &r, bs, st);
            orc_hit* h = &j->hits[i];
            if (res.found) {
                h->t = res.t;
                h->u = res.u;
                h->v = res.v;
                h->prim = res.prim;
                h->inst = 0;
                h->packed = orc_pack_index(0, res.prim);
            } else {
                miss_hit(h);
            }
        } else {
            tlas_traverse(j->tnodes, j->inst, j->blas_nodes, j->blas_tris, &r, ts, bs, &j->hits[i], st);
        }
    }
    free(bs);
    free(ts);
    j->rc = ORC_OK;
    return NULL;
}

static int run_jobs(job_t proto, uint64_t n, int nthreads)
{
    if (nthreads < 1)
        nthreads = 1;
    if (nthreads > 256)
        nthreads = 256;
    if ((uint64_t)nthreads > n)
        nthreads = n ? (int)n : 1;
    job_t jobs[256];
    pthread_t th[256];
    uint64_t chunk = (n + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; ++t) {
        jobs[t] = proto;
        jobs[t].begin = (uint64_t)t * chunk;
        jobs[t].end = jobs[t].begin + chunk < n ? jobs[t].begin + chunk : n;
        if (jobs[t].begin > n)
            jobs[t].begin = n;
    }
    if (nthreads == 1) {
        run_job(&jobs[0]);
        return jobs[0].rc;
    }
    for (int t = 0; t < nthreads; ++t)
        pthread_create(&th[t], NULL, run_job, &jobs[t]);
    int rc = ORC_OK;
    for (int t = 0; t < nthreads; ++t) {
        pthread_join(th[t], NULL);
        if (jobs[t].rc != ORC_OK)
            rc = jobs[t].rc;
    }
    return rc;
}

int orc_bvh_trace(const orc_node* nodes, const float* tris, const orc_ray* rays, uint64_t n,
                  orc_hit* hits, orc_stats* stats, int nthreads)
{
    if (!nodes || !tris || (!rays && n) || (!hits && n))
        return ORC_EINVAL;
    job_t p;
    memset(&p, 0, sizeof p);
    p.kind = 0;
    p.nodes = nodes;
    p.tris = tris;
    p.rays = rays;
    p.hits = hits;
    p.stats = stats;
    p.depth = orc_bvh_depth(nodes);
    p.tdepth = 1;
    return run_jobs(p, n, nthreads);
}

int orc_tlas_trace(const orc_tnode* tnodes, const orc_instance* inst,
                   const orc_node* const* blas_nodes, const float* const* blas_tris,
                   const orc_ray* rays, uint64_t n, orc_hit* hits, orc_stats* stats, int nthreads)
{
    if (!tnodes || !inst || !blas_nodes || !blas_tris || (!rays && n) || (!hits && n))
        return ORC_EINVAL;
    job_t p;
    memset(&p, 0, sizeof p);
    p.kind = 1;
    p.tnodes = tnodes;
    p.inst = inst;
    p.blas_nodes = blas_nodes;
    p.blas_tris = blas_tris;
    p.rays = rays;
    p.hits = hits;
    p.stats = stats;
    p.depth = max_blas_depth(tnodes, inst, blas_nodes);
    p.tdepth = tlas_depth(tnodes);
    return run_jobs(p, n, nthreads);
}

/* Brute force nearest hit over all triangles in array order, with the same
 * per-triangle test and strict-< acceptance (no BVH): pins the traversal's
 * result up to the choice among exactly equal t.  ties[i] receives the
 * number of triangles whose returned t equals the minimum. */
int orc_brute_trace(const float* tris, uint32_t ntris, const orc_ray* rays, uint64_t n, orc_hit* hits,
                    uint32_t* ties)
{
    if ((!tris && ntris) || (!rays && n) || (!hits && n))
        return ORC_EINVAL;
    for (uint64_t i = 0; i < n; ++i) {
        ray_t r = ray_from_rec(&rays[i]);
        float closest = r.t, bu = 0, bv = 0;
        uint32_t best = 0, nt = 0;
        int have = 0;
        for (uint32_t k = 0; k < ntris; ++k) {
            float t, u, v;
            if (tri_hit(tris + 9u * (size_t)k, &r, &t, &u, &v)) {
                if (t < closest) {
                    closest = t;
                    bu = u;
                    bv = v;
                    best = k;
                    have = 1;
                    nt = 1;
                } else if (have && t == closest) {
                    nt++;
                }
            }
        }
        if (have && closest < FLT_MAX) {
            hits[i].t = closest;
            hits[i].u = bu;
            hits[i].v = bv;
            hits[i].prim = best;
            hits[i].inst = 0;
            hits[i].packed = orc_pack_index(0, best);
        } else {
            miss_hit(&hits[i]);
            nt = 0;
        }
        if (ties)
            ties[i] = nt;
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Camera (camera/camera.rs:994-1010)                                   */
/* ------------------------------------------------------------------ */

/* Camera::get_ray_eye then get_ray_world: pixel = 0 + tl + (tr-tl)u + (bl-tl)v,
 * direction normalised, both carried to world space by view_matrix_inv,
 * t = f32::MAX (Ray::from_origin_dir, query/ray.rs:33-35). */
void orc_camera_ray(const orc_camera* cam, float u, float v, orc_ray* out)
{
    v3 origin = mk(0.0f, 0.0f, 0.0f);
    v3 tl = ld3(cam->tl), tr = ld3(cam->tr), bl = ld3(cam->bl);
    v3 pixel = add(add(add(origin, tl), scl(sub(tr, tl), u)), scl(sub(bl, tl), v));
    v3 dir = normalize(sub(pixel, origin));
    v3 ow = mat_mul4(cam->view_inv, origin, 1.0f);
    v3 dw = mat_mul4(cam->view_inv, dir, 0.0f);
    out->o[0] = ow.x; out->o[1] = ow.y; out->o[2] = ow.z;
    out->tmax = FLT_MAX;
    out->d[0] = dw.x; out->d[1] = dw.y; out->d[2] = dw.z;
    out->pad = 0.0f;
}

/* PathTracer::evaluate's per-pixel mapping u = x/W, v = y/H
 * (renderer.rs:358-362); rays are written in pixel (row-major) order. */
int orc_raygen(const orc_camera* cam, uint32_t width, uint32_t height, orc_ray* rays)
{
    if (!cam || !rays || width == 0 || height == 0)
        return ORC_EINVAL;
    for (uint32_t y = 0; y < height; ++y)
        for (uint32_t x = 0; x < width; ++x)
            orc_camera_ray(cam, (float)x / (float)width, (float)y / (float)height,
                           &rays[(size_t)y * width + x]);
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Differentiable depth/shading checker                                 */
/* ------------------------------------------------------------------ */
/* No reference counterpart (the reference has no differentiable path).
 * These restate the GPU library's forward/adjoint pair so that it can be
 * checked: the forward in binary32 with the same operation order (bitwise
 * comparison), the gradient in binary64 (tolerance comparison; its own
 * correctness is pinned by central finite differences in the tests). */

void orc_mesh_normals(const float* V, uint32_t nv, const uint32_t* F, uint32_t nf, float* N)
{
    for (size_t k = 0; k < 3 * (size_t)nv; ++k)
        N[k] = 0.0f;
    for (uint32_t f = 0; f < nf; ++f) {
        v3 a = ld3(V + 3 * (size_t)F[3 * (size_t)f]);
        v3 b = ld3(V + 3 * (size_t)F[3 * (size_t)f + 1]);
        v3 c = ld3(V + 3 * (size_t)F[3 * (size_t)f + 2]);
        v3 n = cross(sub(b, a), sub(c, a));
        for (int k = 0; k < 3; ++k) {
            float* acc = N + 3 * (size_t)F[3 * (size_t)f + k];
            v3 s = add(ld3(acc), n);
            acc[0] = s.x;
            acc[1] = s.y;
            acc[2] = s.z;
        }
    }
    for (uint32_t i = 0; i < nv; ++i) {
        float* p = N + 3 * (size_t)i;
        v3 a = ld3(p);
        float len = sqrtf(dot(a, a));
        if (len > 0.0f) {
            p[0] = a.x / len;
            p[1] = a.y / len;
            p[2] = a.z / len;
        } else {
            p[0] = p[1] = p[2] = 0.0f;
        }
    }
}

/* Lambert term of the barycentric-interpolated vertex normal:
 * max(0, m.L/|m|), m = ((1-u)-v) N0 + u N1 + v N2. */
static float diff_shade(const float* normals, const uint32_t* vid, uint32_t prim, float u, float v, v3 light)
{
    v3 n0 = ld3(normals + 3 * (size_t)vid[3 * (size_t)prim]);
    v3 n1 = ld3(normals + 3 * (size_t)vid[3 * (size_t)prim + 1]);
    v3 n2 = ld3(normals + 3 * (size_t)vid[3 * (size_t)prim + 2]);
    float w = (1.0f - u) - v;
    v3 m = add(add(scl(n0, w), scl(n1, u)), scl(n2, v));
    float ml = dot(m, light);
    float len = sqrtf(dot(m, m));
    return (ml > 0.0f && len > 0.0f) ? ml / len : 0.0f;
}

static void diff_forward_one(const job_t* j, uint64_t i, uint32_t* stack)
{
    const orc_ray* in = &j->rays[i];
    v3 o = mat_mul4(j->minv, ld3(in->o), 1.0f);
    v3 d = mat_mul4(j->minv, ld3(in->d), 0.0f);
    ray_t r = ray_new(o, d, in->tmax);
    blas_result res = bvh_traverse(j->nodes, j->tris, &r, stack, NULL);
    orc_hit* h = &j->hits[i];
    float* img = j->image + 2 * i;
    if (res.found) {
        h->t = res.t;
        h->u = res.u;
        h->v = res.v;
        h->prim = res.prim;
        h->inst = 0;
        h->packed = orc_pack_index(0, res.prim);
        img[0] = res.t;
        img[1] = diff_shade(j->normals, j->vid, res.prim, res.u, res.v, ld3(j->light));
    } else {
        miss_hit(h);
        img[0] = FLT_MAX;
        img[1] = 0.0f;
    }
}

int orc_diff_forward(const orc_node* nodes, const float* tris, const uint32_t* vid, const float* normals,
                     const float minv[16], const float light[3], const orc_ray* rays, uint64_t n, orc_hit* hits,
                     float* image, int nthreads)
{
    if (!nodes || !tris || !vid || !normals || !minv || !light || (n && (!rays || !hits || !image)))
        return ORC_EINVAL;
    job_t p;
    memset(&p, 0, sizeof p);
    p.kind = 2;
    p.nodes = nodes;
    p.tris = tris;
    p.vid = vid;
    p.normals = normals;
    p.minv = minv;
    p.light = light;
    p.rays = rays;
    p.hits = hits;
    p.image = image;
    p.depth = orc_bvh_depth(nodes);
    p.tdepth = 1;
    return run_jobs(p, n, nthreads);
}

typedef struct {
    double x, y, z;
} d3;

static inline d3 dk(double x, double y, double z) { d3 v = {x, y, z}; return v; }
static inline d3 dld(const double* p) { return dk(p[0], p[1], p[2]); }
static inline d3 dsub(d3 a, d3 b) { return dk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline d3 dadd(d3 a, d3 b) { return dk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline d3 dscl(d3 a, double s) { return dk(a.x * s, a.y * s, a.z * s); }
static inline double ddot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline d3 dcross(d3 a, d3 b)
{
    return dk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline d3 dmul4(const double* m, d3 p, double w)
{
    return dk(m[0] * p.x + m[4] * p.y + m[8] * p.z + m[12] * w, m[1] * p.x + m[5] * p.y + m[9] * p.z + m[13] * w,
              m[2] * p.x + m[6] * p.y + m[10] * p.z + m[14] * w);
}

/* Loss and gradient in binary64 for fixed per-ray hit triangles prims[i]
 * (BLAS build order; 0xFFFFFFFF = miss).  Ray i contributes iff it hits and
 * its target depth is below FLT_MAX:
 *   loss += w_depth (t - t*)^2 + w_shade (s - s*)^2
 * with t, u, v the exact solution of o' + t d' = v0 + u e1 + v e2
 * (o' = M^-1 (o,1), d' = M^-1 (d,0)) and s the Lambert term of the
 * interpolated normals N (held fixed).  grad_v (3 nv) and grad_minv (16,
 * column-major; bottom row 0) receive dLoss/dV and dLoss/dM^-1 by the
 * closed-form derivatives of that linear solve; image (2n, may be NULL)
 * receives {t, s} (0 where the ray does not contribute). */
int orc_diff_grad64(const double* V, uint32_t nv, const uint32_t* vid, uint32_t ntris, const double* N,
                    const double minv[16], const double light[3], const orc_ray* rays, const uint32_t* prims,
                    uint64_t n, const float* target, double w_depth, double w_shade, double* loss, double* grad_v,
                    double* grad_minv, double* image)
{
    if (!V || !vid || !N || !minv || !light || !loss || !grad_v || !grad_minv || (n && (!rays || !prims || !target)))
        return ORC_EINVAL;
    for (size_t k = 0; k < 3 * (size_t)ntris; ++k)
        if (vid[k] >= nv)
            return ORC_EINVAL;
    memset(grad_v, 0, sizeof(double) * 3 * (size_t)nv);
    memset(grad_minv, 0, sizeof(double) * 16);
    double L = 0.0;
    d3 lt = dld(light);
    for (uint64_t i = 0; i < n; ++i) {
        if (image) {
            image[2 * i] = 0.0;
            image[2 * i + 1] = 0.0;
        }
        uint32_t p = prims[i];
        if (p >= ntris || !(target[2 * i] < FLT_MAX))
            continue;
        d3 ow = dk(rays[i].o[0], rays[i].o[1], rays[i].o[2]);
        d3 dw = dk(rays[i].d[0], rays[i].d[1], rays[i].d[2]);
        d3 o = dmul4(minv, ow, 1.0), d = dmul4(minv, dw, 0.0);
        const uint32_t* tv = vid + 3 * (size_t)p;
        d3 v0 = dld(V + 3 * (size_t)tv[0]), v1 = dld(V + 3 * (size_t)tv[1]), v2 = dld(V + 3 * (size_t)tv[2]);
        d3 e1 = dsub(v1, v0), e2 = dsub(v2, v0), s = dsub(o, v0);
        d3 a0 = dscl(d, -1.0);
        d3 c12 = dcross(e1, e2), c20 = dcross(e2, a0), c01 = dcross(a0, e1);
        double det = ddot(a0, c12);
        if (det == 0.0)
            continue;
        /* Cramer: x_k = (row k of A^-1) . s, rows = c12/det, c20/det, c01/det */
        double t = ddot(c12, s) / det, u = ddot(c20, s) / det, v = ddot(c01, s) / det;
        d3 n0 = dld(N + 3 * (size_t)tv[0]), n1 = dld(N + 3 * (size_t)tv[1]), n2 = dld(N + 3 * (size_t)tv[2]);
        double w = 1.0 - u - v;
        d3 m = dadd(dadd(dscl(n0, w), dscl(n1, u)), dscl(n2, v));
        double ml = ddot(m, lt), len = sqrt(ddot(m, m));
        int lit = ml > 0.0 && len > 0.0;
        double sh = lit ? ml / len : 0.0;
        double rt = t - (double)target[2 * i], rs = sh - (double)target[2 * i + 1];
        L += w_depth * rt * rt + w_shade * rs * rs;
        if (image) {
            image[2 * i] = t;
            image[2 * i + 1] = sh;
        }
        double gt = 2.0 * w_depth * rt, gs = 2.0 * w_shade * rs, gu = 0.0, gv = 0.0;
        if (lit) {
            d3 dsdm = dscl(dsub(lt, dscl(m, sh / len)), 1.0 / len);
            gu = gs * ddot(dsdm, dsub(n1, n0));
            gv = gs * ddot(dsdm, dsub(n2, n0));
        }
        d3 lam = dscl(dadd(dadd(dscl(c12, gt), dscl(c20, gu)), dscl(c01, gv)), 1.0 / det);
        double cw[3] = {-w, -u, -v};
        for (int k = 0; k < 3; ++k) {
            double* g = grad_v + 3 * (size_t)tv[k];
            g[0] += cw[k] * lam.x;
            g[1] += cw[k] * lam.y;
            g[2] += cw[k] * lam.z;
        }
        double lr[3] = {lam.x, lam.y, lam.z};
        double oc[4] = {ow.x, ow.y, ow.z, 1.0}, dc[4] = {dw.x, dw.y, dw.z, 0.0};
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 4; ++c)
                grad_minv[c * 4 + r] += lr[r] * oc[c] + t * lr[r] * dc[c];
    }
    *loss = L;
    return ORC_OK;
}

/* ========================================================================
 * Physics: World / ForceRegistry / RigidBody (bvhtracer/src/physics/) and
 * SceneObject::update_transform.  Binary32 in the reference's statement
 * order; cglinalg's quaternion product, quaternion norm s*s + dot(v, v),
 * 3x3 and 4x4 products are written left to right (cglinalg is not vendored).
 * ======================================================================== */

static const double ORC_LN2_HI = 6.93147180369123816490e-01;
static const double ORC_LN2_LO = 1.90821492927058770002e-10;

static uint64_t d2u(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }
static double u2d(uint64_t u) { double x; memcpy(&x, &u, 8); return x; }

/* ln x for x > 0: x = m 2^e, m in [sqrt(1/2), sqrt(2)),
 * ln m = 2 atanh(s) = 2 (s + s^3/3 + ...), s = (m-1)/(m+1). */
static double orc_ln_pos(double x)
{
    const double ln2 = ORC_LN2_HI + ORC_LN2_LO;
    uint64_t b = d2u(x);
    int e = (int)((b >> 52) & 0x7FF);
    if (e == 0) {
        x *= 18014398509481984.0; /* 2^54 */
        b = d2u(x);
        e = (int)((b >> 52) & 0x7FF) - 54;
    }
    e -= 1022;
    double m = u2d((b & 0x000FFFFFFFFFFFFFull) | 0x3FE0000000000000ull);
    if (m < 0.70710678118654752440) {
        m = m * 2.0;
        e -= 1;
    }
    double s = (m - 1.0) / (m + 1.0);
    double s2 = s * s, term = s, sum = s;
    for (int k = 3; k <= 41; k += 2) {
        term = term * s2;
        sum = sum + term / (double)k;
    }
    return (double)e * ln2 + 2.0 * sum;
}

/* exp z, |z| < 700: z = k ln2 + r (Cody-Waite), Taylor series of exp r,
 * scaled by 2^k through the exponent field. */
static double orc_exp_small(double z)
{
    const double ln2 = ORC_LN2_HI + ORC_LN2_LO;
    double kf = z / ln2;
    double k = kf < 0.0 ? -(double)(long long)(0.5 - kf) : (double)(long long)(kf + 0.5);
    double r = (z - k * ORC_LN2_HI) - k * ORC_LN2_LO;
    double p = 1.0;
    for (int i = 24; i >= 1; --i)
        p = 1.0 + (r / (double)i) * p;
    long long ki = (long long)k, k1 = ki / 2, k2 = ki - k1;
    return (p * u2d((uint64_t)(k1 + 1023) << 52)) * u2d((uint64_t)(k2 + 1023) << 52);
}

/* f32::powf (rigid_body.rs:160-161,190): x >= 0; x < 0 gives NaN. */
float orc_powf(float x, float y)
{
    if (y == 0.0f || x == 1.0f)
        return 1.0f;
    if (isnan(x) || isnan(y) || x < 0.0f)
        return NAN;
    if (x == 0.0f)
        return y > 0.0f ? 0.0f : INFINITY;
    if (isinf(x))
        return y > 0.0f ? x : 0.0f;
    if (isinf(y))
        return ((x > 1.0f) == (y > 0.0f)) ? INFINITY : 0.0f;
    double z = (double)y * orc_ln_pos((double)x);
    if (z > 700.0)
        return INFINITY;
    if (z < -800.0)
        return 0.0f;
    return (float)orc_exp_small(z);
}

/* Transform3::transform_point (transform.rs:219-223) */
static void orc_body_point(const float m[16], const float p[3], float out[3])
{
    for (int r = 0; r < 3; ++r)
        out[r] = m[0 + r] * p[0] + m[4 + r] * p[1] + m[8 + r] * p[2] + m[12 + r] * 1.0f;
}

/* RigidBody::calculate_derived_data (rigid_body.rs:120-135) with
 * Transform3::from_translation_rotation (transform.rs:190-217) and
 * _transform_inertia_tensor (rigid_body.rs:28-64). */
static void orc_derived_data(orc_body* b)
{
    float* q = b->orientation;
    float norm = sqrtf(q[0] * q[0] + ((q[1] * q[1] + q[2] * q[2]) + q[3] * q[3]));
    for (int k = 0; k < 4; ++k)
        q[k] = q[k] / norm;
    float s = q[0], x = q[1], y = q[2], z = q[3];
    float one = 1.0f, two = one + one;
    float* t = b->transform; /* t[c*4+r] = matrix[c][r] */
    t[0] = one - two * y * y - two * z * z;
    t[1] = two * x * y + two * s * z;
    t[2] = two * x * z - two * s * y;
    t[3] = 0.0f;
    t[4] = two * x * y - two * s * z;
    t[5] = one - two * x * x - two * z * z;
    t[6] = two * y * z + two * s * x;
    t[7] = 0.0f;
    t[8] = two * x * z + two * s * y;
    t[9] = two * y * z - two * s * x;
    t[10] = one - two * x * x - two * y * y;
    t[11] = 0.0f;
    t[12] = b->position[0];
    t[13] = b->position[1];
    t[14] = b->position[2];
    t[15] = one;
    const float* ib = b->inverse_inertia_tensor; /* ib[c*3+r] = iit_body[c][r] */
    float mm[3][3]; /* mm[c][r] = m{r}{c} of rigid_body.rs:34-48 ("m" + row + col) */
    for (int c = 0; c < 3; ++c)     /* rot_mat row c */
        for (int r = 0; r < 3; ++r) /* iit_body column r */
            mm[c][r] = t[0 * 4 + c] * ib[r * 3 + 0] + t[1 * 4 + c] * ib[r * 3 + 1] + t[2 * 4 + c] * ib[r * 3 + 2];
    /* iit_world[j][i] = m0i*rot[0][j] + m1i*rot[1][j] + m2i*rot[2][j] (rigid_body.rs:50-63) */
    float* w = b->inverse_inertia_tensor_world;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            w[j * 3 + i] = mm[i][0] * t[0 * 4 + j] + mm[i][1] * t[1 * 4 + j] + mm[i][2] * t[2 * 4 + j];
}

/* RigidBody::integrate (rigid_body.rs:138-203). */
static void orc_integrate(orc_body* b, float duration)
{
    if (!b->is_awake)
        return;
    for (int k = 0; k < 3; ++k) {
        b->last_frame_acceleration[k] = b->acceleration[k];
        b->last_frame_acceleration[k] += b->force_accumulator[k] * b->inverse_mass;
    }
    const float* w = b->inverse_inertia_tensor_world;
    float ang[3];
    for (int r = 0; r < 3; ++r)
        ang[r] = w[0 + r] * b->torque_accumulator[0] + w[3 + r] * b->torque_accumulator[1] +
                 w[6 + r] * b->torque_accumulator[2];
    for (int k = 0; k < 3; ++k)
        b->velocity[k] += b->last_frame_acceleration[k] * duration;
    for (int k = 0; k < 3; ++k)
        b->rotation[k] += ang[k] * duration;
    float ld = orc_powf(b->linear_drag, duration), ad = orc_powf(b->angular_drag, duration);
    for (int k = 0; k < 3; ++k)
        b->velocity[k] *= ld;
    for (int k = 0; k < 3; ++k)
        b->rotation[k] *= ad;
    for (int k = 0; k < 3; ++k)
        b->position[k] += b->velocity[k] * duration;
    /* q = Quaternion::from_pure(rotation * duration) * orientation */
    float a[4] = {0.0f, b->rotation[0] * duration, b->rotation[1] * duration, b->rotation[2] * duration};
    float* o = b->orientation;
    float q[4];
    q[0] = a[0] * o[0] - a[1] * o[1] - a[2] * o[2] - a[3] * o[3];
    q[1] = a[0] * o[1] + a[1] * o[0] + a[2] * o[3] - a[3] * o[2];
    q[2] = a[0] * o[2] + a[2] * o[0] + a[3] * o[1] - a[1] * o[3];
    q[3] = a[0] * o[3] + a[3] * o[0] + a[1] * o[2] - a[2] * o[1];
    float half = 0.5f;
    for (int k = 0; k < 4; ++k)
        o[k] += q[k] * half;
    orc_derived_data(b);
    for (int k = 0; k < 3; ++k)
        b->force_accumulator[k] = b->torque_accumulator[k] = 0.0f;
    if (b->can_sleep) {
        const float *v = b->velocity, *r = b->rotation;
        float motion_now = ((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]) + ((r[0] * r[0] + r[1] * r[1]) + r[2] * r[2]);
        float bias = orc_powf(half, duration);
        b->motion = bias * b->motion + (1.0f - bias) * motion_now;
        float threshold = 0.3f, ten = 10.0f;
        if (b->motion < threshold) {
            b->is_awake = 0;
            for (int k = 0; k < 3; ++k)
                b->velocity[k] = b->rotation[k] = 0.0f;
        } else if (b->motion > ten * threshold) {
            b->motion = ten * threshold;
        }
    }
}

void orc_world_start_frame(orc_body* bodies, uint32_t n)
{
    for (uint32_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k)
            bodies[i].force_accumulator[k] = bodies[i].torque_accumulator[k] = 0.0f;
        orc_derived_data(&bodies[i]);
    }
}

void orc_world_run_physics(orc_body* bodies, uint32_t n, const orc_force* forces, uint32_t nf, float duration)
{
    /* ForceRegistry::apply_forces (force_registry.rs:72-77) */
    for (uint32_t j = 0; j < nf; ++j) {
        const orc_force* f = &forces[j];
        orc_body* b = &bodies[f->body];
        if (f->kind == 1) { /* Gravity::apply_force (force_generator.rs:50-55) */
            if (b->inverse_mass >= 0.0f) {
                float mass = b->inverse_mass == 0.0f ? FLT_MAX : 1.0f / b->inverse_mass;
                for (int k = 0; k < 3; ++k)
                    b->force_accumulator[k] += f->v0[k] * mass;
                b->is_awake = 1;
            }
        } else if (f->kind == 2) { /* Spring::apply_force (force_generator.rs:101-120) */
            float lws[3], ows[3], d[3], force[3], pcm[3];
            orc_body_point(b->transform, f->v0, lws);
            orc_body_point(bodies[f->other].transform, f->v1, ows);
            for (int k = 0; k < 3; ++k)
                d[k] = lws[k] - ows[k];
            float len = sqrtf((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
            float magnitude = f->spring_constant * fabsf(len - f->rest_length);
            for (int k = 0; k < 3; ++k)
                force[k] = (d[k] / len) * (-magnitude);
            /* RigidBody::apply_force_at_point (rigid_body.rs:434-439) */
            for (int k = 0; k < 3; ++k)
                pcm[k] = lws[k] - b->position[k];
            for (int k = 0; k < 3; ++k)
                b->force_accumulator[k] += force[k];
            b->torque_accumulator[0] += pcm[1] * force[2] - pcm[2] * force[1];
            b->torque_accumulator[1] += pcm[2] * force[0] - pcm[0] * force[2];
            b->torque_accumulator[2] += pcm[0] * force[1] - pcm[1] * force[0];
            b->is_awake = 1;
        }
    }
    for (uint32_t i = 0; i < n; ++i) /* World::integrate (world.rs:51-55) */
        orc_integrate(&bodies[i], duration);
}

/* 4x4 inverse by the adjugate (cofactor expansion), binary64. */
static void orc_inverse4(const float* mf, float* out)
{
    double m[16], inv[16];
    for (int k = 0; k < 16; ++k)
        m[k] = (double)mf[k];
    inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] +
             m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] -
             m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] +
             m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] -
              m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] -
             m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] +
             m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] -
             m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] +
              m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] +
             m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] -
             m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] +
              m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] -
              m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] -
             m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] +
             m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] -
              m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] +
              m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
    for (int k = 0; k < 16; ++k)
        out[k] = (float)(inv[k] / det);
}

void orc_objects_from_bodies(const orc_body* bodies, const uint32_t* object_body, const float* init, uint32_t n,
                             float* m_out, float* minv_out)
{
    for (uint32_t i = 0; i < n; ++i) {
        if (object_body[i] == 0xFFFFFFFFu)
            continue;
        const float* A = bodies[object_body[i]].transform;
        const float* B = init + 16 * (size_t)i;
        float* M = m_out + 16 * (size_t)i;
        /* Matrix4x4 product: (A B)[c][r] = sum_k A[k][r] B[c][k] */
        for (int c = 0; c < 4; ++c)
            for (int r = 0; r < 4; ++r)
                M[c * 4 + r] = A[0 + r] * B[c * 4 + 0] + A[4 + r] * B[c * 4 + 1] + A[8 + r] * B[c * 4 + 2] +
                               A[12 + r] * B[c * 4 + 3];
        orc_inverse4(M, minv_out + 16 * (size_t)i);
    }
}
