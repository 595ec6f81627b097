// gfx950 kernels of the wavefront path tracer that reproduces Atrc's `pt`.
//
// Pipeline per bounce (one HIP stream, queues compacted with wave-aggregated
// atomics):
//   raygen  -> trace_closest (continuation + MIS-BSDF rays) -> trace_shadow (NEE rays)
//           -> shade (resolve NEE/MIS of the previous vertex, shade the new vertex,
//              emit next rays, Russian roulette) -> ... -> film
// Each path owns fixed slots in the ray/hit/NEE arrays, so queue order never
// changes a result: every path's arithmetic and random numbers are those of
// Atrc's trace_std for the same PCG32 stream.
//
// Reference semantics (paths relative to src/tracer/):
//   trace_std                         src/render/path_tracing.cpp:15-225
//   mis_sample_light / mis_sample_bsdf src/render/direct_illum.cpp:11-192
//   closest / any hit                 src/core/aggregate/{native_aggregate,entity_bvh}.cpp,
//                                     src/core/geometry/{triangle_bvh,quad,triangle,sphere,disk}.cpp
// The file is compiled with -ffp-contract=off: the FP32 operation order is the
// one written here, which is Atrc's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include <cstdint>

#include "../../include/atrc_gpu.h"
#include "agx_kernels.h"

#pragma clang fp contract(off)

namespace agx {

namespace {

constexpr float PI_F = 3.14159265358979323846f;
constexpr float INF_F = __builtin_huge_valf();

// ---------------------------------------------------------------- small vector math
struct V {
    float x, y, z;
};
__device__ __forceinline__ V mk(float x, float y, float z) { return V{x, y, z}; }
__device__ __forceinline__ V mk(float4 a) { return V{a.x, a.y, a.z}; }
__device__ __forceinline__ V mk(const vec4 &a) { return V{a.x, a.y, a.z}; }
__device__ __forceinline__ float4 ld(const vec4 &a) { return make_float4(a.x, a.y, a.z, a.w); }
__device__ __forceinline__ V operator+(V a, V b) { return V{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V operator-(V a, V b) { return V{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V operator*(V a, V b) { return V{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ V operator*(V a, float s) { return V{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V operator/(V a, float s) { return V{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ V operator-(V a) { return V{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ float dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V cross(V a, V b) {
    return V{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float length(V a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ V normalize(V a) { return a / length(a); }
__device__ __forceinline__ bool black(V a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
__device__ __forceinline__ float cosv(V a, V b) { return dot(a, b) / (length(a) * length(b)); }
__device__ __forceinline__ float4 f4(V a, float w) { return make_float4(a.x, a.y, a.z, w); }
// std::min / std::max operand semantics (matter for NaN slab values)
__device__ __forceinline__ float smin(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float smax(float a, float b) { return a < b ? b : a; }
__device__ __forceinline__ uint32_t fbits(float f) { return __float_as_uint(f); }

struct Frame {
    V x, y, z;
};
// Duff et al. 2017 orthonormal basis around z.
__device__ Frame frame_from_z(V zin) {
    const V n = normalize(zin);
    const float sign = copysignf(1.0f, n.z);
    const float a = -1.0f / (sign + n.z);
    const float b = n.x * n.y * a;
    return Frame{mk(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x), mk(b, sign + n.y * n.y * a, -n.y), n};
}
__device__ __forceinline__ V to_global(const Frame &f, V v) { return f.x * v.x + f.y * v.y + f.z * v.z; }
__device__ __forceinline__ V to_local(const Frame &f, V v) { return mk(dot(v, f.x), dot(v, f.y), dot(v, f.z)); }
__device__ __forceinline__ bool pos_hemi(const Frame &f, V v) { return dot(v, f.z) > 0.0f; }
// re-orient a frame to a new z, keeping x as close as possible (Gram-Schmidt)
__device__ Frame rotate_to_new_z(const Frame &f, V new_z) {
    const V nz = normalize(new_z);
    const V xp = f.x - nz * dot(f.x, nz);
    const float l = length(xp);
    if (!(l > 1e-6f)) return frame_from_z(nz);
    const V nx = xp / l;
    return Frame{nx, cross(nz, nx), nz};
}

__device__ __forceinline__ V xf_point(const vec4 *m, V p) {
    return mk(m[0].x * p.x + m[0].y * p.y + m[0].z * p.z + m[0].w, m[1].x * p.x + m[1].y * p.y + m[1].z * p.z + m[1].w,
              m[2].x * p.x + m[2].y * p.y + m[2].z * p.z + m[2].w);
}
__device__ __forceinline__ V xf_vec(const vec4 *m, V v) {
    return mk(m[0].x * v.x + m[0].y * v.y + m[0].z * v.z, m[1].x * v.x + m[1].y * v.y + m[1].z * v.z,
              m[2].x * v.x + m[2].y * v.y + m[2].z * v.z);
}
__device__ __forceinline__ V xf_tvec(const vec4 *m, V v) {
    return mk(m[0].x * v.x + m[1].x * v.y + m[2].x * v.z, m[0].y * v.x + m[1].y * v.y + m[2].y * v.z,
              m[0].z * v.x + m[1].z * v.y + m[2].z * v.z);
}
// frame of a transformed geometry in world space: normal by the inverse transpose,
// x by the linear part, re-orthonormalised
__device__ Frame frame_to_world(const DevEntity &e, const Frame &c) {
    const V nz = normalize(xf_tvec(e.w2l, c.z));
    const V mx = xf_vec(e.l2w, c.x);
    const V xp = mx - nz * dot(mx, nz);
    const float l = length(xp);
    if (!(l > 1e-6f)) return frame_from_z(nz);
    const V nx = xp / l;
    return Frame{nx, cross(nz, nx), nz};
}

// ---------------------------------------------------------------- RNG (PCG32 XSH-RR)
__device__ __forceinline__ unsigned long long splitmix64(unsigned long long x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
struct Rng {
    unsigned long long state, inc;
    __device__ __forceinline__ uint32_t next() {
        const unsigned long long old = state;
        state = old * 6364136223846793005ull + inc;
        const uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = uint32_t(old >> 59u);
        return (xs >> rot) | (xs << ((0u - rot) & 31u));
    }
    __device__ __forceinline__ float uniform() { return float(next() >> 8) * 0x1p-24f; }
};
__device__ __forceinline__ unsigned long long path_seq(unsigned long long pixel, unsigned long long sample) {
    return (pixel << 24) ^ sample;
}
__device__ Rng rng_seed(unsigned long long seed, unsigned long long seq) {
    Rng r;
    r.state = 0;
    r.inc = (seq << 1u) | 1u;
    r.next();
    r.state += splitmix64(seed ^ splitmix64(seq));
    r.next();
    return r;
}

// ---------------------------------------------------------------- queues
// Appends one entry for every active lane with `pred` set: a single atomic per
// wave on the queue counter (per-lane atomics on one address serialise in L2),
// then each lane takes base + its rank among the appending lanes.
__device__ __forceinline__ uint32_t wave_append(uint32_t *counter, bool pred) {
    const unsigned long long mask = __ballot(pred);
    const uint32_t lane = __lane_id();
    const uint32_t rank = uint32_t(__popcll(mask & ((1ull << lane) - 1ull)));
    uint32_t base = 0;
    if (mask) {
        const int leader = __ffsll((long long)mask) - 1;
        if (int(lane) == leader) base = atomicAdd(counter, uint32_t(__popcll(mask)));
        base = __shfl(base, leader);
    }
    return base + rank;
}

// ---------------------------------------------------------------- sampling distributions
__device__ void unit_disk(float u, float v, float *ox, float *oy) {  // Shirley-Chiu concentric map
    const float sx = 2 * u - 1, sy = 2 * v - 1;
    if (sx == 0.0f && sy == 0.0f) { *ox = 0; *oy = 0; return; }
    float r, th;
    if (fabsf(sx) > fabsf(sy)) { r = sx; th = (PI_F / 4) * (sy / sx); }
    else { r = sy; th = PI_F / 2 - (PI_F / 4) * (sx / sy); }
    *ox = r * cosf(th);
    *oy = r * sinf(th);
}
__device__ V cos_hemisphere(float u, float v, float *pdf) {  // Malley's method
    float x, y;
    unit_disk(u, v, &x, &y);
    const float z = sqrtf(smax(0.0f, 1 - x * x - y * y));
    *pdf = z / PI_F;
    return mk(x, y, z);
}
__device__ __forceinline__ float cos_hemisphere_pdf(float c) { return c > 0 ? c / PI_F : 0.0f; }
__device__ V uniform_sphere(float u, float v) {
    const float z = 1 - 2 * u;
    const float r = sqrtf(smax(0.0f, 1 - z * z));
    const float phi = 2 * PI_F * v;
    return mk(r * cosf(phi), r * sinf(phi), z);
}
constexpr float SPHERE_PDF = 1.0f / (4 * PI_F);
__device__ V uniform_cone(float cmax, float u, float v) {
    const float c = (1 - u) + u * cmax;
    const float s = sqrtf(smax(0.0f, 1 - c * c));
    const float phi = 2 * PI_F * v;
    return mk(cosf(phi) * s, sinf(phi) * s, c);
}
__device__ __forceinline__ float cone_pdf(float cmax) { return 1 / (2 * PI_F * (1 - cmax)); }

// ---------------------------------------------------------------- intersection primitives
// closest_intersection_with_triangle (utility/triangle_aux.h:38-69)
__device__ __forceinline__ bool mt_closest(V o, V d, float tmin, float tmax, V A, V BA, V CA, float *t, float *u,
                                           float *v) {
    const V s1 = cross(d, CA);
    const float div = dot(s1, BA);
    if (!(div != 0.0f)) return false;
    const float inv = 1 / div;
    const V oa = o - A;
    const float alpha = dot(oa, s1) * inv;
    if (alpha < 0) return false;
    const V s2 = cross(oa, BA);
    const float beta = dot(d, s2) * inv;
    if (beta < 0 || alpha + beta > 1) return false;
    const float tt = dot(CA, s2) * inv;
    if (!(tmin <= tt && tt <= tmax)) return false;
    *t = tt; *u = alpha; *v = beta;
    return true;
}
// has_intersection_with_triangle (triangle_aux.h:9-30)
__device__ __forceinline__ bool mt_any(V o, V d, float tmin, float tmax, V A, V BA, V CA) {
    const V s1 = cross(d, CA);
    const float div = dot(s1, BA);
    if (!(div != 0.0f)) return false;
    const float inv = 1 / div;
    const V oa = o - A;
    const float alpha = dot(oa, s1) * inv;
    if (alpha < 0 || alpha > 1) return false;
    const V s2 = cross(oa, BA);
    const float beta = dot(d, s2) * inv;
    if (beta < 0 || alpha + beta > 1) return false;
    const float tt = dot(CA, s2) * inv;
    return tmin <= tt && tt <= tmax;
}
// Node::has_intersection (triangle_bvh.cpp:75-97)
__device__ __forceinline__ bool slab(float4 lo, float4 hi, V o, V inv, float tmin, float tmax, float *tenter) {
    const float nx = inv.x * (lo.x - o.x), ny = inv.y * (lo.y - o.y), nz = inv.z * (lo.z - o.z);
    const float fx = inv.x * (hi.x - o.x), fy = inv.y * (hi.y - o.y), fz = inv.z * (hi.z - o.z);
    tmin = smax(tmin, smin(nx, fx));
    tmin = smax(tmin, smin(ny, fy));
    tmin = smax(tmin, smin(nz, fz));
    tmax = smin(tmax, smax(nx, fx));
    tmax = smin(tmax, smax(ny, fy));
    tmax = smin(tmax, smax(nz, fz));
    *tenter = tmin;
    return tmin <= tmax;
}

struct Counts {
    uint32_t visits = 0, tests = 0;
};

struct HitRec {
    float t, u, v;
    int32_t ent, prim;
};

// threads per block of the traversal kernels (the Atrc-mode LDS stack columns hold 256)
constexpr unsigned TRACE_BLOCK = 256;
static_assert(TRACE_BLOCK <= 256 && TRACE_BLOCK % 64 == 0, "traversal block: whole waves, at most the LDS columns");

// Traversal stacks.  TLAS and BLAS share one stack: a BLAS traversal uses the
// entries above the TLAS entries pending at its entity.  LdsStack keeps one
// column per thread in LDS ([depth][thread], conflict-free at equal depth);
// PrivStack is the private-memory fallback for scenes whose BVH depth exceeds
// the LDS column (chosen at upload).
struct LdsStack {
    uint32_t *col;  // &lds[0][threadIdx.x]
    int top;
    __device__ __forceinline__ void push(uint32_t v) { col[top * 256] = v; ++top; }
    __device__ __forceinline__ uint32_t pop() { --top; return col[top * 256]; }
};
struct PrivStack {
    uint32_t s[BLAS_STACK + TLAS_STACK];
    int top;
    __device__ __forceinline__ void push(uint32_t v) { s[top++] = v; }
    __device__ __forceinline__ uint32_t pop() { return s[--top]; }
};

// UntransformedTriangleBVH::closest_intersection (triangle_bvh.cpp:403-488)
template <bool STATS, class Stack>
__device__ bool blas_closest(const DevScene &s, const DevEntity &e, V o, V d, float tmin, float tmax, HitRec *h,
                             Counts &c, Stack &stk) {
    const V inv = mk(1 / d.x, 1 / d.y, 1 / d.z);
    const uint32_t root = e.node_base;
    float te;
    if (STATS) c.visits++;
    if (!slab(ld(s.blas_nodes[2 * root]), ld(s.blas_nodes[2 * root + 1]), o, inv, tmin, tmax, &te)) return false;
    const int bottom = stk.top;
    stk.push(root);
    float best = INF_F, bu = 0, bv = 0;
    uint32_t bslot = 0;
    while (stk.top > bottom) {
        const uint32_t ni = stk.pop();
        const float4 lo = ld(s.blas_nodes[2 * ni]), hi = ld(s.blas_nodes[2 * ni + 1]);
        const uint32_t start = fbits(lo.w), end = fbits(hi.w);
        if (start != INTERIOR) {
            for (uint32_t i = start; i < end; ++i) {
                if (STATS) c.tests++;
                const vec4 p0 = s.blas_prims[3 * i], p1 = s.blas_prims[3 * i + 1], p2 = s.blas_prims[3 * i + 2];
                float tt, uu, vv;
                if (mt_closest(o, d, tmin, tmax, mk(p0), mk(p1), mk(p2), &tt, &uu, &vv)) {
                    best = tt; bu = uu; bv = vv; bslot = i;
                    tmax = tt;
                }
            }
        } else {
            const uint32_t l = ni + 1, r = end;
            float tl, tr;
            if (STATS) c.visits += 2;
            const bool al = slab(ld(s.blas_nodes[2 * l]), ld(s.blas_nodes[2 * l + 1]), o, inv, tmin, tmax, &tl);
            const bool ar = slab(ld(s.blas_nodes[2 * r]), ld(s.blas_nodes[2 * r + 1]), o, inv, tmin, tmax, &tr);
            if (al && ar) {
                if (tl < tr) { stk.push(r); stk.push(l); }
                else { stk.push(l); stk.push(r); }
            } else if (al) {
                stk.push(l);
            } else if (ar) {
                stk.push(r);
            }
        }
    }
    if (isinf(best)) return false;
    h->t = best; h->u = bu; h->v = bv; h->prim = int32_t(bslot);
    return true;
}

// UntransformedTriangleBVH::has_intersection (triangle_bvh.cpp:363-401)
template <bool STATS, class Stack>
__device__ bool blas_any(const DevScene &s, const DevEntity &e, V o, V d, float tmin, float tmax, Counts &c,
                         Stack &stk) {
    const V inv = mk(1 / d.x, 1 / d.y, 1 / d.z);
    const uint32_t root = e.node_base;
    float te;
    if (STATS) c.visits++;
    if (!slab(ld(s.blas_nodes[2 * root]), ld(s.blas_nodes[2 * root + 1]), o, inv, tmin, tmax, &te)) return false;
    const int bottom = stk.top;
    stk.push(root);
    while (stk.top > bottom) {
        const uint32_t ni = stk.pop();
        const float4 lo = ld(s.blas_nodes[2 * ni]), hi = ld(s.blas_nodes[2 * ni + 1]);
        const uint32_t start = fbits(lo.w), end = fbits(hi.w);
        if (start != INTERIOR) {
            for (uint32_t i = start; i < end; ++i) {
                if (STATS) c.tests++;
                const vec4 p0 = s.blas_prims[3 * i], p1 = s.blas_prims[3 * i + 1], p2 = s.blas_prims[3 * i + 2];
                if (mt_any(o, d, tmin, tmax, mk(p0), mk(p1), mk(p2))) { stk.top = bottom; return true; }
            }
        } else {
            const uint32_t l = ni + 1, r = end;
            if (STATS) c.visits += 2;
            if (slab(ld(s.blas_nodes[2 * l]), ld(s.blas_nodes[2 * l + 1]), o, inv, tmin, tmax, &te)) stk.push(l);
            if (slab(ld(s.blas_nodes[2 * r]), ld(s.blas_nodes[2 * r + 1]), o, inv, tmin, tmax, &te)) stk.push(r);
        }
    }
    return false;
}

// sphere_aux.h:10-16
__device__ __forceinline__ void sphere_abc(V o, V d, float r, float *A, float *B, float *C) {
    *A = dot(d, d);
    *B = 2 * dot(d, o);
    *C = dot(o, o) - r * r;
}

// Closest hit of one analytic entity (quad, triangle, sphere, disk); on success
// *tmax is lowered to the hit distance.
template <bool STATS>
__device__ bool analytic_closest(const DevEntity &e, int32_t ei, V o, V d, float tmin, float *tmax, HitRec *h,
                                 Counts &c) {
    switch (e.kind) {
    case AGX_GEOM_QUAD: {  // quad.cpp:30-58
        float t, u, v;
        if (STATS) c.tests++;
        if (mt_closest(o, d, tmin, *tmax, mk(e.a), mk(e.b_a), mk(e.c_a), &t, &u, &v)) {
            h->t = t; h->u = u; h->v = v; h->prim = 0; h->ent = ei; *tmax = t;
            return true;
        }
        if (STATS) c.tests++;
        if (mt_closest(o, d, tmin, *tmax, mk(e.a), mk(e.c_a), mk(e.d_a), &t, &u, &v)) {
            h->t = t; h->u = u; h->v = v; h->prim = 1; h->ent = ei; *tmax = t;
            return true;
        }
        return false;
    }
    case AGX_GEOM_TRIANGLE: {  // triangle.cpp:31-51
        const V lo = xf_point(e.w2l, o), ld = xf_vec(e.w2l, d);
        float t, u, v;
        if (STATS) c.tests++;
        if (!mt_closest(lo, ld, tmin, *tmax, mk(e.a), mk(e.b_a), mk(e.c_a), &t, &u, &v)) return false;
        h->t = t; h->u = u; h->v = v; h->prim = 0; h->ent = ei; *tmax = t;
        return true;
    }
    case AGX_GEOM_SPHERE: {  // sphere_aux.h:66-87
        const V lo = xf_point(e.w2l, o), ld = xf_vec(e.w2l, d);
        float A, B, C;
        sphere_abc(lo, ld, e.a.x, &A, &B, &C);
        float delta = B * B - 4 * A * C;
        if (delta < 0) return false;
        delta = sqrtf(delta);
        const float inv2A = 0.5f / A;
        const float t0 = (-B - delta) * inv2A, t1 = (-B + delta) * inv2A;
        if (*tmax < t0 || t1 < tmin) return false;
        const bool not_t0 = t0 < tmin;
        if (not_t0 && t1 > *tmax) return false;
        const float t = not_t0 ? t1 : t0;
        h->t = t; h->u = 0; h->v = 0; h->prim = 0; h->ent = ei; *tmax = t;
        return true;
    }
    case AGX_GEOM_DISK: {  // disk.cpp:63-83
        const V lo = xf_point(e.w2l, o), ld = xf_vec(e.w2l, d);
        const float t = -lo.z / ld.z;
        if (isinf(t) || !(tmin <= t && t <= *tmax)) return false;
        const V p = lo + ld * t;
        const float rad = sqrtf(p.x * p.x + p.y * p.y);
        if (rad > e.a.x) return false;
        h->t = t; h->u = 0; h->v = 0; h->prim = 0; h->ent = ei; *tmax = t;
        return true;
    }
    default:
        return false;
    }
}

// Closest hit of one entity; on success *tmax is lowered to the hit distance.
template <bool STATS, class Stack>
__device__ bool entity_closest(const DevScene &s, int32_t ei, V o, V d, float tmin, float *tmax, HitRec *h,
                               Counts &c, Stack &stk) {
    const DevEntity &e = s.ents[ei];
    if (e.kind != AGX_GEOM_TRIANGLE_BVH) return analytic_closest<STATS>(e, ei, o, d, tmin, tmax, h, c);
    if (!blas_closest<STATS>(s, e, o, d, tmin, *tmax, h, c, stk)) return false;
    h->ent = ei;
    *tmax = h->t;
    return true;
}

template <bool STATS>
__device__ bool analytic_any(const DevEntity &e, V o, V d, float tmin, float tmax, Counts &c) {
    switch (e.kind) {
    case AGX_GEOM_QUAD:  // quad.cpp:24-28
        if (STATS) c.tests++;
        if (mt_any(o, d, tmin, tmax, mk(e.a), mk(e.b_a), mk(e.c_a))) return true;
        if (STATS) c.tests++;
        return mt_any(o, d, tmin, tmax, mk(e.a), mk(e.c_a), mk(e.d_a));
    case AGX_GEOM_TRIANGLE: {
        if (STATS) c.tests++;
        return mt_any(xf_point(e.w2l, o), xf_vec(e.w2l, d), tmin, tmax, mk(e.a), mk(e.b_a), mk(e.c_a));
    }
    case AGX_GEOM_SPHERE: {  // sphere_aux.h:50-64
        const V lo = xf_point(e.w2l, o), ld = xf_vec(e.w2l, d);
        float A, B, C;
        sphere_abc(lo, ld, e.a.x, &A, &B, &C);
        float delta = B * B - 4 * A * C;
        if (delta < 0) return false;
        delta = sqrtf(delta);
        const float inv2A = 0.5f / A;
        const float t0 = (-B - delta) * inv2A, t1 = (-B + delta) * inv2A;
        return (tmin <= t0 && t0 <= tmax) || (tmin <= t1 && t1 <= tmax);
    }
    case AGX_GEOM_DISK: {  // disk.cpp:49-61
        const V lo = xf_point(e.w2l, o), ld = xf_vec(e.w2l, d);
        const float t = -lo.z / ld.z;
        if (isinf(t) || !(tmin <= t && t <= tmax)) return false;
        const V p = lo + ld * t;
        return p.x * p.x + p.y * p.y <= e.a.y;
    }
    default:
        return false;
    }
}

template <bool STATS, class Stack>
__device__ bool entity_any(const DevScene &s, int32_t ei, V o, V d, float tmin, float tmax, Counts &c, Stack &stk) {
    const DevEntity &e = s.ents[ei];
    if (e.kind != AGX_GEOM_TRIANGLE_BVH) return analytic_any<STATS>(e, o, d, tmin, tmax, c);
    return blas_any<STATS>(s, e, o, d, tmin, tmax, c, stk);
}

// AABB::intersect (common.h:384-395) with std::min/max operand order
__device__ __forceinline__ bool tlas_box(float4 lo, float4 hi, V o, V inv, float tmin, float tmax) {
    float te;
    return slab(lo, hi, o, inv, tmin, tmax, &te);
}

// Scene closest hit: NativeAggregate (native_aggregate.cpp:31-45) or EntityBVH
// (entity_bvh.cpp:141-172, left subtree before right, node bound tested on entry).
template <bool STATS, class Stack>
__device__ bool scene_closest(const DevScene &s, V o, V d, float tmin, float tmax, HitRec *h, Counts &c, Stack &stk) {
    bool hit = false;
    stk.top = 0;
    if (s.aggregate == AGX_AGGREGATE_BVH) {
        const V inv = mk(1 / d.x, 1 / d.y, 1 / d.z);
        stk.push(0);
        while (stk.top > 0) {
            const uint32_t ni = stk.pop();
            const float4 lo = ld(s.tlas_nodes[2 * ni]), hi = ld(s.tlas_nodes[2 * ni + 1]);
            if (STATS) c.visits++;
            if (!tlas_box(lo, hi, o, inv, tmin, tmax)) continue;
            const uint32_t a = fbits(lo.w), b = fbits(hi.w);
            if (a & TLAS_LEAF) {
                for (uint32_t i = a & ~TLAS_LEAF; i < b; ++i)
                    hit = entity_closest<STATS>(s, s.tlas_prims[i], o, d, tmin, &tmax, h, c, stk) || hit;
            } else {
                stk.push(b);
                stk.push(a);
            }
        }
    } else {
        for (int32_t i = 0; i < s.n_ent; ++i) hit = entity_closest<STATS>(s, i, o, d, tmin, &tmax, h, c, stk) || hit;
    }
    return hit;
}

template <bool STATS, class Stack>
__device__ bool scene_any(const DevScene &s, V o, V d, float tmin, float tmax, Counts &c, Stack &stk) {
    stk.top = 0;
    if (s.aggregate == AGX_AGGREGATE_BVH) {
        const V inv = mk(1 / d.x, 1 / d.y, 1 / d.z);
        stk.push(0);
        while (stk.top > 0) {
            const uint32_t ni = stk.pop();
            const float4 lo = ld(s.tlas_nodes[2 * ni]), hi = ld(s.tlas_nodes[2 * ni + 1]);
            if (STATS) c.visits++;
            if (!tlas_box(lo, hi, o, inv, tmin, tmax)) continue;
            const uint32_t a = fbits(lo.w), b = fbits(hi.w);
            if (a & TLAS_LEAF) {
                for (uint32_t i = a & ~TLAS_LEAF; i < b; ++i)
                    if (entity_any<STATS>(s, s.tlas_prims[i], o, d, tmin, tmax, c, stk)) return true;
            } else {
                stk.push(b);
                stk.push(a);
            }
        }
        return false;
    }
    for (int32_t i = 0; i < s.n_ent; ++i)
        if (entity_any<STATS>(s, i, o, d, tmin, tmax, c, stk)) return true;
    return false;
}

// ---------------------------------------------------------------- watertight triangle test
// Woop, Benthin and Wald, "Watertight Ray/Triangle Intersection", JCGT 2(1),
// 2013.  The ray is sheared so that it runs along +z through the origin;
// each vertex is projected with the same per-ray arithmetic, so an edge shared
// by two triangles gives both the same edge function with opposite sign (the
// products commute exactly and no FMA contraction is allowed), and the
// decision is consistent.  Edge functions that come out exactly zero are
// recomputed in FP64 from the same projected coordinates.
struct WtRay {
    V o;
    int kx, ky, kz;
    float sx, sy, sz;
};
__device__ __forceinline__ float vcomp(V a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
__device__ __forceinline__ WtRay wt_ray(V o, V d) {
    WtRay r;
    r.o = o;
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    r.kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    r.kx = r.kz == 2 ? 0 : r.kz + 1;
    r.ky = r.kx == 2 ? 0 : r.kx + 1;
    const float dz = vcomp(d, r.kz);
    if (dz < 0) { const int t = r.kx; r.kx = r.ky; r.ky = t; }  // keep the winding
    r.sx = vcomp(d, r.kx) / dz;
    r.sy = vcomp(d, r.ky) / dz;
    r.sz = 1.0f / dz;
    return r;
}
// hit: t and the barycentrics of B (u) and C (v), as the Moller-Trumbore test reports them
__device__ __forceinline__ bool wt_test(const WtRay &r, V A, V B, V C, float tmin, float tmax, float *t, float *u,
                                        float *v) {
    const V a = A - r.o, b = B - r.o, c = C - r.o;
    const float az = vcomp(a, r.kz), bz = vcomp(b, r.kz), cz = vcomp(c, r.kz);
    const float ax = vcomp(a, r.kx) - r.sx * az, ay = vcomp(a, r.ky) - r.sy * az;
    const float bx = vcomp(b, r.kx) - r.sx * bz, by = vcomp(b, r.ky) - r.sy * bz;
    const float cx = vcomp(c, r.kx) - r.sx * cz, cy = vcomp(c, r.ky) - r.sy * cz;
    float U = cx * by - cy * bx;
    float W0 = ax * cy - ay * cx;  // weight of B
    float W1 = bx * ay - by * ax;  // weight of C
    if (U == 0.0f || W0 == 0.0f || W1 == 0.0f) {
        U = float(double(cx) * double(by) - double(cy) * double(bx));
        W0 = float(double(ax) * double(cy) - double(ay) * double(cx));
        W1 = float(double(bx) * double(ay) - double(by) * double(ax));
    }
    if ((U < 0 || W0 < 0 || W1 < 0) && (U > 0 || W0 > 0 || W1 > 0)) return false;
    const float det = U + W0 + W1;
    if (det == 0.0f) return false;
    const float T = U * (r.sz * az) + W0 * (r.sz * bz) + W1 * (r.sz * cz);
    const float inv = 1.0f / det;
    const float tt = T * inv;
    if (!(tmin <= tt && tt <= tmax)) return false;
    *t = tt; *u = W0 * inv; *v = W1 * inv;
    return true;
}

// ---------------------------------------------------------------- BVH4 (production mode)
// Slab test of one child box with Atrc's arithmetic, NaN axes ignored (fminf /
// fmaxf) and the exit distance widened by 2^-21 relative, so that a ray meeting
// a primitive exactly on its bound is never culled.  Returns the entry
// distance, or +INF for a miss.
__device__ __forceinline__ float b4_slab(float lx, float hx, float ly, float hy, float lz, float hz, V o, V inv,
                                         float tmin, float tmax) {
    const float nx = inv.x * (lx - o.x), fx = inv.x * (hx - o.x);
    const float ny = inv.y * (ly - o.y), fy = inv.y * (hy - o.y);
    const float nz = inv.z * (lz - o.z), fz = inv.z * (hz - o.z);
    const float t0 = fmaxf(fmaxf(tmin, fminf(nx, fx)), fmaxf(fminf(ny, fy), fminf(nz, fz)));
    const float t1 = fminf(fminf(tmax, fmaxf(nx, fx)), fminf(fmaxf(ny, fy), fmaxf(nz, fz)));
    return t0 <= t1 * 1.000000477f ? t0 : INF_F;
}

// child bound j of a quantized node: origin + byte j of w times the step
__device__ __forceinline__ float b4_q(float org, uint32_t w, int j, float step) {
    const float m = float((w >> (8 * j)) & 0xffu) * step;
    return org + m;
}

struct B4Node {
    float4 h, cc, qa, qb;  // origin + step exponents, child codes, quantized x / y and z bounds
};
__device__ __forceinline__ B4Node b4_load(const float4 *n) { return B4Node{n[0], n[1], n[2], n[3]}; }

// 64-byte node: decode the four children's (conservative) boxes and slab-test them
__device__ __forceinline__ void b4_children(const B4Node &nd, V o, V inv, float tmin, float tmax, float t[4],
                                            uint32_t k[4]) {
    const float4 h = nd.h, cc = nd.cc, qa = nd.qa, qb = nd.qb;
    const uint32_t eb = __float_as_uint(h.w);
    const float sx = __uint_as_float((eb & 0xffu) << 23), sy = __uint_as_float(((eb >> 8) & 0xffu) << 23),
                sz = __uint_as_float(((eb >> 16) & 0xffu) << 23);
    const uint32_t wlx = __float_as_uint(qa.x), whx = __float_as_uint(qa.y), wly = __float_as_uint(qa.z),
                   why = __float_as_uint(qa.w), wlz = __float_as_uint(qb.x), whz = __float_as_uint(qb.y);
    k[0] = __float_as_uint(cc.x); k[1] = __float_as_uint(cc.y);
    k[2] = __float_as_uint(cc.z); k[3] = __float_as_uint(cc.w);
#pragma unroll
    for (int j = 0; j < 4; ++j)
        t[j] = k[j] != BVH4_EMPTY
                   ? b4_slab(b4_q(h.x, wlx, j, sx), b4_q(h.x, whx, j, sx), b4_q(h.y, wly, j, sy),
                             b4_q(h.y, why, j, sy), b4_q(h.z, wlz, j, sz), b4_q(h.z, whz, j, sz), o, inv, tmin, tmax)
                   : INF_F;
}

// the same for the node at n
__device__ __forceinline__ void b4_children(const float4 *n, V o, V inv, float tmin, float tmax, float t[4],
                                            uint32_t k[4]) {
    b4_children(b4_load(n), o, inv, tmin, tmax, t, k);
}

__device__ __forceinline__ void cswap(float t[4], uint32_t k[4], int a, int b) {
    if (t[b] < t[a]) {
        const float tt = t[a]; t[a] = t[b]; t[b] = tt;
        const uint32_t kk = k[a]; k[a] = k[b]; k[b] = kk;
    }
}

// Per-lane traversal stack whose newest S entries live in an LDS ring and
// whose older entries spill to a private array.  The ring is a
// [slot][TB] array (TB = threads per block), so the lanes of a wave touch consecutive words at
// any mix of depths and never conflict on a bank; an entry reaches private
// memory only while the stack is deeper than S.  A push onto a full ring moves
// the oldest ring entry (the same slot) out to the spill array, and a pop that
// leaves the stack at depth >= S brings it back, so the ring always holds the
// top min(depth, S) entries.  Each lane touches only its own column: no
// barrier is needed.  (Free functions over separate locals, so that the depth
// and the column address stay in registers.)
template <int S, int TB, class E>
__device__ __forceinline__ void ring_push(E *col, E *spill, int &sp, E e) {
    const int slot = sp & (S - 1);
    if (sp >= S) spill[sp - S] = col[slot * TB];
    col[slot * TB] = e;
    ++sp;
}
template <int S, int TB, class E>
__device__ __forceinline__ E ring_pop(E *col, const E *spill, int &sp) {
    --sp;
    const int slot = sp & (S - 1);
    const E e = col[slot * TB];
    if (sp >= S) col[slot * TB] = spill[sp - S];
    return e;
}
static_assert((BVH4_RING_CLOSEST & (BVH4_RING_CLOSEST - 1)) == 0 && (BVH4_RING_ANY & (BVH4_RING_ANY - 1)) == 0,
              "ring sizes are powers of two");

// Closest hit over the BVH4: children visited near-first, pending children
// kept with their entry distance and skipped once a closer hit is known.
// ring = this lane's column of a [BVH4_RING_CLOSEST][TB] LDS array.  The state
// of one lane's traversal: start() takes a ray, step() does one node visit, or
// one leaf and the pops that follow it, and returns true once the traversal is
// over (hit() then tells whether it found a hit, rec the hit).  The stack
// entries below the ring live in the caller's spill array (a separate local,
// so that the state itself stays in registers).
template <bool STATS, int CAP, bool WT, int TB>
struct B4Closest {
    static constexpr int SPILL = CAP - BVH4_RING_CLOSEST;  // (node code, entry distance bits)
    V o, d, inv;
    float tmin, tmax;
    WtRay wr;
    uint32_t code;
    int sp;
    int32_t best_rank;  // tie rank of the entity of the current hit, -1 = no hit yet
    HitRec rec;

    __device__ __forceinline__ void start(V o_, V d_, float tmin_, float tmax_) {
        o = o_; d = d_; tmin = tmin_; tmax = tmax_;
        inv = mk(1 / d.x, 1 / d.y, 1 / d.z);
        if (WT) wr = wt_ray(o, d);
        code = 0; sp = 0; best_rank = -1;
        rec.ent = -1; rec.prim = -1; rec.t = -1; rec.u = 0; rec.v = 0;
    }
    __device__ __forceinline__ bool hit() const { return best_rank >= 0; }

    // The first step, with the root node's rows already loaded (the root is
    // always an interior node): its entered children are pushed farthest
    // first and the nearest becomes the current node.  True if the ray enters
    // none (the traversal is over).
    __device__ __forceinline__ bool visit_root(const B4Node &root, Counts &c, uint2 *ring, uint2 (&spill)[SPILL]) {
        if (STATS) c.visits++;
        float t[4];
        uint32_t k[4];
        b4_children(root, o, inv, tmin, tmax, t, k);
        cswap(t, k, 0, 1); cswap(t, k, 2, 3); cswap(t, k, 0, 2); cswap(t, k, 1, 3); cswap(t, k, 1, 2);
        if (t[0] == INF_F) return true;
        if (t[3] != INF_F)
            ring_push<BVH4_RING_CLOSEST, TB>(ring, spill, sp, make_uint2(k[3], __float_as_uint(t[3])));
        if (t[2] != INF_F)
            ring_push<BVH4_RING_CLOSEST, TB>(ring, spill, sp, make_uint2(k[2], __float_as_uint(t[2])));
        if (t[1] != INF_F)
            ring_push<BVH4_RING_CLOSEST, TB>(ring, spill, sp, make_uint2(k[1], __float_as_uint(t[1])));
        code = k[0];
        return false;
    }

    __device__ __forceinline__ bool step(const DevScene &s, Counts &c, uint2 *ring, uint2 (&spill)[SPILL]) {
        const float4 *N = reinterpret_cast<const float4 *>(s.b4_nodes);
        const float4 *P = reinterpret_cast<const float4 *>(s.b4_prims);
        if (!(code & BVH4_LEAF)) {
            if (STATS) c.visits++;
            float t[4];
            uint32_t k[4];
            b4_children(N + size_t(code) * BVH4_NODE_VEC4, o, inv, tmin, tmax, t, k);
            cswap(t, k, 0, 1); cswap(t, k, 2, 3); cswap(t, k, 0, 2); cswap(t, k, 1, 3); cswap(t, k, 1, 2);
            if (t[0] != INF_F) {
                if (t[3] != INF_F && sp < CAP)
                    ring_push<BVH4_RING_CLOSEST, TB>(ring, spill, sp, make_uint2(k[3], __float_as_uint(t[3])));
                if (t[2] != INF_F && sp < CAP)
                    ring_push<BVH4_RING_CLOSEST, TB>(ring, spill, sp, make_uint2(k[2], __float_as_uint(t[2])));
                if (t[1] != INF_F && sp < CAP)
                    ring_push<BVH4_RING_CLOSEST, TB>(ring, spill, sp, make_uint2(k[1], __float_as_uint(t[1])));
                code = k[0];
                return false;
            }
        } else {
            const uint32_t first = code & BVH4_FIRST_MASK, cnt = ((code >> 27) & 15u) + 1;
            for (uint32_t i = first; i < first + cnt; ++i) {
                // the record's three rows in flight together (the third does not
                // wait for the slot test: every kind of record uses it)
                const float4 p0 = P[3 * size_t(i)], p1 = P[3 * size_t(i) + 1], p2 = P[3 * size_t(i) + 2];
                const int32_t ent = __float_as_int(p0.w);
                const uint32_t slot = __float_as_uint(p1.w);
                if (slot < BVH4_QUAD_ACD || (WT && slot != BVH4_ANALYTIC)) {
                    // a mesh triangle, or (watertight mode) one triangle of a quad
                    if (STATS) c.tests++;
                    float tt, uu, vv;
                    const bool hit_tri = WT ? wt_test(wr, mk(p0), mk(p1), mk(p2), tmin, tmax, &tt, &uu, &vv)
                                            : mt_closest(o, d, tmin, tmax, mk(p0), mk(p1), mk(p2), &tt, &uu, &vv);
                    // an exact tie with another entity's hit goes to the entity that
                    // Atrc's aggregate tests later (its tie rank)
                    const int32_t rank = __float_as_int(p2.w);
                    if (hit_tri && (tt < tmax || ent == rec.ent || rank > best_rank)) {
                        tmax = tt;
                        rec.t = tt; rec.u = uu; rec.v = vv; rec.ent = ent;
                        rec.prim = (WT && slot >= BVH4_QUAD_ACD) ? (slot == BVH4_QUAD_ABC ? 0 : 1) : int32_t(slot);
                        best_rank = rank;
                    }
                } else {
                    // a lower-ranked entity must hit strictly closer: its test runs
                    // against the next float below the current hit distance
                    const int32_t rank = __float_as_int(p2.w);
                    float tb = rank < best_rank ? __uint_as_float(__float_as_uint(tmax) - 1u) : tmax;
                    if (analytic_closest<STATS>(s.ents[ent], ent, o, d, tmin, &tb, &rec, c)) {
                        tmax = tb;
                        best_rank = rank;
                    }
                }
            }
        }
        // next pending child that may still hold a closer hit
        for (;;) {
            if (sp == 0) return true;
            const uint2 e = ring_pop<BVH4_RING_CLOSEST, TB>(ring, spill, sp);
            if (__uint_as_float(e.y) <= tmax) { code = e.x; return false; }
        }
    }
};

template <bool STATS, int CAP, bool WT, int TB>
__device__ bool b4_closest(const DevScene &s, V o, V d, float tmin, float tmax, HitRec *h, Counts &c,
                           uint2 *ring) {
    B4Closest<STATS, CAP, WT, TB> q;
    uint2 spill[B4Closest<STATS, CAP, WT, TB>::SPILL];
    q.start(o, d, tmin, tmax);
    q.rec = *h;
    while (!q.step(s, c, ring, spill)) {
    }
    *h = q.rec;
    return q.hit();
}

// Any hit over the BVH4; ring = this lane's column of a [BVH4_RING_ANY][TB]
// LDS array.  The same shape as B4Closest: step() does one node visit, or one
// leaf and the pop after it, and returns true once the traversal is over
// (occluded then tells whether it found a hit).
template <bool STATS, int CAP, bool WT, int TB>
struct B4Any {
    static constexpr int SPILL = CAP - BVH4_RING_ANY;
    V o, d, inv;
    float tmin, tmax;
    WtRay wr;
    uint32_t code;
    int sp;
    bool occluded;

    __device__ __forceinline__ void start(V o_, V d_, float tmin_, float tmax_) {
        o = o_; d = d_; tmin = tmin_; tmax = tmax_;
        inv = mk(1 / d.x, 1 / d.y, 1 / d.z);
        if (WT) wr = wt_ray(o, d);
        code = 0; sp = 0; occluded = false;
    }

    // Visit of an interior node whose rows are loaded: the first child the ray
    // enters becomes the current node and the others are pushed (true), or
    // none is entered (false: the caller pops).
    __device__ __forceinline__ bool visit(const B4Node &nd, Counts &c, uint32_t *ring, uint32_t (&spill)[SPILL]) {
        if (STATS) c.visits++;
        float t[4];
        uint32_t k[4];
        b4_children(nd, o, inv, tmin, tmax, t, k);
        uint32_t next = BVH4_EMPTY;
        for (int j = 0; j < 4; ++j) {
            if (t[j] == INF_F) continue;
            if (next == BVH4_EMPTY) next = k[j];
            else if (sp < CAP) ring_push<BVH4_RING_ANY, TB>(ring, spill, sp, k[j]);
        }
        if (next == BVH4_EMPTY) return false;
        code = next;
        return true;
    }
    // the next pending node; true once none is left
    __device__ __forceinline__ bool pop_next(uint32_t *ring, const uint32_t (&spill)[SPILL]) {
        if (sp == 0) return true;
        code = ring_pop<BVH4_RING_ANY, TB>(ring, spill, sp);
        return false;
    }

    __device__ __forceinline__ bool step(const DevScene &s, Counts &c, uint32_t *ring, uint32_t (&spill)[SPILL]) {
        const float4 *N = reinterpret_cast<const float4 *>(s.b4_nodes);
        const float4 *P = reinterpret_cast<const float4 *>(s.b4_prims);
        if (!(code & BVH4_LEAF)) {
            if (visit(b4_load(N + size_t(code) * BVH4_NODE_VEC4), c, ring, spill)) return false;
        } else {
            const uint32_t first = code & BVH4_FIRST_MASK, cnt = ((code >> 27) & 15u) + 1;
            for (uint32_t i = first; i < first + cnt; ++i) {
                // the third row is loaded with the others (analytic records
                // leave it unused; triangles need it)
                const float4 p0 = P[3 * size_t(i)], p1 = P[3 * size_t(i) + 1], p2 = P[3 * size_t(i) + 2];
                const uint32_t slot = __float_as_uint(p1.w);
                if (slot < BVH4_QUAD_ACD || (WT && slot != BVH4_ANALYTIC)) {
                    if (STATS) c.tests++;
                    if (WT) {
                        float tt, uu, vv;
                        if (wt_test(wr, mk(p0), mk(p1), mk(p2), tmin, tmax, &tt, &uu, &vv)) { occluded = true; return true; }
                    } else if (mt_any(o, d, tmin, tmax, mk(p0), mk(p1), mk(p2))) {
                        occluded = true;
                        return true;
                    }
                } else if (analytic_any<STATS>(s.ents[__float_as_int(p0.w)], o, d, tmin, tmax, c)) {
                    occluded = true;
                    return true;
                }
            }
        }
        return pop_next(ring, spill);
    }
};

template <bool STATS, int CAP, bool WT, int TB>
__device__ bool b4_any(const DevScene &s, V o, V d, float tmin, float tmax, Counts &c, uint32_t *ring) {
    B4Any<STATS, CAP, WT, TB> q;
    uint32_t spill[B4Any<STATS, CAP, WT, TB>::SPILL];
    q.start(o, d, tmin, tmax);
    while (!q.step(s, c, ring, spill)) {
    }
    return q.occluded;
}

// Traversal variants: Atrc's structures with the private or the LDS stack, or
// the BVH4 with the small or the large stack (LDS ring + private spill), with
// Atrc's triangle test or the watertight one.
enum { TM_ATRC_PRIV = 0, TM_ATRC_LDS = 1, TM_B4_SMALL = 2, TM_B4_LARGE = 3, TM_B4W_SMALL = 4, TM_B4W_LARGE = 5 };

// LDS words per thread of a traversal kernel: the Atrc-mode LDS column, or
// the BVH4 rings (the closest-hit ring of (code, distance) pairs and the
// any-hit ring of codes share the space; k_trace_rays runs either).
constexpr int lds_stack_words(int tm) {
    return tm == TM_ATRC_LDS ? LDS_STACK
                             : (tm >= TM_B4_SMALL ? (2 * BVH4_RING_CLOSEST > BVH4_RING_ANY ? 2 * BVH4_RING_CLOSEST
                                                                                          : BVH4_RING_ANY)
                                                  : 1);
}

// lds = the block's [lds_stack_words(TM)][TB] word array (the Atrc-mode LDS
// column needs TB = 256)
template <bool STATS, int TM, int TB>
__device__ __forceinline__ bool query_closest(const DevScene &s, V o, V d, float tmin, float tmax, HitRec *h,
                                              Counts &c, uint32_t *lds) {
    if constexpr (TM == TM_ATRC_LDS) {
        LdsStack stk{lds + threadIdx.x, 0};
        return scene_closest<STATS>(s, o, d, tmin, tmax, h, c, stk);
    } else if constexpr (TM == TM_ATRC_PRIV) {
        PrivStack stk;
        return scene_closest<STATS>(s, o, d, tmin, tmax, h, c, stk);
    } else if constexpr (TM == TM_B4_SMALL || TM == TM_B4W_SMALL) {
        return b4_closest<STATS, BVH4_STACK_SMALL, TM == TM_B4W_SMALL, TB>(s, o, d, tmin, tmax, h, c,
                                                                        reinterpret_cast<uint2 *>(lds) + threadIdx.x);
    } else {
        return b4_closest<STATS, BVH4_STACK_LARGE, TM == TM_B4W_LARGE, TB>(s, o, d, tmin, tmax, h, c,
                                                                        reinterpret_cast<uint2 *>(lds) + threadIdx.x);
    }
}

template <bool STATS, int TM, int TB>
__device__ __forceinline__ bool query_any(const DevScene &s, V o, V d, float tmin, float tmax, Counts &c,
                                          uint32_t *lds) {
    if constexpr (TM == TM_ATRC_LDS) {
        LdsStack stk{lds + threadIdx.x, 0};
        return scene_any<STATS>(s, o, d, tmin, tmax, c, stk);
    } else if constexpr (TM == TM_ATRC_PRIV) {
        PrivStack stk;
        return scene_any<STATS>(s, o, d, tmin, tmax, c, stk);
    } else if constexpr (TM == TM_B4_SMALL || TM == TM_B4W_SMALL) {
        return b4_any<STATS, BVH4_STACK_SMALL, TM == TM_B4W_SMALL, TB>(s, o, d, tmin, tmax, c, lds + threadIdx.x);
    } else {
        return b4_any<STATS, BVH4_STACK_LARGE, TM == TM_B4W_LARGE, TB>(s, o, d, tmin, tmax, c, lds + threadIdx.x);
    }
}

// The block's traversal stack memory in LDS (one word when the variant has none)
#define AGX_LDS_COLUMN(TM, TB)                                                 \
    __shared__ __align__(16) uint32_t lds_stack_[lds_stack_words(TM) * (TB)]; \
    uint32_t *lds_col = lds_stack_

// ---------------------------------------------------------------- surface reconstruction
struct Surf {
    V pos;
    float u, v;  // texture coordinates
    Frame geo, user;
    V wr;
};

// sphere_aux.h:18-48 (local frame and uv of a point on the sphere)
__device__ void sphere_local(V p, float radius, float eps, Surf *sp) {
    float phi = (p.x == 0.0f && p.y == 0.0f) ? 0.0f : atan2f(p.y, p.x);
    if (phi < 0) phi += 2 * PI_F;
    const float st = smin(1.0f, smax(-1.0f, p.z / radius));
    const float th = asinf(st);
    sp->u = 0.5f / PI_F * phi;
    sp->v = th / PI_F + 0.5f;
    V ex, ey, ez = normalize(p);
    if (fabsf(ez.z - 1) < eps) { ex = mk(1, 0, 0); ey = mk(0, 1, 0); ez = mk(0, 0, 1); }
    else if (fabsf(ez.z + 1) < eps) { ex = mk(0, 1, 0); ey = mk(1, 0, 0); ez = mk(0, 0, -1); }
    else { ex = cross(mk(0, 0, 1), ez); ey = cross(ez, ex); }
    sp->geo = Frame{ex, ey, ez};
    sp->user = sp->geo;
}
// disk.cpp:11-28
__device__ void disk_local(V p, float rad, float radius, Surf *sp) {
    float phi = (p.y == 0.0f && p.x == 0.0f) ? 0.0f : atan2f(p.y, p.x);
    if (phi < 0) phi += 2 * PI_F;
    sp->pos = p;
    sp->u = phi / (2 * PI_F);
    sp->v = rad / radius;
    const V cx = normalize(mk(p.x, p.y, 0));
    const V cz = mk(0, 0, 1);
    sp->geo = Frame{cx, cross(cz, cx), cz};
    sp->user = sp->geo;
}
__device__ void mesh_frame(const DevScene &s, uint32_t slot, float bu, float bv, Surf *sp) {
    const vec4 *in = s.blas_info + size_t(slot) * 6;
    const vec4 i0 = in[0], i1 = in[1], i2 = in[2], i3 = in[3], i4 = in[4], i5 = in[5];
    const V x = mk(i3), z = mk(i4);
    sp->geo = Frame{x, cross(z, x), z};
    // uv = t_a + t_b_a * u + t_c_a * v
    sp->u = (i0.w + i2.w * bu) + i4.w * bv;
    sp->v = (i1.w + i3.w * bu) + i5.x * bv;
    const V user_z = mk(i0) + mk(i1) * bu + mk(i2) * bv;
    sp->user = rotate_to_new_z(sp->geo, user_z);
}

__device__ Surf surface_at_raw(const DevScene &s, const HitRec &h, V o, V d, float eps) {
    const DevEntity &e = s.ents[h.ent];
    Surf sp;
    switch (e.kind) {
    case AGX_GEOM_QUAD: {
        sp.pos = o + d * h.t;
        const V z = mk(e.z);
        if (h.prim == 0) {
            const V x = mk(e.x_abc);
            sp.geo = Frame{x, cross(z, x), z};
            sp.u = (e.t_ab.x + e.t_ab.z * h.u) + e.t_cd.x * h.v;
            sp.v = (e.t_ab.y + e.t_ab.w * h.u) + e.t_cd.y * h.v;
        } else {
            const V x = mk(e.x_acd);
            sp.geo = Frame{x, cross(z, x), z};
            sp.u = (e.t_ab.x + e.t_cd.x * h.u) + e.t_cd.z * h.v;
            sp.v = (e.t_ab.y + e.t_cd.y * h.u) + e.t_cd.w * h.v;
        }
        sp.user = sp.geo;
        sp.wr = -d;
        return sp;
    }
    case AGX_GEOM_TRIANGLE: {
        const V lo = xf_point(e.w2l, o), ld = xf_vec(e.w2l, d);
        const V z = mk(e.z), x = mk(e.x_abc);
        Frame f{x, cross(z, x), z};
        sp.pos = xf_point(e.l2w, lo + ld * h.t);
        sp.u = (e.t_ab.x + e.t_ab.z * h.u) + e.t_cd.x * h.v;
        sp.v = (e.t_ab.y + e.t_ab.w * h.u) + e.t_cd.y * h.v;
        sp.geo = frame_to_world(e, f);
        sp.user = frame_to_world(e, f);
        sp.wr = xf_vec(e.l2w, -ld);
        return sp;
    }
    case AGX_GEOM_SPHERE: {
        const V lo = xf_point(e.w2l, o), ld = xf_vec(e.w2l, d);
        const V p = lo + ld * h.t;
        sphere_local(p, e.a.x, eps, &sp);
        sp.pos = xf_point(e.l2w, p);
        const Frame g = sp.geo;
        sp.geo = frame_to_world(e, g);
        sp.user = frame_to_world(e, g);
        sp.wr = xf_vec(e.l2w, -ld);
        return sp;
    }
    case AGX_GEOM_DISK: {
        const V lo = xf_point(e.w2l, o), ld = xf_vec(e.w2l, d);
        const V p = lo + ld * h.t;
        disk_local(p, sqrtf(p.x * p.x + p.y * p.y), e.a.x, &sp);
        sp.pos = xf_point(e.l2w, p);
        const Frame g = sp.geo;
        sp.geo = frame_to_world(e, g);
        sp.user = frame_to_world(e, g);
        sp.wr = xf_vec(e.l2w, -ld);
        return sp;
    }
    default: {
        sp.pos = o + d * h.t;
        mesh_frame(s, uint32_t(h.prim), h.u, h.v, &sp);
        sp.wr = -d;
        return sp;
    }
    }
}

__device__ __forceinline__ Frame flip(const Frame &f) { return Frame{-f.x, -f.y, -f.z}; }

// The surface point of a hit; a double-sided entity hit from the back turns its
// geometry and user frames around (DoubleSidedGeometry::closest_intersection,
// double_sided.cpp:26-38)
__device__ __forceinline__ Surf surface_at(const DevScene &s, const HitRec &h, V o, V d, float eps) {
    Surf sp = surface_at_raw(s, h, o, d, eps);
    if ((s.ents[h.ent].flags & EF_DOUBLE_SIDED) && dot(sp.geo.z, sp.wr) < 0) {
        sp.geo = flip(sp.geo);
        sp.user = flip(sp.user);
    }
    return sp;
}

__device__ __forceinline__ V eps_offset(const Surf &sp, V dir, float eps) {  // intersection.h:17-22
    if (dot(dir, sp.geo.z) > 0) return sp.pos + sp.geo.z * eps;
    return sp.pos - sp.geo.z * eps;
}

// ---------------------------------------------------------------- lights
__device__ __forceinline__ V sky(const DevScene &s, V d) {  // native_sky.cpp:13-20
    const float ct = smin(1.0f, smax(-1.0f, normalize(d).z));
    const float t = (ct + 1) / 2;
    return mk(s.sky_top) * t + mk(s.sky_bottom) * (1 - t);
}
// plane through centre + R d with normal d, met by ref + t d
__device__ __forceinline__ V env_emit_pos(const DevScene &s, V ref, V dir) {
    const V d = normalize(dir);
    const float t = s.world.w + dot(mk(s.world) - ref, d);
    return ref + d * t;
}

// Geometry::pdf(ref, pos) in area measure
__device__ float geom_pdf(const DevEntity &e, V ref) {
    if (e.kind == AGX_GEOM_SPHERE) {  // sphere.cpp:131-143
        const V lr = xf_point(e.w2l, ref);
        const float d = length(lr);
        if (d <= e.a.x) return 1 / e.emit.w;
        const float ct = smin(e.a.x / d, 1.0f);
        return cone_pdf(ct) / (e.a.z * e.a.z);
    }
    return 1 / e.emit.w;
}

// GeometryToDiffuseLight::pdf (geometry_to_diffuse_light.cpp:75-90); a
// double-sided geometry halves its pdf (double_sided.cpp:89-99)
__device__ float area_light_pdf(const DevEntity &e, V ref, V pos, V nor) {
    if (dot(nor, ref - pos) <= 0) return 0;
    float ap = geom_pdf(e, ref);
    if (e.flags & EF_DOUBLE_SIDED) ap = ap * 0.5f;
    const V s2r = ref - pos;
    const float dist2 = dot(s2r, s2r);
    const float ac = fabsf(cosv(nor, s2r));
    return ap * (dist2 / ac);
}

struct LightSmp {
    V pos, nor;
    float pdf;  // solid angle; 0 = no sample
};

// Geometry::sample(ref) per kind + GeometryToDiffuseLight::sample (:14-31)
__device__ LightSmp area_light_sample(const DevScene &s, const DevEntity &e, V ref, float su, float sv, float sw,
                                      float eps) {
    LightSmp ls;
    ls.pdf = 0;
    V pos, z;
    float pdf_area;
    // a double-sided geometry picks its side with the first number and reuses
    // it, rescaled, for the wrapped geometry (double_sided.cpp:66-87)
    const bool two = (e.flags & EF_DOUBLE_SIDED) != 0;
    const bool back = two && su < 0.5f;
    if (two) su = back ? 2 * su : 2 * (su - 0.5f);
    switch (e.kind) {
    case AGX_GEOM_QUAD: {  // quad.cpp:80-105
        const float sq = sqrtf(su);
        const float b1 = 1 - sq, b2 = sv * sq;
        if (sw < e.sample_prob.x) pos = mk(e.a) + mk(e.b_a) * b1 + mk(e.c_a) * b2;
        else pos = mk(e.a) + mk(e.c_a) * b1 + mk(e.d_a) * b2;
        z = mk(e.z);
        pdf_area = 1 / e.emit.w;
        break;
    }
    case AGX_GEOM_TRIANGLE: {  // triangle.cpp:72-87
        const float sq = sqrtf(su);
        const float b1 = 1 - sq, b2 = sv * sq;
        const V x = mk(e.x_abc), zl = mk(e.z);
        const Frame f = frame_to_world(e, Frame{x, cross(zl, x), zl});
        pos = xf_point(e.l2w, mk(e.a) + mk(e.b_a) * b1 + mk(e.c_a) * b2);
        z = f.z;
        pdf_area = 1 / e.emit.w;
        break;
    }
    case AGX_GEOM_SPHERE: {  // sphere.cpp:93-123
        const float radius = e.a.x, wr = e.a.z;
        const V lref = xf_point(e.w2l, ref);
        const float d = length(lref);
        V p;
        if (d <= radius) {
            p = uniform_sphere(su, sv) * radius;
            pdf_area = SPHERE_PDF / (wr * wr);
        } else {
            const float ct = smin(radius / d, 1.0f);
            const V dir = uniform_cone(ct, su, sv);
            p = normalize(to_global(frame_from_z(lref), dir)) * radius;
            pdf_area = cone_pdf(ct) / (wr * wr);
        }
        Surf sp;
        sphere_local(p, radius, eps, &sp);
        pos = xf_point(e.l2w, p);
        z = frame_to_world(e, sp.geo).z;
        break;
    }
    case AGX_GEOM_DISK: {  // disk.cpp:100-110
        float dx, dy;
        unit_disk(su, sv, &dx, &dy);
        const float px = dx * e.a.x, py = dy * e.a.x;
        Surf sp;
        disk_local(mk(px, py, 0), sqrtf(px * px + py * py), e.a.x, &sp);
        pos = xf_point(e.l2w, mk(px, py, 0));
        z = frame_to_world(e, sp.geo).z;
        pdf_area = 1 / e.emit.w;
        break;
    }
    default: {  // UntransformedTriangleBVH::sample (triangle_bvh.cpp:495-518), alias table over areas
        const int32_t n = int32_t(e.prim_count);
        const float nu = su * float(n);
        int32_t idx = int32_t(nu);
        if (idx >= n) idx = n - 1;
        const float frac = nu - float(idx);
        const int32_t pick = frac < s.alias_prob[e.alias_base + idx] ? idx : s.alias_idx[e.alias_base + idx];
        const uint32_t slot = e.prim_base + uint32_t(pick);
        const float sq = sqrtf(sv);
        const float b1 = 1 - sq, b2 = sw * sq;
        pos = mk(s.blas_prims[3 * slot]) + mk(s.blas_prims[3 * slot + 1]) * b1 + mk(s.blas_prims[3 * slot + 2]) * b2;
        z = mk(s.blas_info[size_t(slot) * 6 + 4]);
        pdf_area = 1 / e.emit.w;
        break;
    }
    }
    if (two) {
        pdf_area = pdf_area * 0.5f;
        if (back) z = -z;
    }
    if (dot(z, ref - pos) <= 0) return ls;
    const V s2r = ref - pos;
    const float dist2 = dot(s2r, s2r);
    ls.pdf = pdf_area * dist2 / fabsf(cosv(z, s2r));
    ls.pos = pos;
    ls.nor = z;
    return ls;
}

// ---------------------------------------------------------------- BSDFs
struct Bsdf {
    int32_t kind;
    Frame geo, shd;
    V albedo, coef, color, eta2, k2, c_refl, c_refr;
    float ax, ay, ior;
    float eps;
    const DevMaterial *m;  // Disney parameters are read from the record when needed
};

struct BSmp {
    V dir, f;
    float pdf;
    bool delta;
};

__device__ Bsdf make_bsdf(const DevMaterial &m, const Surf &sp, float eps) {
    Bsdf b;
    b.kind = m.kind;
    b.eps = eps;
    b.m = &m;
    b.geo = sp.geo;
    b.shd = sp.user;
    b.albedo = mk(m.albedo);
    b.coef = mk(m.coef);
    b.color = mk(m.color);
    b.eta2 = mk(m.eta2);
    b.k2 = mk(m.k2);
    b.c_refl = mk(m.c_refl);
    b.c_refr = mk(m.c_refr);
    b.ax = m.misc.x;
    b.ay = m.misc.y;
    b.ior = m.misc.z;
    return b;
}

// ConductorPoint::eval (fresnel_point.h:44-64), one channel
__device__ __forceinline__ float conductor1(float ci, float e2, float ek2) {
    const float cos2 = ci * ci;
    const float sin2 = smax(0.0f, 1 - cos2);
    const float t0 = e2 - ek2 - sin2;
    const float a2b2 = sqrtf(t0 * t0 + 4.0f * e2 * ek2);
    const float t1 = a2b2 + cos2;
    const float a = sqrtf(0.5f * (a2b2 + t0));
    const float t2 = 2 * ci * a;
    const float rs = (t1 - t2) / (t1 + t2);
    const float t3 = cos2 * a2b2 + sin2 * sin2;
    const float t4 = t2 * sin2;
    const float rp = rs * (t3 - t4) / (t3 + t4);
    return 0.5f * (rp + rs);
}
__device__ __forceinline__ V conductor(float ci, V e2, V ek2) {
    if (ci <= 0) return mk(0, 0, 0);
    return mk(conductor1(ci, e2.x, ek2.x), conductor1(ci, e2.y, ek2.y), conductor1(ci, e2.z, ek2.z));
}
// refl_aux::dielectric_fresnel (utility/reflection.h:48-72)
__device__ float dielectric(float eta_i, float eta_o, float ci) {
    if (ci < 0) { const float t = eta_i; eta_i = eta_o; eta_o = t; ci = -ci; }
    const float si = sqrtf(smax(0.0f, 1 - ci * ci));
    const float st = eta_o / eta_i * si;
    if (st >= 1) return 1;
    const float ct = sqrtf(smax(0.0f, 1 - st * st));
    const float para = (eta_i * ci - eta_o * ct) / (eta_i * ci + eta_o * ct);
    const float perp = (eta_o * ci - eta_i * ct) / (eta_o * ci + eta_i * ct);
    return 0.5f * (para * para + perp * perp);
}
__device__ __forceinline__ float normal_corr(V g, V sh, V wi, float eps) {  // common.h:239-244
    const float dem = fabsf(cosv(g, wi));
    return dem < eps ? 1.0f : fabsf(cosv(sh, wi) / dem);
}
__device__ __forceinline__ float phi_of(V w) {  // common.h:220-226
    if (w.y == 0.0f && w.x == 0.0f) return 0;
    const float r = atan2f(w.y, w.x);
    return r < 0 ? r + 2 * PI_F : r;
}
__device__ __forceinline__ float tan_theta(V w) {
    const float t = 1 - w.z * w.z;
    if (t <= 0) return 0;
    return sqrtf(t) / w.z;
}
__device__ __forceinline__ float smith(float cp, float sp, float tt, float ax, float ay) {  // microfacet.cpp:54-61
    const float a = ax * cp, b = ay * sp;
    const float t = a * a + b * b;
    const float sv = 1 + t * (tt * tt);
    const float lambda = -0.5f + 0.5f * sqrtf(sv);
    return 1 / (1 + lambda);
}
// GGX reflection eval/pdf (microfacet_refl_comp.cpp:8-65)
__device__ void ggx_eval_pdf(const Bsdf &b, V lwi, V lwo, V *f, float *pdf) {
    if (lwi.z <= 0 || lwo.z <= 0) {
        if (f) *f = mk(0, 0, 0);
        if (pdf) *pdf = 0;
        return;
    }
    const float cti = lwi.z, cto = lwo.z;
    const V lwh = normalize(lwi + lwo);
    const float ctd = dot(lwi, lwh);
    const float ph = phi_of(lwh);
    const float sph = sinf(ph), cph = cosf(ph);
    const float cth = lwh.z;
    const float sth = sqrtf(smax(0.0f, 1 - cth * cth));
    const float qa = cph / b.ax, qb = sph / b.ay;
    const float A = qa * qa + qb * qb;
    const float RD = sth * sth * A + cth * cth;
    const float D = 1 / (PI_F * b.ax * b.ay * (RD * RD));
    const float pi_ = phi_of(lwi), po = phi_of(lwo);
    const float Gi = smith(cosf(pi_), sinf(pi_), tan_theta(lwi), b.ax, b.ay);
    const float Go = smith(cosf(po), sinf(po), tan_theta(lwo), b.ax, b.ay);
    if (f) {
        const V F = conductor(ctd, b.eta2, b.k2) * b.color;
        const float G = Gi * Go;
        *f = F * D * G / fabsf(4 * cti * cto);
    }
    if (pdf) *pdf = Go * D / (4 * lwo.z);
}
// microfacet::sample_anisotropic_gtr2_vnor (microfacet.cpp:104-128)
__device__ V ggx_vndf(const Bsdf &b, V ve, float su, float sv, float eps) {
    const V vh = normalize(mk(b.ax * ve.x, b.ay * ve.y, ve.z));
    const float lensq = vh.x * vh.x + vh.y * vh.y;
    const V t1 = lensq > eps ? mk(-vh.y, vh.x, 0) / sqrtf(lensq) : mk(1, 0, 0);
    const V t2 = cross(vh, t1);
    const float r = sqrtf(su);
    const float phi = 2 * PI_F * sv;
    const float t_1 = r * cosf(phi);
    const float _t_2 = r * sinf(phi);
    const float s = 0.5f * (1 + vh.z);
    const float t_2 = (1 - s) * sqrtf(1 - t_1 * t_1) + s * _t_2;
    const V nh = t1 * t_1 + t2 * t_2 + vh * sqrtf(smax(0.0f, 1 - t_1 * t_1 - t_2 * t_2));
    return normalize(mk(b.ax * nh.x, b.ay * nh.y, smax(0.0f, nh.z)));
}

__device__ __forceinline__ bool fringe(const Bsdf &b, V w) { return pos_hemi(b.shd, w) != pos_hemi(b.geo, w); }

// ---------------------------------------------------------------- Disney (material/disney.cpp:18-457)
__device__ __forceinline__ float sqr(float x) { return x * x; }
__device__ __forceinline__ float aniso_gtr2(float sph, float cph, float sth, float cth, float ax, float ay) {
    const float A = sqr(cph / ax) + sqr(sph / ay);
    const float RD = sqr(sth) * A + sqr(cth);
    return 1 / (PI_F * ax * ay * sqr(RD));
}
__device__ __forceinline__ float smith_aniso(float cp, float sp, float ax, float ay, float tt) {
    const float t = sqr(ax * cp) + sqr(ay * sp);
    const float sv = 1 + t * sqr(tt);
    const float lambda = -0.5f + 0.5f * sqrtf(sv);
    return 1 / (1 + lambda);
}
__device__ __forceinline__ float gtr1(float sth, float cth, float a) {  // microfacet.cpp:84-90
    const float U = sqr(a) - 1;
    const float LD = 2 * PI_F * logf(a);
    const float RD = sqr(a * cth) + sqr(sth);
    return U / (LD * RD);
}
__device__ __forceinline__ float smith_gtr2(float tt, float a) {  // microfacet.cpp:23-29
    if (tt == 0.0f) return 1;
    const float root = a * tt;
    return 2 / (1 + sqrtf(1 + root * root));
}
__device__ V sample_aniso_gtr2(float ax, float ay, float u, float v) {  // microfacet.cpp:63-82
    float sph = ay * sinf(2 * PI_F * u);
    float cph = ax * cosf(2 * PI_F * u);
    const float nor = 1 / sqrtf(sqr(sph) + sqr(cph));
    sph *= nor;
    cph *= nor;
    const float A = sqr(cph / ax) + sqr(sph / ay);
    const float cth = sqrtf(A * (1 - v) / ((1 - A) * v + A));
    const float sth = sqrtf(smax(0.0f, 1 - sqr(cth)));
    return normalize(mk(sth * cph, sth * sph, cth));
}
__device__ V sample_gtr1(float a, float u, float v) {  // microfacet.cpp:92-102
    const float phi = 2 * PI_F * u;
    const float ct = sqrtf((powf(a, 2 - 2 * v) - 1) / (sqr(a) - 1));
    const float st = sqrtf(smax(0.0f, 1 - ct * ct));
    return normalize(mk(st * cosf(phi), st * sinf(phi), ct));
}
__device__ __forceinline__ float one_minus_5(float x) {
    const float t = 1 - x;
    const float t2 = t * t;
    return t2 * t2 * t;
}
__device__ __forceinline__ float cos_2_sin(float c) { return sqrtf(smax(0.0f, 1 - c * c)); }
__device__ __forceinline__ V mixv(V a, V b, float t) { return a * (1 - t) + b * t; }  // math::mix
__device__ __forceinline__ bool zerov(V v) { return v.x == 0.0f && v.y == 0.0f && v.z == 0.0f; }

__device__ void d_angles(V lwi, V lwo, float tax, float tay, V lwh, float *D, float *G) {
    const float ph = phi_of(lwh);
    const float sph = sinf(ph), cph = cosf(ph);
    const float cth = lwh.z, sth = cos_2_sin(cth);
    *D = aniso_gtr2(sph, cph, sth, cth, tax, tay);
    const float pi_ = phi_of(lwi), po = phi_of(lwo);
    const float spi = sinf(pi_), cpi = cosf(pi_);
    const float spo = sinf(po), cpo = cosf(po);
    *G = smith_aniso(cpi, spi, tax, tay, tan_theta(lwi)) * smith_aniso(cpo, spo, tax, tay, tan_theta(lwo));
}

__device__ V disney_eval(const Bsdf &b, V wi, V wo) {
    if (fringe(b, wi) || fringe(b, wo)) {  // bsdf.h:133-139
        if (!pos_hemi(b.geo, wi) || !pos_hemi(b.geo, wo)) return mk(0, 0, 0);
        return b.albedo / PI_F;
    }
    const DevMaterial &m = *b.m;
    const V lwi = normalize(to_local(b.shd, wi));
    const V lwo = normalize(to_local(b.shd, wo));
    if (fabsf(lwi.z) < b.eps || fabsf(lwo.z) < b.eps) return mk(0, 0, 0);
    const float metallic = m.d0.x, trans = m.d1.y, ior = m.misc.z;
    const V C = b.albedo;
    if (lwi.z * lwo.z < 0) {  // f_trans (disney.cpp:114-163)
        if (trans == 0.0f) return mk(0, 0, 0);
        const float ci = lwi.z, co = lwo.z;
        const float eta = co > 0 ? ior : 1 / ior;
        V lwh = normalize(lwo + lwi * eta);
        if (lwh.z < 0) lwh = -lwh;
        const float cd = dot(lwo, lwh);
        const float F = dielectric(ior, 1, cd);
        float D, G;
        d_angles(lwi, lwo, m.d1.z, m.d1.w, lwh, &D, &G);
        const float sdem = cd + eta * dot(lwi, lwh);
        const float corr = 1 / eta;
        const float val = (1 - F) * D * G * eta * eta * dot(lwi, lwh) * dot(lwo, lwh) * corr * corr /
                          (ci * co * sdem * sdem);
        const float tf = co > 0 ? trans : 1.0f;
        const V sq = mk(sqrtf(C.x), sqrtf(C.y), sqrtf(C.z));
        return sq * ((1 - metallic) * tf) * fabsf(val) * normal_corr(b.geo.z, b.shd.z, wi, b.eps);
    }
    if (lwi.z < 0 && lwo.z < 0) {  // f_inner_refl (disney.cpp:165-198)
        if (trans == 0.0f) return mk(0, 0, 0);
        const V lwh = -normalize(lwi + lwo);
        const float cd = dot(lwo, lwh);
        const float F = dielectric(ior, 1, cd);
        float D, G;
        d_angles(lwi, lwo, m.d1.z, m.d1.w, lwh, &D, &G);
        return C * trans * fabsf(F * D * G / (4 * lwi.z * lwo.z)) * normal_corr(b.geo.z, b.shd.z, wi, b.eps);
    }
    if (lwi.z <= 0 || lwo.z <= 0) return mk(0, 0, 0);
    const float ci = lwi.z, co = lwo.z;
    const V lwh = normalize(lwi + lwo);
    const float cd = dot(lwi, lwh);
    const V tint = mk(m.c_refl);
    V diffuse = mk(0, 0, 0), sheen = mk(0, 0, 0);
    if (metallic < 1) {  // f_diffuse, f_sheen (disney.cpp:77-96)
        const V f_lambert = C / PI_F;
        const float FL = one_minus_5(ci), FV = one_minus_5(co);
        const float RR = 2 * m.misc.w * cd * cd;
        const V retro = C / PI_F * RR * (FL + FV + FL * FV * (RR - 1));
        diffuse = f_lambert * (1 - 0.5f * FL) * (1 - 0.5f * FV) + retro;
        if (m.d0.z > 0) sheen = mixv(mk(1, 1, 1), tint, m.d0.w) * (4 * m.d0.z) * one_minus_5(cd);
    }
    V specular;
    {  // f_specular (disney.cpp:200-236)
        const V cspec = mixv(mixv(mk(1, 1, 1), tint, m.d0.y), C, metallic);
        const V diel = cspec * dielectric(ior, 1, cd);
        const V cond = cspec + (mk(1, 1, 1) - cspec) * one_minus_5(cd);
        const V F = mixv(mk(m.c_refr) * diel, cond, metallic);
        float D, G;
        d_angles(lwi, lwo, b.ax, b.ay, lwh, &D, &G);
        specular = F * D * G / fabsf(4 * ci * co);
    }
    V clearcoat = mk(0, 0, 0);
    if (m.d1.x > 0) {  // f_clearcoat (disney.cpp:98-112)
        const float cth = lwh.z;
        const float D = gtr1(cos_2_sin(cth), cth, m.d2.x);
        const float F = 0.04f + (1 - 0.04f) * one_minus_5(cd);
        const float G = smith_gtr2(tan_theta(lwi), 0.25f) * smith_gtr2(tan_theta(lwo), 0.25f);
        const float c = m.d1.x * D * F * G / fabsf(4 * ci * co);
        clearcoat = mk(c, c, c);
    }
    const V value = (diffuse + sheen) * ((1 - metallic) * (1 - trans)) + specular + clearcoat;
    return value * normal_corr(b.geo.z, b.shd.z, wi, b.eps);
}

__device__ float d_pdf_trans(const DevMaterial &m, V lwi, V lwo) {  // disney.cpp:350-377
    const float ior = m.misc.z;
    const float eta = lwo.z > 0 ? ior : 1 / ior;
    V lwh = normalize(lwo + lwi * eta);
    if (lwh.z < 0) lwh = -lwh;
    if (((lwo.z > 0) != (dot(lwh, lwo) > 0)) || ((lwi.z > 0) != (dot(lwh, lwi) > 0))) return 0;
    const float sdem = dot(lwo, lwh) + eta * dot(lwi, lwh);
    const float dwh_to_dwi = eta * eta * dot(lwi, lwh) / (sdem * sdem);
    const float ph = phi_of(lwh);
    const float cth = lwh.z;
    const float D = aniso_gtr2(sinf(ph), cosf(ph), cos_2_sin(cth), cth, m.d1.z, m.d1.w);
    return fabsf(cth * D * dwh_to_dwi);
}
__device__ float d_pdf_inner(const DevMaterial &m, V lwi, V lwo) {  // disney.cpp:379-395
    const V lwh = -normalize(lwi + lwo);
    const float ph = phi_of(lwh);
    const float cth = lwh.z;
    const float cd = dot(lwi, lwh);
    const float D = aniso_gtr2(sinf(ph), cosf(ph), cos_2_sin(cth), cth, m.d1.z, m.d1.w);
    return fabsf(cth * D / (4 * cd));
}

__device__ float disney_pdf(const Bsdf &b, V wi, V wo) {  // disney.cpp:608-644
    if (fringe(b, wi) || fringe(b, wo)) {
        if (pos_hemi(b.geo, wi) != pos_hemi(b.geo, wo)) return 0;
        return cos_hemisphere_pdf(normalize(to_local(b.geo, wi)).z);
    }
    const DevMaterial &m = *b.m;
    const V lwi = normalize(to_local(b.shd, wi));
    const V lwo = normalize(to_local(b.shd, wo));
    if (fabsf(lwi.z) < b.eps || fabsf(lwo.z) < b.eps) return 0;
    if (lwo.z < 0) {
        if (m.d1.y == 0.0f) return 0;
        const float mF = smin(0.9f, smax(0.1f, dielectric(m.misc.z, 1, lwo.z)));
        if (lwi.z > 0) return (1 - mF) * d_pdf_trans(m, lwi, lwo);
        if (mF == 0.0f) return 0;
        return mF * d_pdf_inner(m, lwi, lwo);
    }
    if (lwi.z < 0) return m.d2.z * d_pdf_trans(m, lwi, lwo);
    const float diffuse = cos_hemisphere_pdf(lwi.z);
    // pdf_specular_clearcoat (disney.cpp:317-348): sin(phi_o) taken as |sin|, as Atrc does
    const V lwh = normalize(lwi + lwo);
    const float ph = phi_of(lwh);
    const float sph = sinf(ph), cph = cosf(ph);
    const float cth = lwh.z, sth = cos_2_sin(cth);
    const float cd = dot(lwi, lwh);
    const float cpo = cosf(phi_of(lwo));
    const float spo = cos_2_sin(cpo);
    const float spec_D = aniso_gtr2(sph, cph, sth, cth, b.ax, b.ay);
    const float pdf_spec = smith_aniso(cpo, spo, b.ax, b.ay, tan_theta(lwo)) * spec_D / (4 * lwo.z);
    const float cc_D = gtr1(sth, cth, m.d2.x);
    const float pdf_cc = cth * cc_D / (4 * cd);
    return m.d2.y * diffuse + m.d2.w * pdf_spec + m.d3.x * pdf_cc;
}

__device__ V d_sample_trans(const DevMaterial &m, V lwo, float u, float v) {  // disney.cpp:270-294
    const V lwh = sample_aniso_gtr2(m.d1.z, m.d1.w, u, v);
    if (lwh.z <= 0) return mk(0, 0, 0);
    if ((lwo.z > 0) != (dot(lwh, lwo) > 0)) return mk(0, 0, 0);
    const float eta = lwo.z > 0 ? 1 / m.misc.z : m.misc.z;
    const V owh = dot(lwh, lwo) > 0 ? lwh : -lwh;
    // refl_aux::refract (utility/reflection.h:31-40)
    const float ci = dot(lwo, owh);
    const float s2t = sqr(eta) * (1 - sqr(ci));
    const float c2t = 1 - s2t;
    if (c2t <= 0) return mk(0, 0, 0);
    const float ct = sqrtf(c2t);
    const V lwi = normalize(normalize(owh * (eta * ci - ct) - lwo * eta));
    if (lwi.z * lwo.z > 0 || ((lwi.z > 0) != (dot(lwh, lwi) > 0))) return mk(0, 0, 0);
    return lwi;
}

__device__ BSmp disney_sample(const Bsdf &b, V wo, float su, float sv, float sw, float eps) {  // disney.cpp:527-596
    BSmp r;
    r.dir = mk(0, 0, 0);
    r.f = mk(0, 0, 0);
    r.pdf = 0;
    r.delta = false;
    if (fringe(b, wo)) {  // bsdf.h:142-159
        if (!pos_hemi(b.geo, wo)) return r;
        float p;
        const V lwi = cos_hemisphere(su, sv, &p);
        if (p < eps) return r;
        const V wi = normalize(to_global(b.geo, lwi));
        r.dir = wi;
        r.f = b.albedo / PI_F * normal_corr(b.geo.z, b.shd.z, wi, eps);
        r.pdf = p;
        return r;
    }
    const DevMaterial &m = *b.m;
    const V lwo = normalize(to_local(b.shd, wo));
    if (fabsf(lwo.z) < eps) return r;
    V lwi = mk(0, 0, 0);
    if (lwo.z < 0) {
        if (m.d1.y == 0.0f) return r;
        const float mF = smin(0.9f, smax(0.1f, dielectric(m.misc.z, 1, lwo.z)));
        if (su >= mF) {
            lwi = d_sample_trans(m, lwo, sv, sw);
        } else {  // sample_inner_refl (disney.cpp:296-310)
            const V lwh = sample_aniso_gtr2(m.d1.z, m.d1.w, sv, sw);
            if (lwh.z > 0) {
                const V c = lwh * (2 * dot(lwo, lwh)) - lwo;
                if (!(c.z > 0)) lwi = normalize(c);
            }
        }
    } else {
        // a diffuse pick (sel < w_diffuse) samples the cosine lobe, but the
        // if/else chain that follows always overwrites that direction
        float sel = su;
        if (!(sel < m.d2.y)) sel -= m.d2.y;
        if (sel < m.d2.z) {
            lwi = d_sample_trans(m, lwo, sv, sw);
        } else {
            sel -= m.d2.z;
            const V lwh = sel < m.d2.w ? normalize(ggx_vndf(b, lwo, sv, sw, eps)) : sample_gtr1(m.d2.x, sv, sw);
            if (lwh.z > 0) {
                const V c = normalize(lwh * (2 * dot(lwo, lwh)) - lwo);
                if (c.z > 0) lwi = c;
            }
        }
    }
    if (zerov(lwi)) return r;
    const V wi = to_global(b.shd, lwi);
    r.dir = wi;
    r.f = disney_eval(b, wi, wo);
    r.pdf = disney_pdf(b, wi, wo);
    return r;
}

// BSDF::eval (component/aggregate.h:39-64; glass/mirror are delta: 0)
template <uint32_t F = SF_ALL>
__device__ V bsdf_eval(const Bsdf &b, V wi, V wo) {
    if ((F & SF_DISNEY) && b.kind == AGX_MAT_DISNEY) return disney_eval(b, wi, wo);
    if (b.kind != AGX_MAT_IDEAL_DIFFUSE && b.kind != AGX_MAT_METAL) return mk(0, 0, 0);
    if (fringe(b, wi) || fringe(b, wo)) {  // bsdf.h:133-139
        if (!pos_hemi(b.geo, wi) || !pos_hemi(b.geo, wo)) return mk(0, 0, 0);
        return b.albedo / PI_F;
    }
    const V lwi = normalize(to_local(b.shd, wi));
    const V lwo = normalize(to_local(b.shd, wo));
    if (lwi.z == 0.0f || lwo.z == 0.0f) return mk(0, 0, 0);
    if (b.kind == AGX_MAT_IDEAL_DIFFUSE) {
        if (lwi.z <= 0 || lwo.z <= 0) return mk(0, 0, 0);
        return b.coef;
    }
    V f;
    ggx_eval_pdf(b, lwi, lwo, &f, nullptr);
    return f;
}

// BSDF::pdf (aggregate.h:218-246)
template <uint32_t F = SF_ALL>
__device__ float bsdf_pdf(const Bsdf &b, V wi, V wo) {
    if ((F & SF_DISNEY) && b.kind == AGX_MAT_DISNEY) return disney_pdf(b, wi, wo);
    if (b.kind != AGX_MAT_IDEAL_DIFFUSE && b.kind != AGX_MAT_METAL) return 0;
    if (fringe(b, wi) || fringe(b, wo)) {  // bsdf.h:178-185
        if (pos_hemi(b.geo, wi) != pos_hemi(b.geo, wo)) return 0;
        return cos_hemisphere_pdf(normalize(to_local(b.geo, wi)).z);
    }
    const V lwi = normalize(to_local(b.shd, wi));
    const V lwo = normalize(to_local(b.shd, wo));
    if (lwi.z == 0.0f || lwo.z == 0.0f) return 0;
    float p;
    if (b.kind == AGX_MAT_IDEAL_DIFFUSE) p = (lwi.z <= 0 || lwo.z <= 0) ? 0.0f : cos_hemisphere_pdf(lwi.z);
    else ggx_eval_pdf(b, lwi, lwo, nullptr, &p);
    return 1.0f * p / 1.0f;
}

// BSDF::sample (aggregate.h:66-139, glass.cpp:52-91, mirror.cpp:32-50)
template <uint32_t F = SF_ALL>
__device__ BSmp bsdf_sample(const Bsdf &b, V wo, float su, float sv, float sw, float eps) {
    BSmp r;
    r.dir = mk(0, 0, 0);
    r.f = mk(0, 0, 0);
    r.pdf = 0;
    r.delta = false;
    if ((F & SF_DISNEY) && b.kind == AGX_MAT_DISNEY) return disney_sample(b, wo, su, sv, sw, eps);
    if (b.kind == AGX_MAT_IDEAL_BLACK) return r;
    if ((F & SF_DELTA) && b.kind == AGX_MAT_INVISIBLE_SURFACE) {  // invisible_surface.cpp:26-31
        const float cv = fabsf(cosv(b.geo.z, wo));
        const float d = cv < eps ? 1.0f : cv;
        r.dir = mk(-wo.x, -wo.y, -wo.z);
        r.f = mk(1.0f / d, 1.0f / d, 1.0f / d);
        r.pdf = 1;
        r.delta = true;
        return r;
    }
    if ((F & SF_DELTA) && b.kind == AGX_MAT_MIRROR) {
        const V lo = to_local(b.shd, wo);
        if (lo.z <= 0) return r;
        const V nwo = normalize(lo);
        const V dir = to_global(b.shd, mk(0, 0, 2 * nwo.z) - nwo);
        const V f = conductor(nwo.z, b.eta2, b.k2) * b.albedo / fabsf(nwo.z);
        r.dir = dir; r.f = f * normal_corr(b.geo.z, b.shd.z, dir, eps); r.pdf = 1; r.delta = true;
        return r;
    }
    if ((F & SF_DELTA) && b.kind == AGX_MAT_GLASS) {
        const V nwo = normalize(to_local(b.shd, wo));
        const V nor = nwo.z > 0 ? mk(0, 0, 1) : mk(0, 0, -1);
        const float fr = dielectric(b.ior, 1.0f, nwo.z);
        if (su < fr) {
            const V lwi = mk(-nwo.x, -nwo.y, nwo.z);
            const V wi = to_global(b.shd, lwi);
            const V f = b.c_refl * fr / fabsf(lwi.z);
            r.dir = wi; r.f = f * normal_corr(b.geo.z, b.shd.z, wi, eps); r.pdf = fr; r.delta = true;
            return r;
        }
        const float eta_i = nwo.z > 0 ? 1.0f : b.ior;
        const float eta_t = nwo.z > 0 ? b.ior : 1.0f;
        const float eta = eta_i / eta_t;
        const float cti = fabsf(nwo.z);
        const float st2 = eta * eta * smax(0.0f, 1 - cti * cti);
        if (st2 >= 1) return r;
        const float ctt = sqrtf(1 - st2);
        const V nwi = normalize(nor * (eta * cti - ctt) - nwo * eta);
        const float corr = eta * eta;
        const V dir = to_global(b.shd, nwi);
        const V f = b.c_refr * corr * (1 - fr) / fabsf(nwi.z);
        r.dir = dir; r.f = f * normal_corr(b.geo.z, b.shd.z, dir, eps); r.pdf = 1 - fr; r.delta = true;
        return r;
    }
    if (fringe(b, wo)) {  // bsdf.h:142-159
        if (!pos_hemi(b.geo, wo)) return r;
        float p;
        const V lwi = cos_hemisphere(su, sv, &p);
        if (p < eps) return r;
        const V wi = normalize(to_global(b.geo, lwi));
        r.dir = wi; r.f = b.albedo / PI_F * normal_corr(b.geo.z, b.shd.z, wi, eps); r.pdf = p;
        return r;
    }
    const V lwo = normalize(to_local(b.shd, wo));
    if (lwo.z == 0.0f) return r;
    if (!((su - 0.001f) * 1.0f <= 1.0f)) return r;  // single component, weight 1
    V lwi, f;
    float p;
    if (b.kind == AGX_MAT_IDEAL_DIFFUSE) {  // diffuse_comp.cpp:29-45
        if (lwo.z <= 0) return r;
        lwi = cos_hemisphere(sv, sw, &p);
        if (p < 0.001f) return r;
        f = b.coef;
    } else {  // microfacet_refl_comp.cpp:85-100
        if (lwo.z <= 0) return r;
        const V lwh = normalize(ggx_vndf(b, lwo, sv, sw, eps));
        if (lwh.z <= 0) return r;
        lwi = normalize(lwh * (2 * dot(lwo, lwh)) - lwo);
        if (lwi.z <= 0) return r;
        ggx_eval_pdf(b, lwi, lwo, &f, &p);
    }
    lwi = normalize(lwi);
    p *= 1.0f;
    const V wi = to_global(b.shd, lwi);
    r.dir = wi; r.f = f * normal_corr(b.geo.z, b.shd.z, wi, eps); r.pdf = p;
    return r;
}

// ---------------------------------------------------------------- wave-level counter reduction
__device__ __forceinline__ void flush_counts(const Counts &c, bool ok, DevCounters *cnt) {
    uint32_t v = c.visits, t = c.tests, h = ok ? 1u : 0u;
    for (int off = 32; off > 0; off >>= 1) {
        v += __shfl_xor(v, off);
        t += __shfl_xor(t, off);
        h += __shfl_xor(h, off);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&cnt->node_visits, (unsigned long long)v);
        atomicAdd(&cnt->tri_tests, (unsigned long long)t);
        atomicAdd(&cnt->hits, (unsigned long long)h);
    }
}

// ---------------------------------------------------------------- kernels
// The dispatcher deals blocks to the eight XCDs round-robin by block index.
// The traversal kernels map them so that XCD x takes the x-th contiguous
// eighth of the queue: queues are in path (pixel) order, so an XCD's rays come
// from one band of the image and share BVH nodes in that XCD's own L2.  The
// logical block of this physical block for a queue of n rays, or -1; the grid
// holds at least ceil(n / TB) + 7 blocks, so every logical block is covered once.
template <int TB>
__device__ __forceinline__ int32_t xcd_block(uint32_t n) {
    const uint32_t nb = (n + TB - 1u) / TB;
    const uint32_t per = (nb + 7u) / 8u;
    const uint32_t x = blockIdx.x & 7u, j = blockIdx.x >> 3;
    if (j >= per) return -1;
    const uint32_t lb = x * per + j;
    return lb < nb ? int32_t(lb) : -1;
}

// TM = traversal variant (TM_*).  Eight waves per SIMD: the traversal is bound by
// the latency of its node fetches, and the default allocation (66 VGPRs) left
// room for seven; eight fit in 64 without spilling.
template <bool STATS, int TM, int TB>
__global__ void __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(8, 8))) k_trace_closest(DevScene s, PathBuffers pb, DevCounters *cnt) {
    const uint32_t n = pb.cnt_in[0];
    const int32_t lb = xcd_block<TB>(n);
    if (lb < 0) return;  // whole block past the queue
    AGX_LDS_COLUMN(TM, TB);
    const uint32_t tid = uint32_t(lb) * blockDim.x + threadIdx.x;
    Counts c;
    bool found = false;
    if constexpr (TM >= TM_B4_SMALL) {
        if (tid < n) {
            // the root node's rows do not depend on the ray: they are loaded
            // first, while the queue entry and then the ray arrive
            const B4Node root = b4_load(reinterpret_cast<const float4 *>(s.b4_nodes));
            const uint32_t ri = pb.q_closest[tid];
            const float4 o4 = pb.ray_o[ri], d4 = pb.ray_d[ri];
            constexpr int CAP = (TM == TM_B4_LARGE || TM == TM_B4W_LARGE) ? BVH4_STACK_LARGE : BVH4_STACK_SMALL;
            using Q = B4Closest<STATS, CAP, TM == TM_B4W_SMALL || TM == TM_B4W_LARGE, TB>;
            uint2 *ring = reinterpret_cast<uint2 *>(lds_col) + threadIdx.x;
            Q q;
            uint2 spill[Q::SPILL];
            q.start(mk(o4), mk(d4), o4.w, d4.w);
            if (!q.visit_root(root, c, ring, spill))
                while (!q.step(s, c, ring, spill)) {
                }
            found = q.hit();
            pb.hit[ri] = make_float4(q.rec.t, q.rec.u, q.rec.v, __int_as_float(found ? q.rec.ent : -1));
            pb.hit_prim[ri] = q.rec.prim;
        }
    } else if (tid < n) {
        const uint32_t ri = pb.q_closest[tid];
        const float4 o4 = pb.ray_o[ri], d4 = pb.ray_d[ri];
        HitRec h;
        h.ent = -1; h.prim = -1; h.t = -1; h.u = 0; h.v = 0;
        found = query_closest<STATS, TM, TB>(s, mk(o4), mk(d4), o4.w, d4.w, &h, c, lds_col);
        if (!found) h.ent = -1;
        pb.hit[ri] = make_float4(h.t, h.u, h.v, __int_as_float(h.ent));
        pb.hit_prim[ri] = h.prim;
    }
    if (STATS) flush_counts(c, found, cnt);
}

template <bool STATS, int TM, int TB>
__global__ void __launch_bounds__(TB) k_trace_shadow(DevScene s, PathBuffers pb, DevCounters *cnt) {
    const uint32_t n = pb.cnt_in[1];
    const int32_t lb = xcd_block<TB>(n);
    if (lb < 0) return;
    AGX_LDS_COLUMN(TM, TB);
    const uint32_t tid = uint32_t(lb) * blockDim.x + threadIdx.x;
    Counts c;
    bool occ = false;
    if constexpr (TM >= TM_B4_SMALL) {
        if (tid < n) {
            // the root node's rows do not depend on the ray: they are loaded
            // first, while the queue entry and then the ray arrive, and the
            // root is visited from registers
            const B4Node root = b4_load(reinterpret_cast<const float4 *>(s.b4_nodes));
            const uint32_t ri = pb.q_shadow[tid];
            const float4 o4 = pb.sh_o[ri], d4 = pb.sh_d[ri];
            constexpr int CAP = (TM == TM_B4_LARGE || TM == TM_B4W_LARGE) ? BVH4_STACK_LARGE : BVH4_STACK_SMALL;
            using Q = B4Any<STATS, CAP, TM == TM_B4W_SMALL || TM == TM_B4W_LARGE, TB>;
            uint32_t *ring = lds_col + threadIdx.x;
            Q q;
            uint32_t spill[Q::SPILL];
            q.start(mk(o4), mk(d4), d4.w, o4.w);
            bool done = q.visit(root, c, ring, spill) ? false : q.pop_next(ring, spill);
            while (!done) done = q.step(s, c, ring, spill);
            occ = q.occluded;
            pb.occluded[ri] = occ ? 1 : 0;
        }
    } else if (tid < n) {
        const uint32_t ri = pb.q_shadow[tid];
        const float4 o4 = pb.sh_o[ri], d4 = pb.sh_d[ri];
        occ = query_any<STATS, TM, TB>(s, mk(o4), mk(d4), d4.w, o4.w, c, lds_col);
        pb.occluded[ri] = occ ? 1 : 0;
    }
    if (STATS) flush_counts(c, false, cnt);
}

__global__ void __launch_bounds__(256) k_raygen(DevScene s, RenderConst rc, PathBuffers pb) {
    const int32_t p = int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= rc.n_paths) return;
    // work item = layer * (sampled pixels) + pixel; the sampled grid is the image,
    // or the image with its filter border when samples splat
    const int32_t gw = rc.splat ? rc.EW : rc.W, gh = rc.splat ? rc.EH : rc.H;
    const int64_t wh = int64_t(gw) * gh;
    const int64_t w = rc.work_base + p;
    const int64_t pixel = w % wh;
    const int64_t sample = int64_t(rc.sample_begin) + w / wh;
    const int32_t off = rc.splat ? rc.border0 : 0;
    const int32_t px = int32_t(pixel % gw) - off, py = int32_t(pixel / gw) - off;
    Rng rng = rng_seed(rc.seed, path_seq((unsigned long long)pixel, (unsigned long long)sample));
    // PerPixelRenderer::render_grid (perpixel_renderer.cpp:27-38)
    const float fu = rng.uniform(), fv = rng.uniform();
    const float pxf = float(px) + fu, pyf = float(py) + fv;
    const float au = rng.uniform(), av = rng.uniform();
    const float fx = pxf / float(rc.W), fy = pyf / float(rc.H);
    // ThinLensCamera::sample_we (thin_lens.cpp:74-94)
    const V film = mk((0.5f - fx) * rc.film_w, (fy - 0.5f) * rc.film_h, rc.focal);
    float dx, dy;
    unit_disk(au, av, &dx, &dy);
    const V lens = mk(rc.lens_r * dx, rc.lens_r * dy, 0);
    const V cx = mk(rc.cam_x), cy = mk(rc.cam_y), cz = mk(rc.cam_z);
    const V o = mk(rc.cam_pos) + cx * lens.x + cy * lens.y + cz * lens.z;
    const V l = film - lens;
    const V d = normalize(cx * l.x + cy * l.y + cz * l.z);
    // box filter support: the sample reaches its own pixel only inside radius r
    // (a splatting filter weighs it for every pixel in k_film_splat)
    const bool keep = rc.splat || (fabsf(float(px) + 0.5f - pxf) <= rc.filter_r &&
                                   fabsf(float(py) + 0.5f - pyf) <= rc.filter_r);
    uint32_t flags = keep ? PF_KEEP : 0u;
    if (s.features & SF_MEDIA)  // the transmittance stream is drawn only in media
        pb.rng_tr[p] = rng_seed(rc.seed ^ TR_STREAM_KEY, path_seq((unsigned long long)pixel, (unsigned long long)sample)).state;
    bool alive = true;
    if (1 > rc.min_depth) {  // path_tracing.cpp:30-35 at depth 1 (a survivor's throughput is 1 / cont_prob)
        if (rng.uniform() > rc.cont_prob) alive = false;
    }
    const bool launch = alive && 1 <= rc.max_depth;
    if (!launch) {  // a launched path gets its G-buffer and radiance from k_shade at depth 1
        pb.gb_albedo[p] = make_float4(0, 0, 0, 1);
        pb.gb_normal[p] = make_float4(0, 0, 0, 0);
        pb.radiance[p] = make_float4(0, 0, 0, 0);
    }
    if (launch) {
        flags |= PF_CONT | PF_FRESH;
        pb.ray_o[2 * p] = f4(o, 0.0f);
        pb.ray_d[2 * p] = f4(d, INF_F);
    }
    if (rc.min_depth >= 1 && rc.max_depth >= 1) {
        // no roulette at depth 1 and a first segment to trace: every camera path is
        // queued, at its own index, so no counter is contended; path 0 writes
        // the queue sizes
        pb.q_out[p] = uint32_t(p);
        pb.q_closest[p] = uint32_t(2 * p);
        if (p == 0) {
            pb.cnt_out[2] = uint32_t(rc.n_paths);
            pb.cnt_out[0] = uint32_t(rc.n_paths);
        }
    } else {
        const uint32_t qi = wave_append(&pb.cnt_out[2], launch);
        const uint32_t ci = wave_append(&pb.cnt_out[0], launch);
        if (launch) {
            pb.q_out[qi] = uint32_t(p);
            pb.q_closest[ci] = uint32_t(2 * p);
        }
    }
    // the first path of every block of 2^shift clears the block's record slot
    if (rc.guide_record && (p & ((1u << rc.guide_rec_shift) - 1u)) == 0)
        pb.g_rec_state[p >> rc.guide_rec_shift] = make_int2(0, -1);
    pb.flags[p] = flags;
    pb.rng[p] = rng.state;
}

// ---------------------------------------------------------------- participating media
// Texture3D lookup: image3d = transform, wrap per axis, nearest or trilinear over
// texel centres, ^inv_gamma (core/texture3d.h:208-233); constant3d returns its
// texel unchanged (constant3d.cpp:48-58).
__device__ __forceinline__ float wrap3(int mode, float x) {
    if (mode == AGX_WRAP_REPEAT) return smin(1.0f, smax(0.0f, x - floorf(x)));
    if (mode == AGX_WRAP_MIRROR) {
        const int ip = int(x >= 0 ? x : floorf(x));
        const float v = (ip & 1) ? 1 - x + float(ip) : x - float(ip);
        return smin(1.0f, smax(0.0f, v));
    }
    return smin(1.0f, smax(0.0f, x));
}
// ---------------------------------------------------------------- Texture2D and the ibl environment
// Texture2D::sample_spectrum (core/texture2d.h:104-125): uv transform, wrap per
// axis, the texture (constant2d.cpp; checker_board.cpp:14-24; hdr.cpp:16-36 with
// nearest or bilinear filtering over texel centres), then ^inv_gamma
__device__ __forceinline__ V texel2(const DevScene &s, int x, int y) {
    const float *g = s.env_texels + (size_t(y) * size_t(s.env_tex.w) + size_t(x)) * 3;
    return mk(g[0], g[1], g[2]);
}
__device__ V tex2d_env(const DevScene &s, float u0, float v0) {
    const DevTex2D &t = s.env_tex;
    const float qu = t.xf[0].x * u0 + t.xf[0].y * v0 + t.xf[0].z;
    const float qv = t.xf[1].x * u0 + t.xf[1].y * v0 + t.xf[1].z;
    const float u = wrap3(t.wrap[0], qu), v = wrap3(t.wrap[1], qv);
    V r;
    if (t.kind == AGX_TEX2D_CONSTANT) {
        r = mk(t.value);
    } else if (t.kind == AGX_TEX2D_CHECKER) {
        const float gs = 1.0f / float(t.grid_count);
        const int ui = int(smin(1.0f, smax(0.0f, u)) / gs), vi = int(smin(1.0f, smax(0.0f, v)) / gs);
        r = ((ui + vi) & 1) ? mk(t.value) : mk(t.value2);
    } else if (!t.linear) {
        const int x = min(t.w - 1, int(u * float(t.w))), y = min(t.h - 1, int(v * float(t.h)));
        r = texel2(s, x, y);
    } else {
        const float fu = u * float(t.w) - 0.5f, fv = v * float(t.h) - 0.5f;
        const float bu = floorf(fu), bv = floorf(fv);
        const float du = fu - bu, dv = fv - bv;
        const int x0 = min(t.w - 1, max(0, int(bu))), x1 = min(t.w - 1, max(0, int(bu) + 1));
        const int y0 = min(t.h - 1, max(0, int(bv))), y1 = min(t.h - 1, max(0, int(bv) + 1));
        auto lerp = [](V a, V b, float k) { return a + (b - a) * k; };
        r = lerp(lerp(texel2(s, x0, y0), texel2(s, x1, y0), du), lerp(texel2(s, x0, y1), texel2(s, x1, y1), du), dv);
    }
    if (t.inv_gamma != 1.0f) r = mk(powf(r.x, t.inv_gamma), powf(r.y, t.inv_gamma), powf(r.z, t.inv_gamma));
    return r;
}
// spherical angles of a unit direction: phi = atan2(y, x) in [0, 2 pi), theta = acos(z)
__device__ __forceinline__ float dir_phi(V d) {
    const float p = atan2f(d.y, d.x);
    return p < 0 ? p + 2 * PI_F : p;
}
// radiance arriving from direction d (native_sky.cpp:13-20; IBL::radiance, ibl.cpp:160-170);
// kernels compiled without SF_IBL serve scenes without an ibl
template <uint32_t F = SF_ALL>
__device__ V env_radiance(const DevScene &s, V d) {
    if (s.env_kind == AGX_ENV_NATIVE_SKY) return sky(s, d);
    if (!(F & SF_IBL) || s.env_kind != AGX_ENV_IBL) return mk(0, 0, 0);
    const V dir = normalize(d);
    const float theta = acosf(smin(1.0f, smax(-1.0f, dir.z)));
    const float u = smin(1.0f, smax(0.0f, dir_phi(dir) / (2 * PI_F)));
    const float v = smin(1.0f, smax(0.0f, theta / PI_F));
    return tex2d_env(s, u, v);
}
// solid-angle pdf of one patch (EnvironmentLightSampler, env_sampler.h:125-183)
__device__ __forceinline__ float env_patch_pdf(const DevScene &s, int px, int py, float *cvmin, float *cvmax,
                                               float *u0, float *u1) {
    *u0 = float(px) / float(s.env_pw);
    *u1 = float(px + 1) / float(s.env_pw);
    const float v0 = float(py) / float(s.env_ph), v1 = float(py + 1) / float(s.env_ph);
    const float c1 = cosf(PI_F * v1), c0 = cosf(PI_F * v0);
    *cvmin = smin(c1, c0);
    *cvmax = smax(c1, c0);
    const float in_patch = 1.0f / (2 * PI_F * ((*u1 - *u0) * (*cvmax - *cvmin)));
    return s.env_probs[py * s.env_pw + px] * in_patch;
}
template <uint32_t F = SF_ALL>
__device__ float env_pdf(const DevScene &s, V d) {
    if (!(F & SF_IBL) || s.env_kind != AGX_ENV_IBL) return SPHERE_PDF;
    const V dir = normalize(d);
    const float theta = acosf(smin(1.0f, smax(-1.0f, dir.z)));
    const float u = dir_phi(dir) / (2 * PI_F), v = theta / PI_F;
    const int px = min(s.env_pw - 1, max(0, int(floorf(u * float(s.env_pw)))));
    const int py = min(s.env_ph - 1, max(0, int(floorf(v * float(s.env_ph)))));
    float cvmin, cvmax, u0, u1;
    return env_patch_pdf(s, px, py, &cvmin, &cvmax, &u0, &u1);
}
// direction towards the environment and its pdf (NativeSky: uniform sphere;
// IBL: alias-sampled patch, then uniform in cos theta and phi within it)
template <uint32_t F = SF_ALL>
__device__ V env_sample(const DevScene &s, float su, float sv, float sw, float *pdf) {
    if (!(F & SF_IBL) || s.env_kind != AGX_ENV_IBL) {
        *pdf = SPHERE_PDF;
        return uniform_sphere(su, sv);
    }
    const int32_t n = s.env_pw * s.env_ph;
    const float nu = su * float(n);
    int32_t idx = int32_t(nu);
    if (idx >= n) idx = n - 1;
    const float frac = nu - float(idx);
    const int32_t pick = frac < s.env_alias_p[idx] ? idx : s.env_alias_i[idx];
    const int px = pick % s.env_pw, py = pick / s.env_pw;
    float cvmin, cvmax, u0, u1;
    *pdf = env_patch_pdf(s, px, py, &cvmin, &cvmax, &u0, &u1);
    const float ct = cvmin + sw * (cvmax - cvmin);
    const float st = sqrtf(smax(0.0f, 1 - ct * ct));
    const float phi = 2 * PI_F * (u0 * (1 - sv) + u1 * sv);
    return mk(st * cosf(phi), st * sinf(phi), ct);
}

__device__ __forceinline__ V texel3(const DevScene &s, const DevTex3D &t, int x, int y, int z) {
    const size_t i = t.offset + ((size_t(z) * size_t(t.h) + size_t(y)) * size_t(t.w) + size_t(x)) * size_t(t.ch);
    if (t.f16) {  // FP16 texels widen exactly to the FP32 values they were built from
        const _Float16 *g = reinterpret_cast<const _Float16 *>(s.grids16) + i;
        if (t.ch == 1) {
            const float v = float(g[0]);
            return mk(v, v, v);
        }
        return mk(float(g[0]), float(g[1]), float(g[2]));
    }
    const float *g = s.grids + i;
    if (t.ch == 1) return mk(g[0], g[0], g[0]);
    return mk(g[0], g[1], g[2]);
}
__device__ V tex3d(const DevScene &s, const DevTex3D &t, V uvw) {
    if (t.kind == AGX_TEX3D_CONSTANT) return mk(t.value);
    V r;
    {
        const V q = xf_point(t.xf, uvw);
        const float u = wrap3(t.wrap[0], q.x), v = wrap3(t.wrap[1], q.y), w = wrap3(t.wrap[2], q.z);
        if (!t.linear) {
            const int x = min(t.w - 1, int(u * float(t.w))), y = min(t.h - 1, int(v * float(t.h)));
            const int z = min(t.d - 1, int(w * float(t.d)));
            r = texel3(s, t, x, y, z);
        } else {
            const float fu = u * float(t.w) - 0.5f, fv = v * float(t.h) - 0.5f, fw = w * float(t.d) - 0.5f;
            const float bu = floorf(fu), bv = floorf(fv), bw = floorf(fw);
            const float du = fu - bu, dv = fv - bv, dw = fw - bw;
            const int x0 = min(t.w - 1, max(0, int(bu))), x1 = min(t.w - 1, max(0, int(bu) + 1));
            const int y0 = min(t.h - 1, max(0, int(bv))), y1 = min(t.h - 1, max(0, int(bv) + 1));
            const int z0 = min(t.d - 1, max(0, int(bw))), z1 = min(t.d - 1, max(0, int(bw) + 1));
            auto lerp = [](V a, V b, float k) { return a + (b - a) * k; };
            const V c00 = lerp(texel3(s, t, x0, y0, z0), texel3(s, t, x1, y0, z0), du);
            const V c10 = lerp(texel3(s, t, x0, y1, z0), texel3(s, t, x1, y1, z0), du);
            const V c01 = lerp(texel3(s, t, x0, y0, z1), texel3(s, t, x1, y0, z1), du);
            const V c11 = lerp(texel3(s, t, x0, y1, z1), texel3(s, t, x1, y1, z1), du);
            r = lerp(lerp(c00, c10, dv), lerp(c01, c11, dv), dw);
        }
    }
    if (t.inv_gamma != 1.0f) r = mk(powf(r.x, t.inv_gamma), powf(r.y, t.inv_gamma), powf(r.z, t.inv_gamma));
    return r;
}

// Henyey-Greenstein phase function (utility/phase_function.h:8-90)
__device__ __forceinline__ float hg_phase(float g, float u) {
    const float g2 = g * g;
    const float dem = 1 + g2 - 2 * g * u;
    return (1 - g2) / (4 * PI_F * dem * sqrtf(dem));
}
__device__ __forceinline__ float hg_eval(float g, V wi, V wo) { return hg_phase(g, -cosv(wi, wo)); }
__device__ void hg_sample(float g, V wo, float su, float sv, V *dir, float *f) {
    const float sn = 2 * su - 1;
    float u;
    if (fabsf(g) < 0.001f) u = sn;
    else {
        const float g2 = g * g;
        const float q = (1 - g2) / (1 + g * sn);
        u = (1 + g2 - q * q) / (2 * g);
    }
    const float ct = -u;
    const float st = sqrtf(smax(0.0f, 1 - ct * ct));
    const float phi = 2 * PI_F * sv;
    *dir = to_global(frame_from_z(wo), mk(st * sinf(phi), st * cosf(phi), ct));
    *f = hg_phase(g, u);
}

__device__ __forceinline__ V vlerp(V a, V b, float t) { return a + (b - a) * t; }

// Bound on delta/ratio-tracking steps per segment (expected count = length x max
// density; a segment needing more is outside any scene this path renders).  The
// exit test is !(t < t_max) so that a NaN segment length ends the loop at once.
constexpr int MAX_TRACKING_STEPS = 1 << 20;

// Medium::tr (homogeneous.cpp:48-57, heterogeneous.cpp:52-72); r = the stream the caller designates
__device__ V medium_tr(const DevScene &s, int32_t mi, V a, V b, Rng &r) {
    if (mi < 0) return mk(1, 1, 1);
    const DevMedium &m = s.media[mi];
    if (m.kind == AGX_MEDIUM_HOMOGENEOUS) {
        const float d = length(a - b);
        return mk(expf(-m.st.x * d), expf(-m.st.y * d), expf(-m.st.z * d));
    }
    if (m.kind == AGX_MEDIUM_HETEROGENEOUS) {
        float result = 1, t = 0;
        const float t_max = length(a - b);
        for (int step = 0; step < MAX_TRACKING_STEPS; ++step) {
            t += -logf(1 - r.uniform()) * m.misc.z;
            if (!(t < t_max)) break;
            const V unit = xf_point(m.w2l, vlerp(a, b, t / t_max));
            result *= 1 - tex3d(s, m.density, unit).x / m.misc.y;
        }
        return mk(result, result, result);
    }
    return mk(1, 1, 1);
}

struct MScat {
    bool happened;
    V pos, wr, thr;
    float g;
};

// Medium::sample_scattering (homogeneous.cpp:62-110, heterogeneous.cpp:81-123)
__device__ MScat medium_sample(const DevScene &s, const DevMedium &m, V a, V b, Rng &r, Rng &rtr, bool indirect,
                               float eps) {
    MScat o;
    o.happened = false;
    o.thr = mk(1, 1, 1);
    o.g = 0;
    if (m.kind == AGX_MEDIUM_HOMOGENEOUS) {
        const float sam = r.uniform();
        if (black(mk(m.ss))) {
            const float d = length(a - b);
            o.thr = mk(expf(-m.st.x * d), expf(-m.st.y * d), expf(-m.st.z * d));
            return o;
        }
        const float t_max = length(b - a);
        int ch = int(sam * 3.0f);  // extract_uniform_int(u, 0, 3)
        if (ch > 2) ch = 2;
        const float nsam = sam * 3.0f - float(ch);
        const float stc = ch == 0 ? m.st.x : (ch == 1 ? m.st.y : m.st.z);
        const float stv = -logf(nsam) / stc;
        const bool in_medium = stv < t_max;
        const float tt = smin(stv, t_max);
        const V trv = mk(expf(-m.st.x * tt), expf(-m.st.y * tt), expf(-m.st.z * tt));
        const V dens = in_medium ? mk(m.st) * trv : trv;
        float pdf = 0;
        pdf += dens.x;
        pdf += dens.y;
        pdf += dens.z;
        pdf /= 3;
        pdf = smax(pdf, eps);
        V thr = trv / pdf;
        if (in_medium) thr = thr * mk(m.ss);
        o.thr = thr;
        if (in_medium) {
            o.happened = true;
            o.pos = vlerp(a, b, stv / t_max);
            o.wr = (a - b) / t_max;
            o.g = m.misc.x;
        }
        return o;
    }
    if (m.kind == AGX_MEDIUM_HETEROGENEOUS) {
        const float t_max = length(a - b);
        float t = 0;
        for (int step = 0; step < MAX_TRACKING_STEPS; ++step) {
            t += -logf(1 - r.uniform()) * m.misc.z;
            if (!(t < t_max)) break;
            const V pos = vlerp(a, b, t / t_max);
            const V unit = xf_point(m.w2l, pos);
            const float dn = tex3d(s, m.density, unit).x;
            if (r.uniform() < dn / m.misc.y) {
                o.happened = true;
                o.pos = pos;
                o.wr = (a - b) / t_max;
                o.thr = (m.white_for_indirect && indirect) ? mk(1, 1, 1) : tex3d(s, m.albedo, unit);
                o.g = tex3d(s, m.g3, unit).x;
                return o;
            }
        }
        return o;
    }
    return o;
}

__device__ __forceinline__ int32_t medium_side(const DevEntity &e, V d, V n) { return dot(d, n) >= 0 ? e.med_out : e.med_in; }

// resolve the MIS-BSDF ray of the previous vertex (direct_illum.cpp:130-192, 194-253;
// a medium vertex stores |cos| = 1 and pdf = f, so one formula serves both)
template <uint32_t F>
__device__ V resolve_mis(const DevScene &s, const PathBuffers &pb, uint32_t p, uint32_t flags, float eps, Rng &rtr) {
    const float4 ma = pb.mis_a[p], mb = pb.mis_b[p];
    const V coef = mk(ma), bf = mk(mb);
    const float absdot = ma.w, pdf = mb.w;
    const bool delta = (flags & PF_MIS_DELTA) != 0;
    const float4 h4 = pb.hit[2 * p + 1];
    const int32_t ent = __float_as_int(h4.w);
    const V o = mk(pb.ray_o[2 * p + 1]), d = mk(pb.ray_d[2 * p + 1]);
    if (ent < 0) {
        if (s.env_kind == AGX_ENV_NONE) return mk(0, 0, 0);
        const V lr = env_radiance<F>(s, d);
        if (black(lr)) return mk(0, 0, 0);
        const V f = lr * bf * absdot;
        if (delta) return coef * (f / pdf);
        return coef * (f / (pdf + env_pdf<F>(s, d)));
    }
    const DevEntity &e = s.ents[ent];
    if (e.light < 0) return mk(0, 0, 0);
    HitRec h;
    h.t = h4.x; h.u = h4.y; h.v = h4.z; h.ent = ent; h.prim = pb.hit_prim[2 * p + 1];
    const Surf sp = surface_at(s, h, o, d, eps);
    const V lrad = dot(sp.geo.z, sp.wr) > 0 ? mk(e.emit) : mk(0, 0, 0);
    if (black(lrad)) return mk(0, 0, 0);
    const int32_t med = int32_t(flags >> PF_MIS_MED_SHIFT) - 1;
    const V f = ((F & SF_MEDIA) ? medium_tr(s, med, o, sp.pos, rtr) : mk(1, 1, 1)) * lrad * bf * absdot;
    if (delta) return coef * (f / pdf);
    const float lpdf = area_light_pdf(e, o, sp.pos, sp.geo.z);
    return coef * (f / (pdf + lpdf));
}

// writes the shadow ray of slot p * L + l; k_shade queues the slots at its end
__device__ __forceinline__ void put_shadow(const PathBuffers &pb, size_t slot, V o, V d, float tmin, float tmax) {
    pb.sh_o[slot] = f4(o, tmax);
    pb.sh_d[slot] = f4(d, tmin);
}

// ---------------------------------------------------------------- path guiding
// One path of every block of 2^guide_rec_shift records (the one at offset
// -guide_rec_rot in the block); its record slot is p >> guide_rec_shift.
__device__ __forceinline__ bool guide_recording(const RenderConst &rc, uint32_t p) {
    return ((p + rc.guide_rec_rot) & ((1u << rc.guide_rec_shift) - 1u)) == 0;
}
// Network inputs of a surface vertex, each in [0, 1]: position in the scene's
// bounding sphere, shading normal, and the spherical angles of the direction
// towards the previous vertex.
__device__ void guide_inputs(const DevScene &s, const Surf &sp, float *x) {
    const float inv = 0.5f / s.world.w;
    x[0] = (sp.pos.x - s.world.x) * inv + 0.5f;
    x[1] = (sp.pos.y - s.world.y) * inv + 0.5f;
    x[2] = (sp.pos.z - s.world.z) * inv + 0.5f;
    const V n = normalize(sp.user.z);
    x[3] = 0.5f * (n.x + 1);
    x[4] = 0.5f * (n.y + 1);
    x[5] = 0.5f * (n.z + 1);
    const V w = normalize(sp.wr);
    x[6] = acosf(smin(1.0f, smax(-1.0f, w.z))) / PI_F;
    x[7] = (atan2f(w.y, w.x) + PI_F) / (2 * PI_F);
}
// A training record of a guided vertex of a recording path: its inputs, the
// sampled direction, the mixture pdf, the throughput after the vertex and
// lum(f) |cos| (or the phase value at a medium point) of the sampled direction.
__device__ void guide_record(const PathBuffers &pb, const RenderConst &rc, uint32_t p, uint32_t tid, V dir, float pdf,
                             V coef, float flum) {
    int2 st = pb.g_rec_state[p >> rc.guide_rec_shift];
    if (st.x >= GUIDE_MAX_REC) return;
    float *rec = pb.g_rec + (size_t(p >> rc.guide_rec_shift) * GUIDE_MAX_REC + st.x) * GUIDE_REC_FLOATS;
    const float *x = pb.g_feat + size_t(tid) * 8;
    for (int k = 0; k < 8; ++k) rec[k] = x[k];
    rec[8] = dir.x; rec[9] = dir.y; rec[10] = dir.z;
    rec[11] = pdf;
    rec[12] = coef.x; rec[13] = coef.y; rec[14] = coef.z;
    rec[15] = rec[16] = rec[17] = 0;
    rec[18] = flum;
    st.y = st.x;
    st.x += 1;
    pb.g_rec_state[p >> rc.guide_rec_shift] = st;
}

// Surfaces whose continuation may come from the learned density.  The one-sample
// mixture needs the BSDF's pdf to be the density of its sampler; Disney's
// pdf differs from its sampler (the diffuse pick overwritten, |sin phi_o| in
// the specular term, material/disney.cpp), so Disney vertices keep plain BSDF
// sampling and the image stays Atrc's.
__device__ __forceinline__ bool guidable(int32_t kind) {
    return kind == AGX_MAT_IDEAL_DIFFUSE || kind == AGX_MAT_METAL;
}
// von Mises-Fisher lobe: density kappa / (2 pi (1 - e^-2kappa)) e^{kappa (mu.d - 1)}
__device__ __forceinline__ float vmf_pdf(V mu, float k, V d) {
    return k / (2 * PI_F * -expm1f(-2 * k)) * expf(k * (dot(mu, d) - 1));
}
// mixture density of a decoded network row at the unit direction d
__device__ float guide_pdf(const float *G, V d) {
    float q = 0;
    for (int j = 0; j < 8; ++j) {
        const float4 l = *reinterpret_cast<const float4 *>(G + 4 * j);
        q += G[32 + j] * vmf_pdf(mk(l.x, l.y, l.z), l.w, d);
    }
    return q;
}
// lobe j by inverse CDF of the weights with u0, then the stable vMF sampler
// w = 1 + log(u + (1 - u) e^-2kappa) / kappa around mu
__device__ V guide_sample(const float *G, float u0, float u1, float u2) {
    int j = 0;
    float c = G[32];
    while (j < 7 && u0 >= c) { ++j; c += G[32 + j]; }
    const float4 l = *reinterpret_cast<const float4 *>(G + 4 * j);
    const float k = l.w;
    const float w = 1.0f + logf(u1 + (1.0f - u1) * expf(-2.0f * k)) / k;
    const float sw = sqrtf(smax(0.0f, 1 - w * w));
    const float phi = 2 * PI_F * u2;
    return normalize(to_global(frame_from_z(mk(l.x, l.y, l.z)), mk(sw * cosf(phi), sw * sinf(phi), w)));
}
__device__ __forceinline__ float lum3(V c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

// the increment of path p's streams: its (sampled pixel, sample), as k_raygen
// keys it (the sampled grid includes the filter border when samples splat)
__device__ __forceinline__ unsigned long long path_inc(const RenderConst &rc, uint32_t p) {
    const int64_t wh = rc.splat ? int64_t(rc.EW) * rc.EH : int64_t(rc.W) * rc.H;
    const int64_t w = rc.work_base + int64_t(p);
    return (path_seq((unsigned long long)(w % wh), (unsigned long long)(rc.sample_begin + w / wh)) << 1u) | 1u;
}

// Network inputs of a medium scattering point: position and direction as at a
// surface, and the normal inputs at 0.5, the encoding of a zero normal, which
// no surface vertex has.
__device__ void guide_inputs_medium(const DevScene &s, V pos, V wr, float *x) {
    const float inv = 0.5f / s.world.w;
    x[0] = (pos.x - s.world.x) * inv + 0.5f;
    x[1] = (pos.y - s.world.y) * inv + 0.5f;
    x[2] = (pos.z - s.world.z) * inv + 0.5f;
    x[3] = x[4] = x[5] = 0.5f;
    const V w = normalize(wr);
    x[6] = acosf(smin(1.0f, smax(-1.0f, w.z))) / PI_F;
    x[7] = (atan2f(w.y, w.x) + PI_F) / (2 * PI_F);
}

// Guided renders in media: the medium segment of each vertex about to be
// shaded is sampled here, before inference, so that a scattering point gets a
// network evaluation of its own.  The draws are the ones k_shade would make
// first at this vertex (Medium::sample_scattering, path_tracing.cpp:64-76);
// k_shade takes the stored sample instead of drawing again.
__global__ void __launch_bounds__(256) k_medium_pass(DevScene s, RenderConst rc, PathBuffers pb) {
    const uint32_t n = pb.cnt_in[2];
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= n) return;
    const uint32_t p = pb.q_in[tid];
    uint32_t flags = pb.flags[p] & ~PF_MS;
    const float4 h4 = pb.hit[2 * p];
    const int32_t ent = __float_as_int(h4.w);
    if ((flags & PF_CONT) && ent >= 0) {
        const int4 dep = (flags & PF_FRESH) ? make_int4(1, 1, 0, 0) : pb.depth[p];
        HitRec h;
        h.t = h4.x; h.u = h4.y; h.v = h4.z; h.ent = ent; h.prim = pb.hit_prim[2 * p];
        const V ro = mk(pb.ray_o[2 * p]), rd = mk(pb.ray_d[2 * p]);
        const Surf sp = surface_at(s, h, ro, rd, rc.eps);
        const DevEntity &e = s.ents[ent];
        const int32_t mi = medium_side(e, sp.wr, sp.geo.z);
        if (mi >= 0 && dep.z < s.media[mi].max_scat) {
            Rng rng, rtr;
            rng.state = pb.rng[p];
            rtr.state = pb.rng_tr[p];
            rng.inc = rtr.inc = path_inc(rc, p);
            const MScat ms = medium_sample(s, s.media[mi], ro, sp.pos, rng, rtr, dep.z > 0, rc.eps);
            pb.g_ms[3 * size_t(p)] = f4(ms.pos, ms.g);
            pb.g_ms[3 * size_t(p) + 1] = f4(ms.thr, ms.happened ? 1.0f : 0.0f);
            pb.g_ms[3 * size_t(p) + 2] = f4(ms.wr, 0.0f);
            pb.rng[p] = rng.state;
            pb.rng_tr[p] = rtr.state;
            flags |= PF_MS;
        }
    }
    pb.flags[p] = flags;
}

__global__ void __launch_bounds__(256) k_guide_features(DevScene s, RenderConst rc, PathBuffers pb) {
    const uint32_t n = pb.cnt_in[2];
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= n) return;
    const uint32_t p = pb.q_in[tid];
    float x[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const float4 h4 = pb.hit[2 * p];
    const int32_t ent = __float_as_int(h4.w);
    const uint32_t flags = pb.flags[p];
    if ((flags & PF_MS) && pb.g_ms[3 * size_t(p) + 1].w != 0.0f) {
        guide_inputs_medium(s, mk(pb.g_ms[3 * size_t(p)]), mk(pb.g_ms[3 * size_t(p) + 2]), x);
    } else if ((flags & PF_CONT) && ent >= 0) {
        HitRec h;
        h.t = h4.x; h.u = h4.y; h.v = h4.z; h.ent = ent; h.prim = pb.hit_prim[2 * p];
        const Surf sp = surface_at(s, h, mk(pb.ray_o[2 * p]), mk(pb.ray_d[2 * p]), rc.eps);
        guide_inputs(s, sp, x);
    }
    float4 *dst = reinterpret_cast<float4 *>(pb.g_feat + size_t(tid) * 8);
    dst[0] = make_float4(x[0], x[1], x[2], x[3]);
    dst[1] = make_float4(x[4], x[5], x[6], x[7]);
}

template <uint32_t F>
__global__ void __launch_bounds__(256) k_shade(DevScene s, RenderConst rc, PathBuffers pb) {
    const uint32_t n = pb.cnt_in[2];
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= n) return;
    const uint32_t p = pb.q_in[tid];
    const float eps = rc.eps;
    uint32_t flags = pb.flags[p];
    // a fresh camera path: nothing stored yet but its flags, rays and stream
    const bool fresh = (flags & PF_FRESH) != 0;
    flags &= ~PF_FRESH;
    V L = fresh ? mk(0, 0, 0) : mk(pb.radiance[p]);
    // the transmittance stream is drawn only in media: vacuum variants neither
    // load nor store it
    Rng rtr;
    const bool media = (F & SF_MEDIA) && (s.features & SF_MEDIA);
    rtr.state = media ? pb.rng_tr[p] : 0ull;
    Rng rng;
    rng.state = pb.rng[p];
    rng.inc = rtr.inc = path_inc(rc, p);
    // the medium sample k_medium_pass took for this vertex (guided renders in media)
    const bool ms_done = (F & SF_GUIDE) && (F & SF_MEDIA) && (flags & PF_MS);
    flags &= ~PF_MS;
    // 1. direct illumination of the previous vertex (path_tracing.cpp:85-97, 134-146)
    if (flags & PF_PENDING) {
        V direct = mk(0, 0, 0);
        // two lights' records and occlusion bits in flight at a time (the
        // light count is a run-time value, so the compiler would issue one
        // light's loads per iteration); the sum keeps light order
        const size_t nb = size_t(p) * rc.L;
        int32_t l = 0;
        for (; l + 1 < rc.L; l += 2) {
            const float4 c0 = pb.nee[nb + l], c1 = pb.nee[nb + l + 1];
            const uint8_t o0 = pb.occluded[nb + l], o1 = pb.occluded[nb + l + 1];
            if (c0.w != 0.0f && !o0) direct = direct + mk(c0);
            if (c1.w != 0.0f && !o1) direct = direct + mk(c1);
        }
        if (l < rc.L) {
            const float4 c = pb.nee[nb + l];
            if (c.w != 0.0f && !pb.occluded[nb + l]) direct = direct + mk(c);
        }
        if (flags & PF_MIS) direct = direct + resolve_mis<F>(s, pb, p, flags, eps, rtr);
        L = L + direct;
        if ((F & SF_GUIDE) && rc.guide_record && guide_recording(rc, p) && !black(direct)) {
            // light reaching the previous vertex arrives at every earlier recorded
            // vertex along its sampled direction: incident radiance += direct / T
            const int2 st = pb.g_rec_state[p >> rc.guide_rec_shift];
            float *recs = pb.g_rec + size_t(p >> rc.guide_rec_shift) * GUIDE_MAX_REC * GUIDE_REC_FLOATS;
            for (int j = 0; j < st.x; ++j) {
                if (j == st.y) continue;
                float *rec = recs + j * GUIDE_REC_FLOATS;
                if (rec[12] > 0) rec[15] += direct.x / rec[12];
                if (rec[13] > 0) rec[16] += direct.y / rec[13];
                if (rec[14] > 0) rec[17] += direct.z / rec[14];
            }
        }
        flags &= ~(PF_PENDING | PF_MIS | PF_MIS_DELTA);
        flags &= (1u << PF_MIS_MED_SHIFT) - 1u;
    }
    if (!(flags & PF_CONT)) {
        pb.radiance[p] = f4(L, 0);
        pb.flags[p] = flags;
        if (media) pb.rng_tr[p] = rtr.state;
        return;
    }
    flags &= ~PF_CONT;
    // this vertex has made no training record (yet): its direct light will reach every earlier record
    if ((F & SF_GUIDE) && rc.guide_record && guide_recording(rc, p)) pb.g_rec_state[p >> rc.guide_rec_shift].y = -1;
    V coef;
    int4 dep;
    if (fresh) {  // k_raygen's depth-1 state: throughput 1, or 1 / cont_prob after the depth-1 roulette
        const float c0 = (1 > rc.min_depth) ? 1.0f / rc.cont_prob : 1.0f;
        coef = mk(c0, c0, c0);
        dep = make_int4(1, 1, 0, 0);
    } else {
        coef = mk(pb.coef[p]);
        dep = pb.depth[p];
    }
    const float4 h4 = pb.hit[2 * p];
    const int32_t hit_prim = pb.hit_prim[2 * p];  // loaded with the hit (a miss leaves it unused)
    const int32_t ent = __float_as_int(h4.w);
    const V ro = mk(pb.ray_o[2 * p]), rd = mk(pb.ray_d[2 * p]);
    if (ent < 0) {  // path_tracing.cpp:41-49 (trace_nomis, :250-256: at every depth)
        if ((dep.x == 1 || rc.nomis) && s.env_kind != AGX_ENV_NONE) L = L + coef * env_radiance<F>(s, rd);
        if (dep.x == 1) {  // the camera ray escaped: empty G-buffer (k_raygen leaves it to this branch)
            pb.gb_normal[p] = make_float4(0, 0, 0, 0);
            pb.gb_albedo[p] = make_float4(0, 0, 0, 1);
        }
        pb.radiance[p] = f4(L, 0);
        pb.flags[p] = flags;
        if (media) pb.rng_tr[p] = rtr.state;
        return;
    }
    HitRec h;
    h.t = h4.x; h.u = h4.y; h.v = h4.z; h.ent = ent; h.prim = hit_prim;
    const Surf sp = surface_at(s, h, ro, rd, eps);
    const DevEntity &e = s.ents[ent];
    const Bsdf b = make_bsdf(s.mats[e.material], sp, eps);
    if (dep.x == 1) {  // :54-60
        pb.gb_normal[p] = f4(sp.user.z, 0);
        pb.gb_albedo[p] = f4(b.albedo, (e.flags & EF_NO_DENOISE) ? 0.0f : 1.0f);
    }
    bool pending = false;
    bool cont = false;
    bool push_mis = false;
    uint32_t shadow_mask = 0;  // bit l: the shadow ray of light l is traced
    // medium between the ray origin and the hit (path_tracing.cpp:64-119)
    const int32_t mi = (F & SF_MEDIA) ? medium_side(e, sp.wr, sp.geo.z) : -1;
    bool scattered = false;
    if ((F & SF_MEDIA) && mi >= 0) {
        const DevMedium &m = s.media[mi];
        if (dep.z < m.max_scat) {
            MScat ms;
            if (ms_done) {
                const float4 a = pb.g_ms[3 * size_t(p)], b4 = pb.g_ms[3 * size_t(p) + 1], c4 = pb.g_ms[3 * size_t(p) + 2];
                ms.pos = mk(a); ms.g = a.w; ms.thr = mk(b4); ms.happened = b4.w != 0.0f; ms.wr = mk(c4);
            } else {
                ms = medium_sample(s, m, ro, sp.pos, rng, rtr, dep.z > 0, eps);
            }
            coef = coef * ms.thr;
            if (ms.happened) {
                scattered = true;
                dep.z += 1;
                // NEE at the medium point: area lights only; the environment draws nothing
                // (trace_nomis samples no light)
                for (int32_t l = 0; l < (rc.nomis ? 0 : rc.L); ++l) {
                    const size_t slot = size_t(p) * rc.L + l;
                    const int32_t li = s.lights[l];
                    pb.nee[slot] = make_float4(0, 0, 0, 0);
                    if (li < 0) continue;
                    float s5[5];
                    for (int k = 0; k < 5; ++k) s5[k] = rng.uniform();
                    const DevEntity &le = s.ents[li];
                    const LightSmp ls = area_light_sample(s, le, ms.pos, s5[0], s5[1], s5[3], eps);
                    const V rad = mk(le.emit);
                    if (black(rad) || ls.pdf == 0.0f) continue;
                    const V tl = normalize(ls.pos - ms.pos);
                    const float fp = hg_eval(ms.g, tl, ms.wr);
                    if (fp == 0.0f) continue;
                    const V f = medium_tr(s, mi, ms.pos, ls.pos, rtr) * rad * mk(fp, fp, fp);
                    pb.nee[slot] = f4(coef * (f / (ls.pdf + fp)), 1.0f);
                    pending = true;
                    // Scene::visible(light, point): from the light towards the point
                    const float dis = length(ls.pos - ms.pos);
                    put_shadow(pb, slot, ls.pos, normalize(ms.pos - ls.pos), eps, dis - eps);
                    shadow_mask |= 1u << l;
                }
                if (!rc.nomis) {  // MIS-BSDF ray with the phase function (direct_illum.cpp:194-253)
                    const float u0 = rng.uniform(), u1 = rng.uniform();
                    (void)rng.uniform();
                    V dir;
                    float f;
                    hg_sample(ms.g, ms.wr, u0, u1, &dir, &f);
                    if (f != 0.0f) {
                        dir = normalize(dir);
                        pb.ray_o[2 * p + 1] = f4(ms.pos, 0.0f);
                        pb.ray_d[2 * p + 1] = f4(dir, INF_F);
                        pb.mis_a[p] = f4(coef, 1.0f);
                        pb.mis_b[p] = make_float4(f, f, f, f);
                        flags |= PF_MIS | (uint32_t(mi + 1) << PF_MIS_MED_SHIFT);
                        pending = true;
                        push_mis = true;
                    }
                }
                if ((F & SF_GUIDE) && rc.guide && dep.x <= rc.guide_medium_depth) {
                    // guided: one-sample mixture of the learned density (probability
                    // alpha) and the phase function, weighted by the mixture density
                    const float usel = rng.uniform();
                    const float u0 = rng.uniform(), u1 = rng.uniform(), u2 = rng.uniform();
                    const float *G = pb.g_out + size_t(tid) * GUIDE_ROW;
                    const float a = rc.guide_alpha;
                    V dir;
                    float fp;
                    if (usel < a) {
                        dir = guide_sample(G, u0, u1, u2);
                        fp = hg_eval(ms.g, dir, ms.wr);
                    } else {
                        hg_sample(ms.g, ms.wr, u0, u1, &dir, &fp);
                        dir = normalize(dir);
                    }
                    const float pdf = a * guide_pdf(G, dir) + (1 - a) * fp;  // the phase function's pdf is its value
                    if (fp != 0.0f && !(pdf < eps) && isfinite(pdf)) {
                        pb.ray_o[2 * p] = f4(ms.pos, 0.0f);
                        pb.ray_d[2 * p] = f4(dir, INF_F);
                        coef = coef * (fp / pdf);
                        dep.x += 1;
                        if (dep.x <= rc.max_depth) {
                            cont = true;
                            if (dep.x > rc.min_depth) {
                                if (rng.uniform() > rc.cont_prob) cont = false;
                                else coef = coef / rc.cont_prob;
                            }
                        }
                        if (cont && rc.guide_record && guide_recording(rc, p))
                            guide_record(pb, rc, p, tid, dir, pdf, coef, fp);
                    }
                } else {  // phase-function continuation (path_tracing.cpp:99-108; trace_nomis :285-296
                   // samples around the surface hit's wr, tests only f and keeps the direction as sampled)
                    const float u0 = rng.uniform(), u1 = rng.uniform();
                    (void)rng.uniform();
                    V dir;
                    float f;
                    hg_sample(ms.g, rc.nomis ? sp.wr : ms.wr, u0, u1, &dir, &f);
                    if (f != 0.0f && (rc.nomis || !(f < eps))) {
                        pb.ray_o[2 * p] = f4(ms.pos, 0.0f);
                        pb.ray_d[2 * p] = f4(rc.nomis ? dir : normalize(dir), INF_F);
                        coef = coef * (mk(f, f, f) / f);
                        dep.x += 1;
                        if (dep.x <= rc.max_depth) {
                            cont = true;
                            if (dep.x > rc.min_depth) {
                                if (rng.uniform() > rc.cont_prob) cont = false;
                                else coef = coef / rc.cont_prob;
                            }
                        }
                    }
                }
            }
        } else {
            // medium->ab (absorption only) when the scattering count is exhausted
            if (m.kind == AGX_MEDIUM_HOMOGENEOUS) {
                const float d = length(ro - sp.pos);
                coef = coef * mk(expf(-m.sa.x * d), expf(-m.sa.y * d), expf(-m.sa.z * d));
            } else {
                coef = coef * medium_tr(s, mi, ro, sp.pos, rng);
            }
        }
    }
    if (!scattered) {
        dep.z = 0;
        if ((dep.x == 1 || rc.nomis) && e.light >= 0)  // :123-130 (trace_nomis, :313-318: at every depth)
            if (dot(sp.geo.z, sp.wr) > 0) L = L + coef * mk(e.emit);
        // NEE over every light (mis_sample_light, direct_illum.cpp:11-42, 72-99); none in trace_nomis
        for (int32_t l = 0; l < (rc.nomis ? 0 : rc.L); ++l) {
            float s5[5];
            for (int k = 0; k < 5; ++k) s5[k] = rng.uniform();
            const size_t slot = size_t(p) * rc.L + l;
            const int32_t li = s.lights[l];
            V contrib = mk(0, 0, 0);
            bool valid = false;
            V so, sd;
            float tmax = 0;
            if (li >= 0) {
                const DevEntity &le = s.ents[li];
                const LightSmp ls = area_light_sample(s, le, sp.pos, s5[0], s5[1], s5[3], eps);
                const V rad = mk(le.emit);
                if (!black(rad) && ls.pdf != 0.0f) {
                    const float len = length(ls.pos - sp.pos) - eps;
                    if (len > eps) {
                        const V tl = normalize(ls.pos - sp.pos);
                        const V f = bsdf_eval<F>(b, tl, sp.wr);
                        if (!black(f)) {
                            const V trv = (F & SF_MEDIA) ? medium_tr(s, medium_side(e, tl, sp.geo.z), ls.pos, sp.pos, rtr)
                                                         : mk(1, 1, 1);
                            const V ff = trv * rad * f * fabsf(cosv(tl, sp.geo.z));
                            const float bp = bsdf_pdf<F>(b, tl, sp.wr);
                            contrib = coef * (ff / (ls.pdf + bp));
                            valid = true;
                            so = sp.pos; sd = tl; tmax = len;
                        }
                    }
                }
            } else {
                float epdf;
                const V dir = env_sample<F>(s, s5[0], s5[1], s5[2], &epdf);
                const V lpos = env_emit_pos(s, sp.pos, dir);
                const V rad = env_radiance<F>(s, dir);
                if (!black(rad)) {
                    const V tl = normalize(lpos - sp.pos);
                    const V f = bsdf_eval<F>(b, tl, sp.wr);
                    if (!black(f)) {
                        const V ff = rad * f * fabsf(cosv(tl, sp.geo.z));
                        const float bp = bsdf_pdf<F>(b, tl, sp.wr);
                        contrib = coef * (ff / (epdf + bp));
                        valid = true;
                        // Scene::visible(A, B): [eps, |A-B| - eps] along normalize(B - A)
                        so = sp.pos; sd = normalize(lpos - sp.pos); tmax = length(sp.pos - lpos) - eps;
                    }
                }
            }
            pb.nee[slot] = f4(contrib, valid ? 1.0f : 0.0f);
            if (valid) {
                pending = true;
                put_shadow(pb, slot, so, sd, eps, tmax);
                shadow_mask |= 1u << l;
            }
        }
        // MIS-BSDF ray (mis_sample_bsdf, direct_illum.cpp:130-147); none in trace_nomis
        if (!rc.nomis) {
            const float u0 = rng.uniform(), u1 = rng.uniform(), u2 = rng.uniform();
            const BSmp bs = bsdf_sample<F>(b, sp.wr, u0, u1, u2, eps);
            if (!black(bs.f)) {
                const V dir = normalize(bs.dir);
                pb.ray_o[2 * p + 1] = f4(eps_offset(sp, dir, eps), 0.0f);
                pb.ray_d[2 * p + 1] = f4(dir, INF_F);
                pb.mis_a[p] = f4(coef, fabsf(dot(sp.geo.z, dir)));
                pb.mis_b[p] = f4(bs.f, bs.pdf);
                const int32_t mm = (F & SF_MEDIA) ? medium_side(e, bs.dir, sp.geo.z) : -1;
                flags |= PF_MIS | (bs.delta ? PF_MIS_DELTA : 0u) | (uint32_t(mm + 1) << PF_MIS_MED_SHIFT);
                pending = true;
                push_mis = true;
            }
        }
        // continuation (path_tracing.cpp:150-169) and Russian roulette of the next depth
        if ((F & SF_GUIDE) && rc.guide && guidable(b.kind)) {
            // one-sample mixture of the learned density (probability alpha) and the BSDF
            const float usel = rng.uniform();
            const float u0 = rng.uniform(), u1 = rng.uniform(), u2 = rng.uniform();
            const float *G = pb.g_out + size_t(tid) * GUIDE_ROW;
            const float a = rc.guide_alpha;
            V dir, f;
            float pbsdf;
            if (usel < a) {
                dir = guide_sample(G, u0, u1, u2);
                f = bsdf_eval<F>(b, dir, sp.wr);
                pbsdf = bsdf_pdf<F>(b, dir, sp.wr);
            } else {
                const BSmp bs = bsdf_sample<F>(b, sp.wr, u0, u1, u2, eps);
                dir = normalize(bs.dir);
                f = bs.f;
                pbsdf = bs.pdf;
            }
            const float pdf = a * guide_pdf(G, dir) + (1 - a) * pbsdf;
            if (!black(f) && !(pdf < eps) && isfinite(pdf)) {
                const float abscos = fabsf(cosv(sp.geo.z, dir));
                coef = coef * (f * abscos / pdf);
                pb.ray_o[2 * p] = f4(eps_offset(sp, dir, eps), 0.0f);
                pb.ray_d[2 * p] = f4(dir, INF_F);
                dep.x += 1;
                if (dep.x <= rc.max_depth) {
                    cont = true;
                    if (dep.x > rc.min_depth) {
                        if (rng.uniform() > rc.cont_prob) cont = false;
                        else coef = coef / rc.cont_prob;
                    }
                }
                if (cont && rc.guide_record && guide_recording(rc, p))
                    guide_record(pb, rc, p, tid, dir, pdf, coef, lum3(f) * abscos);
            }
        } else {
            const float u0 = rng.uniform(), u1 = rng.uniform(), u2 = rng.uniform();
            const BSmp bs = bsdf_sample<F>(b, sp.wr, u0, u1, u2, eps);
            // trace_nomis (:320-327) rejects only a black f and keeps the sampled direction
            if (!black(bs.f) && (rc.nomis || !(bs.pdf < eps))) {
                const float abscos = fabsf(cosv(sp.geo.z, bs.dir));
                coef = coef * (bs.f * abscos / bs.pdf);
                pb.ray_o[2 * p] = f4(eps_offset(sp, bs.dir, eps), 0.0f);
                pb.ray_d[2 * p] = f4(rc.nomis ? bs.dir : normalize(bs.dir), INF_F);
                if (bs.delta && dep.x >= 2 && dep.y <= rc.specular_depth) { dep.x -= 1; dep.y += 1; }
                dep.x += 1;
                if (dep.x <= rc.max_depth) {
                    cont = true;
                    if (dep.x > rc.min_depth) {
                        if (rng.uniform() > rc.cont_prob) cont = false;
                        else coef = coef / rc.cont_prob;
                    }
                }
            }
        }
    }
    if (pending) flags |= PF_PENDING;
    pb.radiance[p] = f4(L, 0);
    if (cont) flags |= PF_CONT;
    // the rays of this vertex: 2p (PF_CONT), 2p + 1 (PF_MIS) and the shadow rays of
    // the lights in shadow_mask; k_queue_* build the queues from these after the launch
    (void)push_mis;
    dep.w = int32_t(shadow_mask);
    pb.coef[p] = f4(coef, 0);
    pb.depth[p] = dep;
    pb.flags[p] = flags;
    pb.rng[p] = rng.state;
    if (media) pb.rng_tr[p] = rtr.state;
}

// ---------------------------------------------------------------- queue compaction
// The queues of the next bounce (closest-hit rays, shadow rays, paths to shade)
// are built from the flags k_shade leaves on every path it shaded, in the order
// of the shade queue: per-block counts, one block's exclusive scan of them, and a
// scatter.  No atomics (a device-scope atomic on one counter serialises every
// wave of a launch behind it), and the queues come out in path order, the same
// in every run.
constexpr uint32_t QC_THREADS = 256, QC_ITEMS = 8, QC_SPAN = QC_THREADS * QC_ITEMS;
static_assert(MAX_LIGHTS < 32, "the shadow mask of a path is one 32-bit word");

struct QPred { uint32_t cont, mis, mask, again; };
__device__ __forceinline__ QPred queue_pred(const PathBuffers &pb, uint32_t p, uint32_t L) {
    const uint32_t f = pb.flags[p];
    QPred q;
    q.cont = (f & PF_CONT) ? 1u : 0u;
    q.mis = ((f & PF_MIS) && (f & PF_PENDING)) ? 1u : 0u;
    q.again = (f & (PF_CONT | PF_PENDING)) ? 1u : 0u;
    q.mask = (f & PF_PENDING) ? uint32_t(pb.depth[p].w) & ((1u << L) - 1u) : 0u;  // L <= MAX_LIGHTS
    return q;
}
// exclusive prefix sum over the wave's lanes; *total = the wave's sum
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, uint32_t *total) {
    const int lane = int(__lane_id());
    uint32_t x = v;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    *total = __shfl(x, 63);
    return x - v;
}

// blk[b] = (closest rays, shadow rays, paths) of shade-queue entries [b * QC_SPAN, (b + 1) * QC_SPAN)
__global__ void __launch_bounds__(QC_THREADS) k_queue_count(PathBuffers pb, uint32_t L, uint4 *blk) {
    const uint32_t n = pb.cnt_in[2];
    const uint32_t base = blockIdx.x * QC_SPAN;
    if (base >= n) return;
    uint32_t c0 = 0, c1 = 0, c2 = 0;
    for (uint32_t k = 0; k < QC_ITEMS; ++k) {
        const uint32_t tid = base + k * QC_THREADS + threadIdx.x;
        if (tid < n) {
            const QPred q = queue_pred(pb, pb.q_in[tid], L);
            c0 += q.cont + q.mis;
            c1 += uint32_t(__popc(q.mask));
            c2 += q.again;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        c0 += __shfl_xor(c0, off);
        c1 += __shfl_xor(c1, off);
        c2 += __shfl_xor(c2, off);
    }
    __shared__ uint32_t red[3][QC_THREADS / 64];
    const uint32_t w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { red[0][w] = c0; red[1][w] = c1; red[2][w] = c2; }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint4 t = make_uint4(0, 0, 0, 0);
        for (uint32_t j = 0; j < QC_THREADS / 64; ++j) { t.x += red[0][j]; t.y += red[1][j]; t.z += red[2][j]; }
        blk[blockIdx.x] = t;
    }
}

// one block: blk[b] <- exclusive prefix of the block counts; cnt_out[0..2] <- totals
__global__ void __launch_bounds__(1024) k_queue_scan(PathBuffers pb, uint4 *blk) {
    const uint32_t n = pb.cnt_in[2];
    const uint32_t nb = (n + QC_SPAN - 1) / QC_SPAN;
    __shared__ uint32_t wsum[3][16];
    __shared__ uint32_t carry[3];
    if (threadIdx.x < 3) carry[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t w = threadIdx.x >> 6;
    for (uint32_t b0 = 0; b0 < nb; b0 += 1024) {
        const uint32_t i = b0 + threadIdx.x;
        const uint4 v = i < nb ? blk[i] : make_uint4(0, 0, 0, 0);
        uint32_t s0, s1, s2;
        const uint32_t e0 = wave_excl_scan(v.x, &s0), e1 = wave_excl_scan(v.y, &s1), e2 = wave_excl_scan(v.z, &s2);
        if ((threadIdx.x & 63) == 0) { wsum[0][w] = s0; wsum[1][w] = s1; wsum[2][w] = s2; }
        __syncthreads();
        uint32_t o0 = carry[0], o1 = carry[1], o2 = carry[2];
        for (uint32_t j = 0; j < w; ++j) { o0 += wsum[0][j]; o1 += wsum[1][j]; o2 += wsum[2][j]; }
        if (i < nb) blk[i] = make_uint4(o0 + e0, o1 + e1, o2 + e2, 0);
        __syncthreads();  // every thread has read carry and wsum
        if (threadIdx.x < 3)
            for (uint32_t j = 0; j < 16; ++j) carry[threadIdx.x] += wsum[threadIdx.x][j];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        pb.cnt_out[0] = carry[0];
        pb.cnt_out[1] = carry[1];
        pb.cnt_out[2] = carry[2];
    }
}

// writes the entries of every shaded path at its block's offset + the rank of
// its entries in the block (queue order; a path's 2p before its 2p + 1)
__global__ void __launch_bounds__(QC_THREADS) k_queue_scatter(PathBuffers pb, uint32_t L, const uint4 *blk) {
    const uint32_t n = pb.cnt_in[2];
    const uint32_t base = blockIdx.x * QC_SPAN;
    if (base >= n) return;
    __shared__ uint32_t wsum[2][3][QC_THREADS / 64];
    const uint4 o = blk[blockIdx.x];
    uint32_t run0 = o.x, run1 = o.y, run2 = o.z;
    const uint32_t w = threadIdx.x >> 6;
    for (uint32_t k = 0; k < QC_ITEMS; ++k) {
        const uint32_t tid = base + k * QC_THREADS + threadIdx.x;
        QPred q{0u, 0u, 0u, 0u};
        uint32_t p = 0;
        if (tid < n) {
            p = pb.q_in[tid];
            q = queue_pred(pb, p, L);
        }
        uint32_t s0, s1, s2;
        const uint32_t e0 = wave_excl_scan(q.cont + q.mis, &s0);
        const uint32_t e1 = wave_excl_scan(uint32_t(__popc(q.mask)), &s1);
        const uint32_t e2 = wave_excl_scan(q.again, &s2);
        // double-buffered by k: the next write to this half follows another barrier
        uint32_t (*ws)[QC_THREADS / 64] = wsum[k & 1];
        if ((threadIdx.x & 63) == 0) { ws[0][w] = s0; ws[1][w] = s1; ws[2][w] = s2; }
        __syncthreads();
        uint32_t b0 = run0, b1 = run1, b2 = run2;
        for (uint32_t j = 0; j < QC_THREADS / 64; ++j) {
            const uint32_t a0 = ws[0][j], a1 = ws[1][j], a2 = ws[2][j];
            if (j < w) { b0 += a0; b1 += a1; b2 += a2; }
            run0 += a0; run1 += a1; run2 += a2;
        }
        if (q.cont) pb.q_closest[b0 + e0] = 2 * p;
        if (q.mis) pb.q_closest[b0 + e0 + q.cont] = 2 * p + 1;
        uint32_t si = b1 + e1;
        for (uint32_t l = 0; l < L; ++l)
            if ((q.mask >> l) & 1u) pb.q_shadow[si++] = p * L + l;
        if (q.again) pb.q_out[b2 + e2] = p;
    }
}

// ---------------------------------------------------------------- training samples
// Training samples from the records of the recording paths, written to the
// sample ring from slot ring[0] on, in record order (count, scan, scatter: the
// same samples reach the same slots in every run): weight = luminance of the
// incident radiance estimate x luminance of f |cos| / density the direction was
// drawn with (the product the sampling density should follow).  Records without
// light are skipped: their weight is zero.
__device__ __forceinline__ bool gather_ok(const RenderConst &rc, const PathBuffers &pb, uint32_t gid,
                                          const float **rec_out, float *w_out) {
    const uint32_t r = gid / GUIDE_MAX_REC, j = gid % GUIDE_MAX_REC;
    const uint32_t last_slot = uint32_t(rc.n_paths - 1) >> rc.guide_rec_shift;
    if (r > last_slot || int32_t(j) >= pb.g_rec_state[r].x) return false;
    const float *rec = pb.g_rec + (size_t(r) * GUIDE_MAX_REC + j) * GUIDE_REC_FLOATS;
    const float pdf = rec[11];
    const float w = lum3(mk(rec[15], rec[16], rec[17])) * rec[18] / pdf;
    *rec_out = rec;
    *w_out = w;
    return pdf > 0 && w > 0 && isfinite(w);
}

__global__ void __launch_bounds__(256) k_guide_gather_count(RenderConst rc, PathBuffers pb, uint32_t n, uint32_t *blk) {
    const uint32_t gid = blockIdx.x * 256u + threadIdx.x;
    const float *rec = nullptr;
    float w = 0;
    const bool ok = gid < n && gather_ok(rc, pb, gid, &rec, &w);
    __shared__ uint32_t ws[4];
    const uint32_t c = uint32_t(__popcll(__ballot(ok)));
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) blk[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

// one block: blk[0..nb) <- its exclusive prefix sum, *total <- the sum
__global__ void __launch_bounds__(1024) k_scan_counts(uint32_t *blk, uint32_t nb, uint32_t *total) {
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    const uint32_t w = threadIdx.x >> 6;
    for (uint32_t b0 = 0; b0 < nb; b0 += 1024) {
        const uint32_t i = b0 + threadIdx.x;
        const uint32_t v = i < nb ? blk[i] : 0u;
        uint32_t sw;
        const uint32_t e = wave_excl_scan(v, &sw);
        if ((threadIdx.x & 63) == 0) wsum[w] = sw;
        __syncthreads();
        uint32_t o = carry;
        for (uint32_t j = 0; j < w; ++j) o += wsum[j];
        if (i < nb) blk[i] = o + e;
        __syncthreads();  // every thread has read carry and wsum
        if (threadIdx.x == 0)
            for (uint32_t j = 0; j < 16; ++j) carry += wsum[j];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ void __launch_bounds__(256) k_guide_gather_scatter(RenderConst rc, PathBuffers pb, float *samples,
                                                              uint32_t cap, uint32_t n, const uint32_t *blk,
                                                              const uint32_t *ring) {
    const uint32_t gid = blockIdx.x * 256u + threadIdx.x;
    const float *rec = nullptr;
    float w = 0;
    const bool ok = gid < n && gather_ok(rc, pb, gid, &rec, &w);
    const unsigned long long m = __ballot(ok);
    const uint32_t lane = __lane_id();
    __shared__ uint32_t ws[4];
    if (lane == 0) ws[threadIdx.x >> 6] = uint32_t(__popcll(m));
    __syncthreads();
    uint32_t idx = blk[blockIdx.x];
    for (uint32_t j = 0; j < (threadIdx.x >> 6); ++j) idx += ws[j];
    idx += uint32_t(__popcll(m & ((1ull << lane) - 1ull)));
    if (ok && idx < cap) {
        float *dst = samples + size_t((ring[0] + idx) % cap) * 12;
        for (int k = 0; k < 11; ++k) dst[k] = rec[k];
        dst[11] = w;
    }
}

// accumulate the pass into the film in sample order, Neumaier-compensated
__device__ __forceinline__ void neumaier(float *sum, float *comp, float x) {
    const float s = *sum;
    const float t = s + x;
    if (fabsf(s) >= fabsf(x)) *comp += (s - t) + x;
    else *comp += (x - t) + s;
    *sum = t;
}

__global__ void __launch_bounds__(256) k_film(RenderConst rc, PathBuffers pb) {
    const int64_t wh = int64_t(rc.W) * rc.H;
    const int64_t cover = rc.n_paths < wh ? int64_t(rc.n_paths) : wh;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= cover) return;
    const int64_t pixel = (rc.work_base + i) % wh;
    const int64_t layers = rc.n_paths >= wh ? int64_t(rc.n_paths) / wh : 1;
    // the pixel's sums stay in registers over the layers of the pass and are
    // written back once (same operations in the same order as per-layer updates)
    float sum[10], cmp[10];
#pragma unroll
    for (int k = 0; k < 10; ++k) {
        sum[k] = pb.film_sum[k * wh + pixel];
        cmp[k] = pb.film_comp[k * wh + pixel];
    }
    float count = pb.film_count[pixel];
    for (int64_t j = 0; j < layers; ++j) {
        const int64_t slot = j * wh + i;
        if (slot >= rc.n_paths) break;
        if (!(pb.flags[slot] & PF_KEEP)) continue;
        const float4 v = pb.radiance[slot];
        if (!(isfinite(v.x) && isfinite(v.y) && isfinite(v.z))) continue;
        const float4 a = pb.gb_albedo[slot], n = pb.gb_normal[slot];
        const float vals[10] = {v.x, v.y, v.z, a.x, a.y, a.z, n.x, n.y, n.z, a.w};
#pragma unroll
        for (int k = 0; k < 10; ++k) neumaier(&sum[k], &cmp[k], vals[k]);
        count += 1.0f;
    }
#pragma unroll
    for (int k = 0; k < 10; ++k) {
        pb.film_sum[k * wh + pixel] = sum[k];
        pb.film_comp[k * wh + pixel] = cmp[k];
    }
    pb.film_count[pixel] = count;
}

// GaussianFilter::eval without its normalisation (it cancels in value / weight,
// gaussian_filter.cpp:60-68)
__device__ __forceinline__ float gauss1(float d, float alpha, float expv) {
    return fmaxf(0.0f, expf(-alpha * d * d) - expv);
}

// Wide film filters: each image pixel gathers, in a fixed order (layer, then
// sampled pixel row by row), every sample of this pass that lies within the
// filter radius, weighted by the filter (FilmGrid::apply + apply_image_filter,
// render_target.h:384-395, framebuffer.h:49-80).  A sample's film position is
// re-derived from the first two numbers of its stream.
// The distance Atrc's apply_image_filter (framebuffer.h:49-80) hands the film
// filter when a sample at s (one axis) reaches pixel x of a film grid whose
// pixel range on this axis is [lo, hi], or -1 if it does not reach it: the
// grid's first pixel in reach is a = max(lo, ceil(s - r - 0.5)), its distance
// |a + 0.5 - s| grows by one per pixel after it (so it is not the distance to
// the pixel's centre), and a pixel whose running distance exceeds r is
// skipped.  The additions run in the same order as Atrc's loop.
__device__ __forceinline__ float atrc_filter_rel(int32_t x, int32_t lo, int32_t hi, float r, float s) {
    const int32_t a = max(lo, int32_t(ceilf(s - r - 0.5f)));
    const int32_t b = min(hi, int32_t(floorf(s + r - 0.5f)));
    if (x < a || x > b) return -1.0f;
    float d = fabsf(float(a) + 0.5f - s);
    for (int32_t k = a; k < x; ++k) d += 1.0f;
    return d > r ? -1.0f : d;
}

__global__ void __launch_bounds__(256) k_film_splat(RenderConst rc, PathBuffers pb) {
    const int64_t wh = int64_t(rc.W) * rc.H;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= wh) return;
    const int32_t x = int32_t(i % rc.W), y = int32_t(i / rc.W);
    // the pixel's render tile [tx0, tx1] x [ty0, ty1]: its film grid
    const int32_t tx0 = x - x % rc.tile, ty0 = y - y % rc.tile;
    const int32_t tx1 = min(tx0 + rc.tile, rc.W) - 1, ty1 = min(ty0 + rc.tile, rc.H) - 1;
    const int64_t ewh = int64_t(rc.EW) * rc.EH;
    const int64_t layers = rc.n_paths / ewh;
    const int64_t first_sample = int64_t(rc.sample_begin) + rc.work_base / ewh;
    const float r = rc.filter_r;
    const int32_t reach = int32_t(floorf(r + 0.5f));
    float *sum = pb.film_sum, *cmp = pb.film_comp;
    float wsum = pb.film_count[i];  // running weight sum, one addition per sample in sample order
    for (int64_t j = 0; j < layers; ++j) {
        for (int32_t qy = y - reach; qy <= y + reach; ++qy) {
            if (qy < -rc.border0 || qy >= rc.H + rc.border1) continue;
            for (int32_t qx = x - reach; qx <= x + reach; ++qx) {
                if (qx < -rc.border0 || qx >= rc.W + rc.border1) continue;
                const int64_t q = int64_t(qy + rc.border0) * rc.EW + (qx + rc.border0);
                const int64_t slot = j * ewh + q;
                const float4 v = pb.radiance[slot];
                if (!(isfinite(v.x) && isfinite(v.y) && isfinite(v.z))) continue;
                Rng rs = rng_seed(rc.seed, path_seq((unsigned long long)q, (unsigned long long)(first_sample + j)));
                const float sx = float(qx) + rs.uniform(), sy = float(qy) + rs.uniform();
                const float xr = atrc_filter_rel(x, tx0, tx1, r, sx), yr = atrc_filter_rel(y, ty0, ty1, r, sy);
                if (xr < 0.0f || yr < 0.0f) continue;
                const float wgt = rc.filter_kind == AGX_FILTER_GAUSSIAN
                                      ? gauss1(xr, rc.filter_alpha, rc.filter_exp) * gauss1(yr, rc.filter_alpha,
                                                                                              rc.filter_exp)
                                      : 1.0f;
                const float4 a = pb.gb_albedo[slot], n = pb.gb_normal[slot];
                const float vals[10] = {v.x, v.y, v.z, a.x, a.y, a.z, n.x, n.y, n.z, a.w};
                for (int k = 0; k < 10; ++k) neumaier(&sum[k * wh + i], &cmp[k * wh + i], wgt * vals[k]);
                wsum += wgt;
            }
        }
    }
    pb.film_count[i] = wsum;
}

__global__ void k_film_clear(PathBuffers pb, int64_t wh) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= wh) return;
    for (int k = 0; k < 10; ++k) { pb.film_sum[k * wh + i] = 0; pb.film_comp[k * wh + i] = 0; }
    pb.film_count[i] = 0;
}

__global__ void k_film_resolve(PathBuffers pb, int64_t wh, float *rgb, float *albedo, float *normal, float *denoise) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= wh) return;
    const float c = pb.film_count[i];
    const float inv = c > 0 ? 1.0f / c : 1.0f;
    float v[10];
    for (int k = 0; k < 10; ++k) v[k] = (pb.film_sum[k * wh + i] + pb.film_comp[k * wh + i]) * inv;
    for (int k = 0; k < 3; ++k) {
        rgb[3 * i + k] = v[k];
        albedo[3 * i + k] = v[3 + k];
        normal[3 * i + k] = v[6 + k];
    }
    denoise[i] = v[9];
}

__global__ void k_reset_counters(uint32_t *c) {
    if (threadIdx.x < 4) c[threadIdx.x] = 0;
}

template <bool STATS, int TM, int TB>
__global__ void __launch_bounds__(TB) k_trace_rays(DevScene s, const float4 *o, const float4 *d, uint32_t n,
                                                    int any_hit, int32_t *ent, int32_t *prim, float *t, float *u,
                                                    float *v, DevCounters *cnt) {
    AGX_LDS_COLUMN(TM, TB);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    Counts c;
    bool found = false;
    if (i < n) {
        const float4 o4 = o[i], d4 = d[i];
        if (any_hit) {
            found = query_any<STATS, TM, TB>(s, mk(o4), mk(d4), o4.w, d4.w, c, lds_col);
            ent[i] = found ? 1 : -1;
            found = false;
        } else {
            HitRec h;
            h.ent = -1; h.prim = -1; h.t = -1; h.u = 0; h.v = 0;
            found = query_closest<STATS, TM, TB>(s, mk(o4), mk(d4), o4.w, d4.w, &h, c, lds_col);
            if (found) {
                const DevEntity &e = s.ents[h.ent];
                int32_t pid = h.prim;
                if (e.kind == AGX_GEOM_TRIANGLE_BVH) pid = int32_t(fbits(s.blas_prims[3 * uint32_t(h.prim)].w));
                ent[i] = h.ent; prim[i] = pid; t[i] = h.t; u[i] = h.u; v[i] = h.v;
            } else {
                ent[i] = -1; prim[i] = -1; t[i] = -1; u[i] = 0; v[i] = 0;
            }
        }
    }
    if (STATS) flush_counts(c, found, cnt);
}

// Copies the rays of the current closest-hit (which = 0) or any-hit (1) queue,
// in queue order, into a capture buffer (ray replay measurements).
__global__ void __launch_bounds__(256) k_capture_rays(PathBuffers pb, int which, float4 *o, float4 *d,
                                                      uint32_t *count, uint32_t cap) {
    const uint32_t n = pb.cnt_in[which];
    if (blockIdx.x * blockDim.x >= n) return;
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = tid < n;
    const uint32_t idx = wave_append(count, live);
    if (live && idx < cap) {
        if (which == 0) {
            const uint32_t ri = pb.q_closest[tid];
            o[idx] = pb.ray_o[ri];
            d[idx] = pb.ray_d[ri];
        } else {
            // shadow rays keep (o, t_max) (d, t_min); captured as (o, t_min) (d, t_max)
            const uint32_t ri = pb.q_shadow[tid];
            const float4 so = pb.sh_o[ri], sd = pb.sh_d[ri];
            o[idx] = make_float4(so.x, so.y, so.z, sd.w);
            d[idx] = make_float4(sd.x, sd.y, sd.z, so.w);
        }
    }
}

// plain streaming copy, used to calibrate HBM traffic counters
__global__ void __launch_bounds__(256) k_copy_calibration(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                          uint64_t n) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

inline unsigned grid_of(uint64_t n, unsigned b) { return unsigned((n + b - 1) / b); }

}  // namespace

// ---------------------------------------------------------------- launchers
void launch_raygen(const DevScene &s, const RenderConst &rc, const PathBuffers &pb, hipStream_t st) {
    if (rc.n_paths <= 0) return;
    hipLaunchKernelGGL(k_raygen, dim3(grid_of(uint64_t(rc.n_paths), 256)), dim3(256), 0, st, s, rc, pb);
}
#define AGX_LAUNCH_TM(K, TM, TB, grid, st, ...)                                                  \
    do {                                                                                          \
        if (stats) hipLaunchKernelGGL((K<true, TM, TB>), grid, dim3(TB), 0, st, __VA_ARGS__);     \
        else hipLaunchKernelGGL((K<false, TM, TB>), grid, dim3(TB), 0, st, __VA_ARGS__);          \
    } while (0)
// tb = threads per block: 128 or 256 in the BVH4 modes, 256 in the Atrc modes
#define AGX_LAUNCH_STACKED(K, tb, grid, st, ...)                                                  \
    do {                                                                                          \
        if (s.accel == AGX_ACCEL_BVH4) {                                                          \
            if ((tb) == 128) {                                                                    \
                if (s.bvh4_large_stack) AGX_LAUNCH_TM(K, TM_B4_LARGE, 128, grid, st, __VA_ARGS__);\
                else AGX_LAUNCH_TM(K, TM_B4_SMALL, 128, grid, st, __VA_ARGS__);                   \
            } else {                                                                              \
                if (s.bvh4_large_stack) AGX_LAUNCH_TM(K, TM_B4_LARGE, 256, grid, st, __VA_ARGS__);\
                else AGX_LAUNCH_TM(K, TM_B4_SMALL, 256, grid, st, __VA_ARGS__);                   \
            }                                                                                     \
        } else if (s.accel == AGX_ACCEL_BVH4_WATERTIGHT) {                                        \
            if ((tb) == 128) {                                                                    \
                if (s.bvh4_large_stack) AGX_LAUNCH_TM(K, TM_B4W_LARGE, 128, grid, st, __VA_ARGS__);\
                else AGX_LAUNCH_TM(K, TM_B4W_SMALL, 128, grid, st, __VA_ARGS__);                  \
            } else {                                                                              \
                if (s.bvh4_large_stack) AGX_LAUNCH_TM(K, TM_B4W_LARGE, 256, grid, st, __VA_ARGS__);\
                else AGX_LAUNCH_TM(K, TM_B4W_SMALL, 256, grid, st, __VA_ARGS__);                  \
            }                                                                                     \
        } else if (s.lds_stack) {                                                                 \
            AGX_LAUNCH_TM(K, TM_ATRC_LDS, 256, grid, st, __VA_ARGS__);                            \
        } else {                                                                                  \
            AGX_LAUNCH_TM(K, TM_ATRC_PRIV, 256, grid, st, __VA_ARGS__);                           \
        }                                                                                         \
    } while (0)

// threads per block of the wavefront's traversal launches for this scene
inline unsigned trace_block_of(const DevScene &s) {
    const bool b4 = s.accel == AGX_ACCEL_BVH4 || s.accel == AGX_ACCEL_BVH4_WATERTIGHT;
    return (b4 && s.trace_block == 128) ? 128u : TRACE_BLOCK;
}

void launch_trace_closest(const DevScene &s, const PathBuffers &pb, uint32_t max_n, bool stats, DevCounters *cnt,
                          hipStream_t st) {
    if (!max_n) return;
    const unsigned tb = trace_block_of(s);
    AGX_LAUNCH_STACKED(k_trace_closest, tb, dim3(grid_of(max_n, tb) + 7), st, s, pb, cnt);
}
void launch_trace_shadow(const DevScene &s, const PathBuffers &pb, uint32_t max_n, bool stats, DevCounters *cnt,
                         hipStream_t st) {
    if (!max_n) return;
    const unsigned tb = trace_block_of(s);
    AGX_LAUNCH_STACKED(k_trace_shadow, tb, dim3(grid_of(max_n, tb) + 7), st, s, pb, cnt);
}
void launch_shade(const DevScene &s, const RenderConst &rc, const PathBuffers &pb, uint32_t max_n, hipStream_t st) {
    if (!max_n) return;
    // lean variants serve scenes of Lambert / GGX / black surfaces in vacuum,
    // unguided or guided; every other scene takes the full kernel
    const bool lean = (s.features & (SF_MEDIA | SF_DISNEY | SF_DELTA | SF_IBL)) == 0;
    // media and delta surfaces without Disney, IBL or guiding (the volumetric
    // scenes): 199 VGPRs and no scratch against the full variant's spills
    const bool media_delta = (s.features & (SF_DISNEY | SF_IBL)) == 0;
    if (lean && !rc.guide) hipLaunchKernelGGL(k_shade<0u>, dim3(grid_of(max_n, 256)), dim3(256), 0, st, s, rc, pb);
    else if (lean) hipLaunchKernelGGL(k_shade<SF_GUIDE>, dim3(grid_of(max_n, 256)), dim3(256), 0, st, s, rc, pb);
    else if (media_delta && !rc.guide)
        hipLaunchKernelGGL(k_shade<SF_MEDIA | SF_DELTA>, dim3(grid_of(max_n, 256)), dim3(256), 0, st, s, rc, pb);
    else if (media_delta)  // guided volumetric scenes
        hipLaunchKernelGGL(k_shade<SF_MEDIA | SF_DELTA | SF_GUIDE>, dim3(grid_of(max_n, 256)), dim3(256), 0, st, s,
                           rc, pb);
    else hipLaunchKernelGGL(k_shade<SF_ALL>, dim3(grid_of(max_n, 256)), dim3(256), 0, st, s, rc, pb);
}

uint32_t queue_compact_blocks(uint32_t max_n) { return (max_n + QC_SPAN - 1) / QC_SPAN + 1; }

void launch_queue_compact(const PathBuffers &pb, uint32_t max_n, int L, uint4 *blk, hipStream_t st) {
    const uint32_t nb = (max_n + QC_SPAN - 1) / QC_SPAN;
    const uint32_t nl = uint32_t(L > 0 ? L : 0);
    if (nb) hipLaunchKernelGGL(k_queue_count, dim3(nb), dim3(QC_THREADS), 0, st, pb, nl, blk);
    hipLaunchKernelGGL(k_queue_scan, dim3(1), dim3(1024), 0, st, pb, blk);
    if (nb) hipLaunchKernelGGL(k_queue_scatter, dim3(nb), dim3(QC_THREADS), 0, st, pb, nl, blk);
}

void launch_medium_pass(const DevScene &s, const RenderConst &rc, const PathBuffers &pb, uint32_t max_n,
                        hipStream_t st) {
    if (!max_n) return;
    hipLaunchKernelGGL(k_medium_pass, dim3(grid_of(max_n, 256)), dim3(256), 0, st, s, rc, pb);
}

void launch_guide_features(const DevScene &s, const RenderConst &rc, const PathBuffers &pb, uint32_t max_n,
                           hipStream_t st) {
    if (!max_n) return;
    hipLaunchKernelGGL(k_guide_features, dim3(grid_of(max_n, 256)), dim3(256), 0, st, s, rc, pb);
}

void launch_guide_gather(const RenderConst &rc, const PathBuffers &pb, float *samples, uint32_t cap,
                         uint32_t *count, const uint32_t *ring, uint32_t *blk, hipStream_t st) {
    uint64_t n = rc.n_paths > 0 ? ((uint64_t(rc.n_paths - 1) >> rc.guide_rec_shift) + 1) * GUIDE_MAX_REC : 0;
    // the recording stride keeps at most GUIDE_REC_PATHS record slots (blk holds their blocks)
    n = std::min<uint64_t>(n, uint64_t(GUIDE_REC_PATHS) * GUIDE_MAX_REC);
    const uint32_t nb = grid_of(n, 256);
    if (nb) hipLaunchKernelGGL(k_guide_gather_count, dim3(nb), dim3(256), 0, st, rc, pb, uint32_t(n), blk);
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, st, blk, nb, count);
    if (nb)
        hipLaunchKernelGGL(k_guide_gather_scatter, dim3(nb), dim3(256), 0, st, rc, pb, samples, cap, uint32_t(n), blk,
                           ring);
}
// rays traced by the bounce about to run (closest-hit + shadow queue sizes),
// summed on the device; one thread, stream-ordered
__global__ void k_count_rays(const uint32_t *cin, unsigned long long *acc) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *acc += (unsigned long long)cin[0] + (unsigned long long)cin[1];
}
void launch_count_rays(const uint32_t *cin, unsigned long long *acc, hipStream_t st) {
    hipLaunchKernelGGL(k_count_rays, dim3(1), dim3(64), 0, st, cin, acc);
}
void launch_copy_calibration(const float4 *src, float4 *dst, uint64_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_copy_calibration, dim3(2048), dim3(256), 0, st, src, dst, n);
}
void launch_reset_counters(uint32_t *c, hipStream_t st) {
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, st, c);
}
void launch_film(const RenderConst &rc, const PathBuffers &pb, hipStream_t st) {
    const int64_t wh = int64_t(rc.W) * rc.H;
    if (rc.splat) {
        hipLaunchKernelGGL(k_film_splat, dim3(grid_of(uint64_t(wh), 256)), dim3(256), 0, st, rc, pb);
        return;
    }
    const int64_t cover = rc.n_paths < wh ? int64_t(rc.n_paths) : wh;
    if (cover <= 0) return;
    hipLaunchKernelGGL(k_film, dim3(grid_of(uint64_t(cover), 256)), dim3(256), 0, st, rc, pb);
}
void launch_film_clear(const PathBuffers &pb, int64_t wh, hipStream_t st) {
    hipLaunchKernelGGL(k_film_clear, dim3(grid_of(uint64_t(wh), 256)), dim3(256), 0, st, pb, wh);
}
void launch_film_resolve(const PathBuffers &pb, int64_t wh, float *rgb, float *albedo, float *normal, float *denoise,
                         hipStream_t st) {
    hipLaunchKernelGGL(k_film_resolve, dim3(grid_of(uint64_t(wh), 256)), dim3(256), 0, st, pb, wh, rgb, albedo,
                       normal, denoise);
}
void launch_capture_rays(const PathBuffers &pb, int which, float4 *o, float4 *d, uint32_t *count, uint32_t cap,
                         uint32_t max_n, hipStream_t st) {
    if (!max_n) return;
    hipLaunchKernelGGL(k_capture_rays, dim3(grid_of(max_n, 256)), dim3(256), 0, st, pb, which, o, d, count, cap);
}
// ---------------------------------------------------------------- staging without allocation
// The staging buffer of agx_trace_rays_noalloc.  Layout for n rays (n <=
// NOALLOC_MAX_RAYS), offsets in units of NOALLOC_MAX_RAYS 4-byte words:
//   [0, 4) packed (o, tmin) float4   [4, 8) packed (d, tmax) float4
//   [8, 11) o (3 floats per ray)     [11, 14) d   [14] tmin   [15] tmax
//   [16] entity  [17] primitive  [18] t  [19] u  [20] v
// Every GPU holds its own copy until the process ends; host code uses the
// address hipGetSymbolAddress gives on that device, never the array's name.
static __device__ __align__(256) uint32_t scratch[21 * size_t(NOALLOC_MAX_RAYS)];
static_assert(21 * 4 == NOALLOC_BYTES_PER_RAY, "staging layout");

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

__global__ void __launch_bounds__(256) k_pack_rays(uint32_t *base, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t M = NOALLOC_MAX_RAYS;
    const float *o = reinterpret_cast<const float *>(base + 8 * M), *d = reinterpret_cast<const float *>(base + 11 * M);
    const float *tmin = reinterpret_cast<const float *>(base + 14 * M), *tmax = reinterpret_cast<const float *>(base + 15 * M);
    float4 *po = reinterpret_cast<float4 *>(base), *pd = reinterpret_cast<float4 *>(base + 4 * M);
    po[i] = make_float4(o[3 * i], o[3 * i + 1], o[3 * i + 2], tmin[i]);
    pd[i] = make_float4(d[3 * i], d[3 * i + 1], d[3 * i + 2], tmax[i]);
}

__global__ void k_zero_counters(DevCounters *c) {
    if (threadIdx.x < 2) c[threadIdx.x] = DevCounters{0, 0, 0, 0};
}
}  // namespace

hipError_t noalloc_scratch_init(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = p;
    return hipSuccess;
}

void *noalloc_scratch(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[device];
}

void launch_trace_rays_staged(const DevScene &s, void *scratch_ptr, uint32_t n, int any_hit, DevCounters *cnt,
                              hipStream_t st) {
    if (!n) return;
    uint32_t *base = static_cast<uint32_t *>(scratch_ptr);
    const size_t M = NOALLOC_MAX_RAYS;
    hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, st, cnt);
    hipLaunchKernelGGL(k_pack_rays, dim3(grid_of(n, 256)), dim3(256), 0, st, base, n);
    launch_trace_rays(s, reinterpret_cast<const float4 *>(base), reinterpret_cast<const float4 *>(base + 4 * M), n,
                      any_hit, reinterpret_cast<int32_t *>(base + 16 * M), reinterpret_cast<int32_t *>(base + 17 * M),
                      reinterpret_cast<float *>(base + 18 * M), reinterpret_cast<float *>(base + 19 * M),
                      reinterpret_cast<float *>(base + 20 * M), cnt + (any_hit ? 1 : 0), st);
}

void launch_trace_rays(const DevScene &s, const float4 *o, const float4 *d, uint32_t n, int any_hit, int32_t *ent,
                       int32_t *prim, float *t, float *u, float *v, DevCounters *cnt, hipStream_t st) {
    if (!n) return;
    const bool stats = cnt != nullptr;
    AGX_LAUNCH_STACKED(k_trace_rays, TRACE_BLOCK, dim3(grid_of(n, TRACE_BLOCK)), st, s, o, d, n, any_hit, ent, prim, t, u, v, cnt);
}

}  // namespace agx
