// ORACLE — test infrastructure only.
//
// CPU restatement of Atrc's `pt` renderer (trace_std + PerPixelRenderer) and of
// the scene queries beneath it, templated on the real type: float is the FP32
// restatement that the GPU path is checked against ray for ray and path for
// path; double is the FP64 high-spp reference.  Only tests, the smoke check and
// the CPU-baseline leg of the benchmark use it, as the reference results are
// compared with; the product (atrc_amd/) never links it.
//
// Every function cites the Atrc file:line it follows (paths relative to
// src/tracer/ unless they start with src/).  Where Atrc calls its agz-utils
// submodule, which is not part of the available source, the function states
// the published method it uses instead; those are pinned by the tests in
// tests/test_oracle_*.py against closed forms, not against Atrc.
//
// Deviations from Atrc, each deliberate:
//  * RNG: one PCG32 stream per (pixel, sample) instead of one per worker thread
//    (src/core/renderer/perpixel_renderer.cpp:77-80), so a render is
//    reproducible and the GPU can replay the same random numbers.
//  * uniform float = (pcg32() >> 8) * 2^-24 instead of libstdc++'s
//    uniform_real_distribution<float>.
//  * Film sums use Neumaier-compensated accumulation.
//  * Compiled with -ffp-contract=off so the FP32 arithmetic is the one written.
#pragma once

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <queue>
#include <stack>
#include <stdexcept>
#include <thread>
#include <vector>

#include "../include/atrc_gpu.h"

namespace oracle {

// ------------------------------------------------------------------------------------
// vector math (agz-utils tvec3 / float3 semantics: component-wise, dot = (xx+yy)+zz)
// ------------------------------------------------------------------------------------

template<class T> struct V3 {
    T x = 0, y = 0, z = 0;
    V3() = default;
    V3(T a, T b, T c) : x(a), y(b), z(c) {}
    explicit V3(T a) : x(a), y(a), z(a) {}
    T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    T &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    V3 operator+(const V3 &o) const { return {x + o.x, y + o.y, z + o.z}; }
    V3 operator-(const V3 &o) const { return {x - o.x, y - o.y, z - o.z}; }
    V3 operator*(const V3 &o) const { return {x * o.x, y * o.y, z * o.z}; }
    V3 operator/(const V3 &o) const { return {x / o.x, y / o.y, z / o.z}; }
    V3 operator*(T s) const { return {x * s, y * s, z * s}; }
    V3 operator/(T s) const { return {x / s, y / s, z / s}; }
    V3 operator-() const { return {-x, -y, -z}; }
    V3 &operator+=(const V3 &o) { x += o.x; y += o.y; z += o.z; return *this; }
    V3 &operator*=(const V3 &o) { x *= o.x; y *= o.y; z *= o.z; return *this; }
    V3 &operator*=(T s) { x *= s; y *= s; z *= s; return *this; }
    V3 &operator/=(T s) { x /= s; y /= s; z /= s; return *this; }
    bool is_black() const { return x == 0 && y == 0 && z == 0; }
    bool is_finite() const { return std::isfinite(x) && std::isfinite(y) && std::isfinite(z); }
    T length_square() const { return x * x + y * y + z * z; }
    T length() const { return std::sqrt(length_square()); }
    V3 normalize() const { return *this / length(); }
    T max_elem() const { return std::max(x, std::max(y, z)); }
};
template<class T> inline V3<T> operator*(T s, const V3<T> &v) { return v * s; }
template<class T> inline T dot(const V3<T> &a, const V3<T> &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template<class T> inline V3<T> cross(const V3<T> &a, const V3<T> &b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template<class T> inline T cos_between(const V3<T> &a, const V3<T> &b) { return dot(a, b) / (a.length() * b.length()); }
template<class T> inline V3<T> vsqrt(const V3<T> &a) { return {std::sqrt(a.x), std::sqrt(a.y), std::sqrt(a.z)}; }

template<class T> struct V2 {
    T x = 0, y = 0;
    V2() = default;
    V2(T a, T b) : x(a), y(b) {}
    V2 operator+(const V2 &o) const { return {x + o.x, y + o.y}; }
    V2 operator-(const V2 &o) const { return {x - o.x, y - o.y}; }
    V2 operator*(T s) const { return {x * s, y * s}; }
};
template<class T> inline V2<T> operator*(T s, const V2<T> &v) { return v * s; }

template<class T> constexpr T PI = T(3.14159265358979323846);

// Orthonormal frame (agz-utils tcoord3; unpinned: methods named per function).
template<class T> struct Coord {
    V3<T> x, y, z;
    Coord() : x(1, 0, 0), y(0, 1, 0), z(0, 0, 1) {}
    Coord(const V3<T> &a, const V3<T> &b, const V3<T> &c) : x(a), y(b), z(c) {}
    // Duff, Burgess, Christensen et al., "Building an Orthonormal Basis, Revisited", JCGT 2017.
    static Coord from_z(const V3<T> &zin) {
        const V3<T> n = zin.normalize();
        const T sign = std::copysign(T(1), n.z);
        const T a = T(-1) / (sign + n.z);
        const T b = n.x * n.y * a;
        const V3<T> ex(T(1) + sign * n.x * n.x * a, sign * b, -sign * n.x);
        const V3<T> ey(b, sign + n.y * n.y * a, -n.y);
        return Coord(ex, ey, n);
    }
    V3<T> local_to_global(const V3<T> &v) const { return x * v.x + y * v.y + z * v.z; }
    V3<T> global_to_local(const V3<T> &v) const { return {dot(v, x), dot(v, y), dot(v, z)}; }
    bool in_positive_z_hemisphere(const V3<T> &v) const { return dot(v, z) > 0; }
    // Gram-Schmidt of the old x against the new z; y completes a right-handed frame.
    Coord rotate_to_new_z(const V3<T> &new_z) const {
        const V3<T> nz = new_z.normalize();
        const V3<T> xp = x - nz * dot(x, nz);
        const T l = xp.length();
        if (!(l > T(1e-6))) return from_z(nz);
        const V3<T> nx = xp / l;
        return Coord(nx, cross(nz, nx), nz);
    }
};

// Row-major 3x4 affine transform.
template<class T> struct Affine {
    T m[12];
    static Affine from(const float *f) { Affine a; for (int i = 0; i < 12; ++i) a.m[i] = T(f[i]); return a; }
    V3<T> point(const V3<T> &p) const {
        return {m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3],
                m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
                m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]};
    }
    V3<T> vec(const V3<T> &v) const {
        return {m[0] * v.x + m[1] * v.y + m[2] * v.z,
                m[4] * v.x + m[5] * v.y + m[6] * v.z,
                m[8] * v.x + m[9] * v.y + m[10] * v.z};
    }
    // transpose of the linear part applied to v (normals use the inverse's transpose)
    V3<T> tvec(const V3<T> &v) const {
        return {m[0] * v.x + m[4] * v.y + m[8] * v.z,
                m[1] * v.x + m[5] * v.y + m[9] * v.z,
                m[2] * v.x + m[6] * v.y + m[10] * v.z};
    }
};

// ------------------------------------------------------------------------------------
// RNG: PCG32 (O'Neill 2014, XSH-RR 64/32), one stream per (pixel, sample)
// ------------------------------------------------------------------------------------

inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

struct Pcg32 {
    uint64_t state = 0, inc = 1;
    static constexpr uint64_t MULT = 6364136223846793005ull;
    uint32_t next() {
        const uint64_t old = state;
        state = old * MULT + inc;
        const uint32_t xorshifted = uint32_t(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = uint32_t(old >> 59u);
        return (xorshifted >> rot) | (xorshifted << ((-rot) & 31));
    }
    // pcg32_srandom_r(initstate, initseq)
    void seed(uint64_t initstate, uint64_t initseq) {
        state = 0;
        inc = (initseq << 1u) | 1u;
        next();
        state += initstate;
        next();
    }
    // stream of sample `sample` of pixel `pixel` (pixel = py * width + px)
    void seed_path(uint64_t seed, uint64_t pixel, uint64_t sample) {
        const uint64_t seq = (pixel << 24) ^ sample;
        this->seed(splitmix64(seed ^ splitmix64(seq)), seq);
    }
    float uniform() { return float(next() >> 8) * 0x1p-24f; }
};

// key of the second per-path stream (ratio-tracking transmittance draws)
constexpr uint64_t TR_STREAM_KEY = 0xA5A5F00DC0FFEE11ull;

template<class T> struct Sampler {
    Pcg32 rng;
    T u1() { return T(rng.uniform()); }
};

// ------------------------------------------------------------------------------------
// sampling distributions (agz-utils math::distribution — unpinned; methods named)
// ------------------------------------------------------------------------------------

// Shirley & Chiu 1997 concentric square-to-disk map.
template<class T> inline V2<T> uniform_on_unit_disk(T u, T v) {
    const T sx = 2 * u - 1, sy = 2 * v - 1;
    if (sx == 0 && sy == 0) return {0, 0};
    T r, theta;
    if (std::abs(sx) > std::abs(sy)) { r = sx; theta = (PI<T> / 4) * (sy / sx); }
    else { r = sy; theta = PI<T> / 2 - (PI<T> / 4) * (sx / sy); }
    return {r * std::cos(theta), r * std::sin(theta)};
}
// Malley's method: cosine-weighted hemisphere; pdf = cos/pi.
template<class T> inline void zweighted_on_hemisphere(T u, T v, V3<T> *dir, T *pdf) {
    const V2<T> d = uniform_on_unit_disk(u, v);
    const T z = std::sqrt(std::max(T(0), 1 - d.x * d.x - d.y * d.y));
    *dir = V3<T>(d.x, d.y, z);
    *pdf = z / PI<T>;
}
template<class T> inline T zweighted_on_hemisphere_pdf(T cos_theta) { return cos_theta > 0 ? cos_theta / PI<T> : 0; }
// z = 1 - 2u, phi = 2 pi v; pdf = 1/(4 pi).
template<class T> inline V3<T> uniform_on_sphere(T u, T v) {
    const T z = 1 - 2 * u;
    const T r = std::sqrt(std::max(T(0), 1 - z * z));
    const T phi = 2 * PI<T> * v;
    return {r * std::cos(phi), r * std::sin(phi), z};
}
template<class T> constexpr T uniform_on_sphere_pdf = T(1) / (4 * PI<T>);
// pbrt UniformSampleCone; pdf = 1 / (2 pi (1 - cos_max)).
template<class T> inline V3<T> uniform_on_cone(T cos_max, T u, T v) {
    const T cos_t = (1 - u) + u * cos_max;
    const T sin_t = std::sqrt(std::max(T(0), 1 - cos_t * cos_t));
    const T phi = 2 * PI<T> * v;
    return {std::cos(phi) * sin_t, std::sin(phi) * sin_t, cos_t};
}
template<class T> inline T uniform_on_cone_pdf(T cos_max) { return 1 / (2 * PI<T> * (1 - cos_max)); }
// Barycentric (b1, b2) of a uniform point: (1 - sqrt(u), v sqrt(u)).
template<class T> inline V2<T> uniform_on_triangle(T u, T v) {
    const T su = std::sqrt(u);
    return {1 - su, v * su};
}

// Walker/Vose alias table, single-uniform lookup.
template<class T> struct AliasTable {
    std::vector<T> prob;
    std::vector<int32_t> alias;
    void build(const std::vector<T> &w) {
        const size_t n = w.size();
        prob.assign(n, T(1));
        alias.assign(n, 0);
        double sum = 0;
        for (T x : w) sum += double(x);
        std::vector<double> scaled(n);
        std::vector<int32_t> small, large;
        for (size_t i = 0; i < n; ++i) {
            scaled[i] = double(w[i]) * double(n) / sum;
            alias[i] = int32_t(i);
            (scaled[i] < 1.0 ? small : large).push_back(int32_t(i));
        }
        while (!small.empty() && !large.empty()) {
            const int32_t s = small.back(); small.pop_back();
            const int32_t l = large.back(); large.pop_back();
            prob[s] = T(scaled[s]);
            alias[s] = l;
            scaled[l] = (scaled[l] + scaled[s]) - 1.0;
            (scaled[l] < 1.0 ? small : large).push_back(l);
        }
        for (int32_t i : large) prob[i] = T(1);
        for (int32_t i : small) prob[i] = T(1);
    }
    int32_t sample(T u) const {
        const int32_t n = int32_t(prob.size());
        const T nu = u * T(n);
        int32_t idx = int32_t(nu);
        if (idx >= n) idx = n - 1;
        const T frac = nu - T(idx);
        return frac < prob[idx] ? idx : alias[idx];
    }
};

// ------------------------------------------------------------------------------------
// rays, AABBs, triangles  (common.h:256-395, utility/triangle_aux.h:9-88)
// ------------------------------------------------------------------------------------

template<class T> struct Ray {
    V3<T> o, d;
    T t_min = 0, t_max = std::numeric_limits<T>::infinity();
    Ray() = default;
    Ray(const V3<T> &o_, const V3<T> &d_, T tmin = 0, T tmax = std::numeric_limits<T>::infinity())
        : o(o_), d(d_), t_min(tmin), t_max(tmax) {}
    V3<T> at(T t) const { return o + d * t; }
    bool between(T t) const { return t_min <= t && t <= t_max; }  // common.h:320-323
};

template<class T> struct AABB {
    V3<T> low{std::numeric_limits<T>::max()}, high{std::numeric_limits<T>::lowest()};
    void grow(const V3<T> &p) {
        low.x = std::min(low.x, p.x); low.y = std::min(low.y, p.y); low.z = std::min(low.z, p.z);
        high.x = std::max(high.x, p.x); high.y = std::max(high.y, p.y); high.z = std::max(high.z, p.z);
    }
    void grow(const AABB &b) {
        low.x = std::min(low.x, b.low.x); low.y = std::min(low.y, b.low.y); low.z = std::min(low.z, b.low.z);
        high.x = std::max(high.x, b.high.x); high.y = std::max(high.y, b.high.y); high.z = std::max(high.z, b.high.z);
    }
    // AABB::intersect (common.h:384-395), written with std::min/std::max operand order
    // of triangle_bvh.cpp:87-93 (the SIMD elem_min order of agz-utils is unpinned).
    bool intersect(const V3<T> &o, const V3<T> &inv, T t_min, T t_max) const {
        T tn = t_min, tf = t_max;
        for (int a = 0; a < 3; ++a) {
            const T n = inv[a] * (low[a] - o[a]);
            const T f = inv[a] * (high[a] - o[a]);
            tn = std::max(tn, std::min(n, f));
            tf = std::min(tf, std::max(n, f));
        }
        return tn <= tf;
    }
};

struct Counters {
    uint64_t closest = 0, shadow = 0, node_visits = 0, tri_tests = 0, hits = 0;
};

// has_intersection_with_triangle (triangle_aux.h:9-30)
template<class T>
inline bool mt_any(const Ray<T> &r, const V3<T> &A, const V3<T> &B_A, const V3<T> &C_A) {
    const V3<T> s1 = cross(r.d, C_A);
    const T div = dot(s1, B_A);
    if (!div) return false;
    const T inv_div = 1 / div;
    const V3<T> o_A = r.o - A;
    const T alpha = dot(o_A, s1) * inv_div;
    if (alpha < 0 || alpha > 1) return false;
    const V3<T> s2 = cross(o_A, B_A);
    const T beta = dot(r.d, s2) * inv_div;
    if (beta < 0 || alpha + beta > 1) return false;
    const T t = dot(C_A, s2) * inv_div;
    return r.between(t);
}

// closest_intersection_with_triangle (triangle_aux.h:38-69)
template<class T>
inline bool mt_closest(const Ray<T> &r, const V3<T> &A, const V3<T> &B_A, const V3<T> &C_A,
                       T *t_out, T *u_out, T *v_out) {
    const V3<T> s1 = cross(r.d, C_A);
    const T div = dot(s1, B_A);
    if (!div) return false;
    const T inv_div = 1 / div;
    const V3<T> o_A = r.o - A;
    const T alpha = dot(o_A, s1) * inv_div;
    if (alpha < 0) return false;
    const V3<T> s2 = cross(o_A, B_A);
    const T beta = dot(r.d, s2) * inv_div;
    if (beta < 0 || alpha + beta > 1) return false;
    const T t = dot(C_A, s2) * inv_div;
    if (!r.between(t)) return false;
    *t_out = t; *u_out = alpha; *v_out = beta;
    return true;
}

// Watertight ray/triangle test (Woop, Benthin and Wald, "Watertight
// Ray/Triangle Intersection", JCGT 2(1), 2013), the GPU path's production
// intersection mode (not an Atrc function).  The ray is sheared onto +z
// through the origin; the three 2D edge functions of the projected vertices
// decide the hit, and an edge function that is exactly zero is recomputed in
// double from the same projected coordinates.  Returns t and the barycentrics
// of B and C (the Moller-Trumbore u, v).
template<class T> struct ShearedRay {
    V3<T> o;
    int kx = 0, ky = 1, kz = 2;
    T sx = 0, sy = 0, sz = 1;
    explicit ShearedRay(const Ray<T> &r) : o(r.o) {
        const T ax = std::abs(r.d.x), ay = std::abs(r.d.y), az = std::abs(r.d.z);
        kz = ax > ay ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
        kx = (kz + 1) % 3;
        ky = (kx + 1) % 3;
        if (r.d[kz] < 0) std::swap(kx, ky);
        sx = r.d[kx] / r.d[kz];
        sy = r.d[ky] / r.d[kz];
        sz = T(1) / r.d[kz];
    }
};
template<class T>
inline bool watertight_closest(const ShearedRay<T> &w, T tmin, T tmax, const V3<T> &A, const V3<T> &B,
                               const V3<T> &C, T *t_out, T *u_out, T *v_out) {
    const V3<T> a = A - w.o, b = B - w.o, c = C - w.o;
    const T ax = a[w.kx] - w.sx * a[w.kz], ay = a[w.ky] - w.sy * a[w.kz];
    const T bx = b[w.kx] - w.sx * b[w.kz], by = b[w.ky] - w.sy * b[w.kz];
    const T cx = c[w.kx] - w.sx * c[w.kz], cy = c[w.ky] - w.sy * c[w.kz];
    T ea = cx * by - cy * bx;  // opposite A: weight of A
    T eb = ax * cy - ay * cx;  // weight of B
    T ec = bx * ay - by * ax;  // weight of C
    if (ea == 0 || eb == 0 || ec == 0) {
        ea = T(double(cx) * double(by) - double(cy) * double(bx));
        eb = T(double(ax) * double(cy) - double(ay) * double(cx));
        ec = T(double(bx) * double(ay) - double(by) * double(ax));
    }
    const bool neg = ea < 0 || eb < 0 || ec < 0, pos = ea > 0 || eb > 0 || ec > 0;
    if (neg && pos) return false;
    const T det = ea + eb + ec;
    if (det == 0) return false;
    const T tsc = ea * (w.sz * a[w.kz]) + eb * (w.sz * b[w.kz]) + ec * (w.sz * c[w.kz]);
    const T inv = T(1) / det;
    const T t = tsc * inv;
    if (!(tmin <= t && t <= tmax)) return false;
    *t_out = t; *u_out = eb * inv; *v_out = ec * inv;
    return true;
}

template<class T> inline T triangle_area(const V3<T> &B_A, const V3<T> &C_A) { return cross(B_A, C_A).length() / 2; }

// dpdu_as_ex (triangle_aux.h:76-88)
template<class T>
inline V3<T> dpdu_as_ex(const V3<T> &B_A, const V3<T> &C_A, const V2<T> &b_a, const V2<T> &c_a, const V3<T> &nor) {
    const T m00 = b_a.x, m01 = b_a.y, m10 = c_a.x, m11 = c_a.y;
    const T det = m00 * m11 - m01 * m10;
    if (!det) return Coord<T>::from_z(nor).x;
    const T inv_det = 1 / det;
    return (B_A * (m11 * inv_det) - C_A * (m01 * inv_det)).normalize();
}

// ------------------------------------------------------------------------------------
// intersections  (core/intersection.h:10-93)
// ------------------------------------------------------------------------------------

template<class T> struct SurfacePoint {
    V3<T> pos;
    V2<T> uv;
    Coord<T> geometry_coord, user_coord;
    V3<T> eps_offset(const V3<T> &dir, T eps) const {  // intersection.h:17-22
        if (dot(dir, geometry_coord.z) > 0) return pos + geometry_coord.z * eps;
        return pos - geometry_coord.z * eps;
    }
};

template<class T> struct GeometryHit : SurfacePoint<T> {
    T t = -1;
    V3<T> wr;
    int32_t prim = 0;
    T bu = 0, bv = 0;  // barycentrics of the hit triangle (parity output)
};

template<class T> struct EntityHit : GeometryHit<T> {
    int32_t entity = -1;
};

// ------------------------------------------------------------------------------------
// geometry
// ------------------------------------------------------------------------------------

// Atrc's default triangle BVH (core/geometry/triangle_bvh.cpp:24-488).
template<class T> struct TriangleBVH {
    struct Node { T low[3], high[3]; uint32_t start, end_or_right; bool is_leaf() const { return start != 0xffffffffu; } };
    struct Prim { V3<T> a, b_a, c_a; };
    struct PrimInfo { V3<T> n_a, n_b_a, n_c_a; V2<T> t_a, t_b_a, t_c_a; V3<T> x, z; };
    std::vector<Node> nodes;
    std::vector<Prim> prims;
    std::vector<PrimInfo> infos;
    std::vector<uint32_t> orig_id;  // original triangle index of BVH slot i
    AliasTable<T> area_table;
    T surface_area = 0;
    AABB<T> world_bound;
    int max_depth = 0;

    // Node::has_intersection (triangle_bvh.cpp:75-97)
    static bool node_hit(const Node &nd, const T *ori, const T *inv, T t_min, T t_max, T *t_enter) {
        const T nx = inv[0] * (nd.low[0] - ori[0]);
        const T ny = inv[1] * (nd.low[1] - ori[1]);
        const T nz = inv[2] * (nd.low[2] - ori[2]);
        const T fx = inv[0] * (nd.high[0] - ori[0]);
        const T fy = inv[1] * (nd.high[1] - ori[1]);
        const T fz = inv[2] * (nd.high[2] - ori[2]);
        t_min = std::max(t_min, std::min(nx, fx));
        t_min = std::max(t_min, std::min(ny, fy));
        t_min = std::max(t_min, std::min(nz, fz));
        t_max = std::min(t_max, std::max(nx, fx));
        t_max = std::min(t_max, std::max(ny, fy));
        t_max = std::min(t_max, std::max(nz, fz));
        *t_enter = t_min;
        return t_min <= t_max;
    }

    void build(const agx_vertex *v, uint32_t n) {
        struct BTri { uint32_t id; V3<T> p[3]; V3<T> nor[3]; V2<T> uv[3]; V3<T> centroid; };
        std::vector<BTri> tris(n);
        surface_area = 0;
        world_bound = AABB<T>();
        for (uint32_t i = 0; i < n; ++i) {
            BTri &b = tris[i];
            b.id = i;
            for (int k = 0; k < 3; ++k) {
                const agx_vertex &vv = v[3 * i + k];
                b.p[k] = V3<T>(T(vv.pos[0]), T(vv.pos[1]), T(vv.pos[2]));
                b.nor[k] = V3<T>(T(vv.nor[0]), T(vv.nor[1]), T(vv.nor[2]));
                b.uv[k] = V2<T>(T(vv.uv[0]), T(vv.uv[1]));
            }
            b.centroid = (b.p[0] + b.p[1] + b.p[2]) / T(3);            // :331-334
            surface_area += triangle_area(b.p[1] - b.p[0], b.p[2] - b.p[0]);
            world_bound.grow(b.p[0]); world_bound.grow(b.p[1]); world_bound.grow(b.p[2]);
        }
        // build_bvh (triangle_bvh.cpp:122-223): breadth-first, midpoint split on the
        // largest centroid extent while depth < 64, then sort + median.
        struct BNode { AABB<T> bound; int left = -1, right = -1; uint32_t start = 0, end = 0; int depth = 0; };
        std::vector<BNode> bn;
        struct Task { int *fill; int fill_node; bool fill_left; uint32_t start, end, depth; };
        std::queue<Task> tasks;
        int root = -1;
        tasks.push({&root, -1, false, 0, n, 0});
        const uint32_t leaf_threshold = 5, depth_threshold = 64;
        max_depth = 0;
        while (!tasks.empty()) {
            const Task task = tasks.front(); tasks.pop();
            AABB<T> all, cb;
            for (uint32_t i = task.start; i < task.end; ++i) {
                all.grow(tris[i].p[0]); all.grow(tris[i].p[1]); all.grow(tris[i].p[2]);
                cb.grow(tris[i].centroid);
            }
            const int idx = int(bn.size());
            bn.push_back(BNode());
            bn[idx].bound = all;
            bn[idx].depth = int(task.depth);
            max_depth = std::max(max_depth, int(task.depth));
            if (task.fill_node < 0) root = idx;
            else if (task.fill_left) bn[task.fill_node].left = idx;
            else bn[task.fill_node].right = idx;
            const uint32_t cnt = task.end - task.start;
            if (cnt <= leaf_threshold) {
                bn[idx].start = task.start; bn[idx].end = task.end;
                continue;
            }
            const V3<T> cd = cb.high - cb.low;
            const int axis = cd[0] > cd[1] ? (cd[0] > cd[2] ? 0 : 2) : (cd[1] > cd[2] ? 1 : 2);
            uint32_t mid;
            if (task.depth < depth_threshold) {
                const T split = T(0.5) * (cb.high[axis] + cb.low[axis]);
                mid = task.start;
                for (uint32_t i = task.start; i < task.end; ++i)
                    if (tris[i].centroid[axis] < split) std::swap(tris[i], tris[mid++]);
                if (mid == task.start || mid == task.end) mid = task.start + cnt / 2;
            } else {
                std::sort(tris.begin() + task.start, tris.begin() + task.end,
                          [axis](const BTri &L, const BTri &R) { return L.centroid[axis] < R.centroid[axis]; });
                mid = task.start + cnt / 2;
            }
            tasks.push({nullptr, idx, true, task.start, mid, task.depth + 1});
            tasks.push({nullptr, idx, false, mid, task.end, task.depth + 1});
        }
        // compact_bvh (triangle_bvh.cpp:225-304): depth-first, left child at idx+1.
        nodes.clear(); prims.clear(); infos.clear(); orig_id.clear();
        nodes.reserve(bn.size()); prims.resize(n); infos.resize(n); orig_id.resize(n);
        struct CT { int b; int64_t fill; };
        std::stack<CT> st;
        st.push({root, -1});
        uint32_t next_prim = 0;
        while (!st.empty()) {
            const CT ct = st.top(); st.pop();
            const BNode &b = bn[ct.b];
            const uint32_t my = uint32_t(nodes.size());
            if (ct.fill >= 0) nodes[size_t(ct.fill)].end_or_right = my;
            Node nd;
            for (int k = 0; k < 3; ++k) { nd.low[k] = b.bound.low[k]; nd.high[k] = b.bound.high[k]; }
            if (b.left >= 0 && b.right >= 0) {
                nd.start = 0xffffffffu; nd.end_or_right = 0;
                nodes.push_back(nd);
                st.push({b.right, int64_t(my)});
                st.push({b.left, -1});
            } else {
                const uint32_t s = next_prim, e = next_prim + (b.end - b.start);
                nd.start = s; nd.end_or_right = e;
                nodes.push_back(nd);
                for (uint32_t i = s, j = b.start; i < e; ++i, ++j) {
                    const BTri &tri = tris[j];
                    Prim &pr = prims[i];
                    PrimInfo &pi = infos[i];
                    orig_id[i] = tri.id;
                    pr.a = tri.p[0]; pr.b_a = tri.p[1] - tri.p[0]; pr.c_a = tri.p[2] - tri.p[0];
                    const V3<T> na = tri.nor[0].normalize(), nb = tri.nor[1].normalize(), nc = tri.nor[2].normalize();
                    pi.n_a = na; pi.n_b_a = nb - na; pi.n_c_a = nc - na;
                    pi.t_a = tri.uv[0]; pi.t_b_a = tri.uv[1] - tri.uv[0]; pi.t_c_a = tri.uv[2] - tri.uv[0];
                    pi.z = cross(pr.b_a, pr.c_a).normalize();
                    if (dot(na + nb + nc, pi.z) < 0) pi.z = -pi.z;
                    pi.x = dpdu_as_ex(pr.b_a, pr.c_a, pi.t_b_a, pi.t_c_a, pi.z);
                }
                next_prim = e;
            }
        }
        std::vector<T> areas(n);
        for (uint32_t i = 0; i < n; ++i) areas[i] = triangle_area(prims[i].b_a, prims[i].c_a);
        area_table.build(areas);
        // TriangleBVH::TriangleBVH world bound with flat-axis padding (:570-583)
        AABB<T> wb;
        for (auto &p : prims) { wb.grow(p.a); wb.grow(p.a + p.b_a); wb.grow(p.a + p.c_a); }
        for (int i = 0; i < 3; ++i)
            if (wb.low[i] >= wb.high[i]) wb.low[i] = wb.high[i] - T(0.1) * std::abs(wb.high[i]);
        world_bound = wb;
    }

    // UntransformedTriangleBVH::has_intersection (:363-401)
    bool any_hit(const Ray<T> &r, Counters *c) const {
        const T inv[3] = {1 / r.d.x, 1 / r.d.y, 1 / r.d.z};
        const T ori[3] = {r.o.x, r.o.y, r.o.z};
        T t;
        if (c) c->node_visits++;
        if (!node_hit(nodes[0], ori, inv, r.t_min, r.t_max, &t)) return false;
        uint32_t stack[256];
        int top = 0;
        stack[top++] = 0;
        while (top) {
            const uint32_t ni = stack[--top];
            const Node &nd = nodes[ni];
            if (nd.is_leaf()) {
                for (uint32_t i = nd.start; i < nd.end_or_right; ++i) {
                    if (c) c->tri_tests++;
                    if (mt_any(r, prims[i].a, prims[i].b_a, prims[i].c_a)) return true;
                }
            } else {
                if (c) c->node_visits += 2;
                if (node_hit(nodes[ni + 1], ori, inv, r.t_min, r.t_max, &t)) stack[top++] = ni + 1;
                if (node_hit(nodes[nd.end_or_right], ori, inv, r.t_min, r.t_max, &t)) stack[top++] = nd.end_or_right;
            }
        }
        return false;
    }

    // UntransformedTriangleBVH::closest_intersection (:403-488); fills slot index,
    // t and barycentrics; surface() reconstructs the rest.
    bool closest(Ray<T> r, T *t_out, T *u_out, T *v_out, uint32_t *slot_out, Counters *c) const {
        const T ori[3] = {r.o.x, r.o.y, r.o.z};
        const T inv[3] = {1 / r.d.x, 1 / r.d.y, 1 / r.d.z};
        T tmp;
        if (c) c->node_visits++;
        if (!node_hit(nodes[0], ori, inv, r.t_min, r.t_max, &tmp)) return false;
        uint32_t stack[256];
        int top = 0;
        stack[top++] = 0;
        T best_t = std::numeric_limits<T>::infinity(), bu = 0, bv = 0;
        uint32_t best = 0;
        while (top) {
            const uint32_t ni = stack[--top];
            const Node &nd = nodes[ni];
            if (nd.is_leaf()) {
                for (uint32_t i = nd.start; i < nd.end_or_right; ++i) {
                    if (c) c->tri_tests++;
                    T tt, uu, vv;
                    if (mt_closest(r, prims[i].a, prims[i].b_a, prims[i].c_a, &tt, &uu, &vv)) {
                        best_t = tt; bu = uu; bv = vv; best = i;
                        r.t_max = tt;
                    }
                }
            } else {
                if (c) c->node_visits += 2;
                T tl, tr;
                const bool al = node_hit(nodes[ni + 1], ori, inv, r.t_min, r.t_max, &tl);
                const bool ar = node_hit(nodes[nd.end_or_right], ori, inv, r.t_min, r.t_max, &tr);
                if (al && ar) {
                    if (tl < tr) { stack[top++] = nd.end_or_right; stack[top++] = ni + 1; }
                    else { stack[top++] = ni + 1; stack[top++] = nd.end_or_right; }
                } else if (al) stack[top++] = ni + 1;
                else if (ar) stack[top++] = nd.end_or_right;
            }
        }
        if (std::isinf(best_t)) return false;
        *t_out = best_t; *u_out = bu; *v_out = bv; *slot_out = best;
        return true;
    }

    void surface(uint32_t slot, T bu, T bv, SurfacePoint<T> *sp) const {  // :472-485
        const PrimInfo &pi = infos[slot];
        sp->geometry_coord = Coord<T>(pi.x, cross(pi.z, pi.x), pi.z);
        sp->uv = pi.t_a + pi.t_b_a * bu + pi.t_c_a * bv;
        const V3<T> user_z = pi.n_a + pi.n_b_a * bu + pi.n_c_a * bv;
        sp->user_coord = sp->geometry_coord.rotate_to_new_z(user_z);
    }

    // UntransformedTriangleBVH::sample (:495-518)
    SurfacePoint<T> sample(T *pdf, T su, T sv, T sw) const {
        const int32_t idx = area_table.sample(su);
        const Prim &pr = prims[idx];
        const V2<T> uv = uniform_on_triangle(sv, sw);
        SurfacePoint<T> sp;
        sp.pos = pr.a + pr.b_a * uv.x + pr.c_a * uv.y;
        surface(uint32_t(idx), uv.x, uv.y, &sp);
        *pdf = 1 / surface_area;
        return sp;
    }
};

template<class T> struct Geometry {
    int kind = AGX_GEOM_QUAD;
    bool double_sided = false;  // DoubleSidedGeometry around this geometry (double_sided.cpp)
    // quad (core/geometry/quad.cpp): world-space a, edges; triangle (triangle.cpp): local
    V3<T> a, b_a, c_a, d_a;
    V2<T> t_a, t_b_a, t_c_a, t_d_a;
    V3<T> x_abc, x_acd, z;
    T area = 1, sample_abc_prob = 0;
    // transformed geometry (transformed_geometry.h:31-82)
    Affine<T> l2w, w2l;
    T ratio = 1;
    T radius = 1, radius2 = 1, world_radius = 1;
    std::unique_ptr<TriangleBVH<T>> mesh;
    AABB<T> bound;

    Ray<T> to_local(const Ray<T> &r) const { return Ray<T>(w2l.point(r.o), w2l.vec(r.d), r.t_min, r.t_max); }
    Coord<T> coord_to_world(const Coord<T> &c) const {
        const V3<T> nz = w2l.tvec(c.z).normalize();
        const V3<T> mx = l2w.vec(c.x);
        const V3<T> xp = mx - nz * dot(mx, nz);
        const T l = xp.length();
        if (!(l > T(1e-6))) return Coord<T>::from_z(nz);
        const V3<T> nx = xp / l;
        return Coord<T>(nx, cross(nz, nx), nz);
    }
    void sp_to_world(SurfacePoint<T> *sp) const {
        sp->pos = l2w.point(sp->pos);
        sp->geometry_coord = coord_to_world(sp->geometry_coord);
        sp->user_coord = coord_to_world(sp->user_coord);
    }

    void init(const agx_entity_desc &e) {
        kind = e.geometry;
        double_sided = e.double_sided != 0;
        auto P = [&](int i) { return V3<T>(T(e.p[i][0]), T(e.p[i][1]), T(e.p[i][2])); };
        auto TT = [&](int i) { return V2<T>(T(e.t[i][0]), T(e.t[i][1])); };
        l2w = Affine<T>::from(e.local_to_world);
        w2l = Affine<T>::from(e.world_to_local);
        ratio = l2w.vec(V3<T>(1, 0, 0)).length();
        bound = AABB<T>();
        if (kind == AGX_GEOM_QUAD) {                            // quad.cpp:138-164, 60-73
            a = P(0); b_a = P(1) - a; c_a = P(2) - a; d_a = P(3) - a;
            t_a = TT(0); t_b_a = TT(1) - t_a; t_c_a = TT(2) - t_a; t_d_a = TT(3) - t_a;
            z = cross(b_a, c_a).normalize();
            x_abc = dpdu_as_ex(b_a, c_a, t_b_a, t_c_a, z);
            x_acd = dpdu_as_ex(c_a, d_a, t_c_a, t_d_a, z);
            const T a1 = triangle_area(b_a, c_a), a2 = triangle_area(c_a, d_a);
            area = a1 + a2;
            sample_abc_prob = a1 / area;
            bound.grow(a); bound.grow(a + b_a); bound.grow(a + c_a); bound.grow(a + d_a);
            pad_flat();
        } else if (kind == AGX_GEOM_TRIANGLE) {                 // triangle.cpp:107-125, 53-65
            a = P(0); b_a = P(1) - a; c_a = P(2) - a;
            t_a = TT(0); t_b_a = TT(1) - t_a; t_c_a = TT(2) - t_a;
            z = cross(b_a, c_a).normalize();
            x_abc = dpdu_as_ex(b_a, c_a, t_b_a, t_c_a, z);
            area = triangle_area(l2w.vec(b_a), l2w.vec(c_a));
            bound.grow(l2w.point(a)); bound.grow(l2w.point(a + b_a)); bound.grow(l2w.point(a + c_a));
            pad_flat();
        } else if (kind == AGX_GEOM_SPHERE) {                   // sphere.cpp:14-24, 60-67
            radius = T(e.radius);
            if (!(radius > 0)) throw std::runtime_error("invalid sphere radius");
            world_radius = ratio * radius;
            area = 4 * PI<T> * world_radius * world_radius;
            const V3<T> c = l2w.point(V3<T>(0));
            bound.low = c - V3<T>(world_radius + T(3e-4));
            bound.high = c + V3<T>(world_radius + T(3e-4));
        } else if (kind == AGX_GEOM_DISK) {                     // disk.cpp:33-47, 85-93
            radius = T(e.radius); radius2 = radius * radius;
            if (!(radius > 0)) throw std::runtime_error("invalid disk radius");
            area = ratio * ratio * PI<T> * radius2;
            const T vs = std::min(T(0.1) * radius, T(0.01));
            const V3<T> lo(-radius, -radius, -vs), hi(radius, radius, vs);
            for (int i = 0; i < 8; ++i)
                bound.grow(l2w.point(V3<T>(i & 1 ? hi.x : lo.x, i & 2 ? hi.y : lo.y, i & 4 ? hi.z : lo.z)));
        } else if (kind == AGX_GEOM_TRIANGLE_BVH) {
            if (!e.tris || !e.tri_count) throw std::runtime_error("empty triangle mesh");
            mesh.reset(new TriangleBVH<T>());
            mesh->build(e.tris, e.tri_count);
            area = mesh->surface_area;
            bound = mesh->world_bound;
        } else {
            throw std::runtime_error("unknown geometry kind");
        }
    }
    void pad_flat() {
        for (int i = 0; i < 3; ++i)
            if (bound.low[i] >= bound.high[i]) bound.low[i] = bound.high[i] - T(0.1) * std::abs(bound.high[i]);
    }

    bool any_hit(const Ray<T> &r, Counters *c) const {
        switch (kind) {
        case AGX_GEOM_QUAD:                                      // quad.cpp:24-28
            if (c) c->tri_tests += 1;
            if (mt_any(r, a, b_a, c_a)) return true;
            if (c) c->tri_tests += 1;
            return mt_any(r, a, c_a, d_a);
        case AGX_GEOM_TRIANGLE: {
            if (c) c->tri_tests += 1;
            return mt_any(to_local(r), a, b_a, c_a);
        }
        case AGX_GEOM_SPHERE: return sphere_any(to_local(r));
        case AGX_GEOM_DISK: {                                    // disk.cpp:49-61
            const Ray<T> lr = to_local(r);
            const T t = -lr.o.z / lr.d.z;
            if (std::isinf(t) || !lr.between(t)) return false;
            const V3<T> p = lr.at(t);
            return p.x * p.x + p.y * p.y <= radius2;
        }
        default: return mesh->any_hit(r, c);
        }
    }

    // sphere_aux.h:10-16, 50-64
    void sphere_abc(const Ray<T> &r, T *A, T *B, T *C) const {
        *A = r.d.length_square(); *B = 2 * dot(r.d, r.o); *C = r.o.length_square() - radius * radius;
    }
    bool sphere_any(const Ray<T> &r) const {
        T A, B, C; sphere_abc(r, &A, &B, &C);
        T delta = B * B - 4 * A * C;
        if (delta < 0) return false;
        delta = std::sqrt(delta);
        const T inv2A = T(0.5) / A;
        const T t0 = (-B - delta) * inv2A, t1 = (-B + delta) * inv2A;
        return r.between(t0) || r.between(t1);
    }
    // sphere_aux.h:18-48
    void sphere_local_uv_coord(const V3<T> &p, SurfacePoint<T> *sp, T eps) const {
        T phi = (!p.x && !p.y) ? T(0) : std::atan2(p.y, p.x);
        if (phi < 0) phi += 2 * PI<T>;
        const T sin_theta = std::min(T(1), std::max(T(-1), p.z / radius));
        const T theta = std::asin(sin_theta);
        sp->uv = V2<T>(T(0.5) / PI<T> * phi, theta / PI<T> + T(0.5));
        V3<T> ex, ey, ez = p.normalize();
        if (std::abs(ez.z - 1) < eps) { ex = V3<T>(1, 0, 0); ey = V3<T>(0, 1, 0); ez = V3<T>(0, 0, 1); }
        else if (std::abs(ez.z + 1) < eps) { ex = V3<T>(0, 1, 0); ey = V3<T>(1, 0, 0); ez = V3<T>(0, 0, -1); }
        else { ex = cross(V3<T>(0, 0, 1), ez); ey = cross(ez, ex); }
        sp->geometry_coord = Coord<T>(ex, ey, ez);
        sp->user_coord = sp->geometry_coord;
    }
    // disk.cpp:11-28
    void disk_local_sp(const V3<T> &p, T rad, SurfacePoint<T> *sp) const {
        T phi = (!p.y && !p.x) ? T(0) : std::atan2(p.y, p.x);
        if (phi < 0) phi += 2 * PI<T>;
        sp->pos = p;
        sp->uv = V2<T>(phi / (2 * PI<T>), rad / radius);
        const V3<T> cx = V3<T>(p.x, p.y, 0).normalize();
        const V3<T> cz(0, 0, 1);
        sp->geometry_coord = Coord<T>(cx, cross(cz, cx), cz);
        sp->user_coord = sp->geometry_coord;
    }

    // DoubleSidedGeometry::closest_intersection (double_sided.cpp:26-38): a hit on
    // the back side turns both frames around
    bool closest(const Ray<T> &r, GeometryHit<T> *h, T eps, Counters *c) const {
        if (!closest_one_sided(r, h, eps, c)) return false;
        if (double_sided && dot(h->geometry_coord.z, h->wr) < 0) {
            h->geometry_coord = flip(h->geometry_coord);
            h->user_coord = flip(h->user_coord);
        }
        return true;
    }
    static Coord<T> flip(const Coord<T> &f) { return Coord<T>(-f.x, -f.y, -f.z); }

    bool closest_one_sided(const Ray<T> &r, GeometryHit<T> *h, T eps, Counters *c) const {
        switch (kind) {
        case AGX_GEOM_QUAD: {                                   // quad.cpp:30-58
            T t, u, v;
            if (c) c->tri_tests += 1;
            if (mt_closest(r, a, b_a, c_a, &t, &u, &v)) {
                h->pos = r.at(t);
                h->geometry_coord = Coord<T>(x_abc, cross(z, x_abc), z);
                h->uv = t_a + t_b_a * u + t_c_a * v;
                h->user_coord = h->geometry_coord;
                h->wr = -r.d; h->t = t; h->prim = 0; h->bu = u; h->bv = v;
                return true;
            }
            if (c) c->tri_tests += 1;
            if (mt_closest(r, a, c_a, d_a, &t, &u, &v)) {
                h->pos = r.at(t);
                h->geometry_coord = Coord<T>(x_acd, cross(z, x_acd), z);
                h->uv = t_a + t_c_a * u + t_d_a * v;
                h->user_coord = h->geometry_coord;
                h->wr = -r.d; h->t = t; h->prim = 1; h->bu = u; h->bv = v;
                return true;
            }
            return false;
        }
        case AGX_GEOM_TRIANGLE: {                               // triangle.cpp:31-51
            const Ray<T> lr = to_local(r);
            T t, u, v;
            if (c) c->tri_tests += 1;
            if (!mt_closest(lr, a, b_a, c_a, &t, &u, &v)) return false;
            h->pos = lr.at(t);
            h->geometry_coord = Coord<T>(x_abc, cross(z, x_abc), z);
            h->uv = t_a + t_b_a * u + t_c_a * v;
            h->user_coord = h->geometry_coord;
            h->t = t; h->prim = 0; h->bu = u; h->bv = v;
            sp_to_world(h);
            h->wr = l2w.vec(-lr.d);
            return true;
        }
        case AGX_GEOM_SPHERE: {                                 // sphere.cpp:33-58
            const Ray<T> lr = to_local(r);
            T A, B, C; sphere_abc(lr, &A, &B, &C);
            T delta = B * B - 4 * A * C;
            if (delta < 0) return false;
            delta = std::sqrt(delta);
            const T inv2A = T(0.5) / A;
            const T t0 = (-B - delta) * inv2A, t1 = (-B + delta) * inv2A;
            if (lr.t_max < t0 || t1 < lr.t_min) return false;
            const bool not_t0 = t0 < lr.t_min;
            if (not_t0 && t1 > lr.t_max) return false;
            const T t = not_t0 ? t1 : t0;
            const V3<T> p = lr.at(t);
            h->pos = p;
            sphere_local_uv_coord(p, h, eps);
            h->t = t; h->prim = 0; h->bu = 0; h->bv = 0;
            sp_to_world(h);
            h->wr = l2w.vec(-lr.d);
            return true;
        }
        case AGX_GEOM_DISK: {                                   // disk.cpp:63-83
            const Ray<T> lr = to_local(r);
            const T tv = -lr.o.z / lr.d.z;
            if (std::isinf(tv) || !lr.between(tv)) return false;
            const V3<T> p = lr.at(tv);
            const T rad = std::sqrt(p.x * p.x + p.y * p.y);
            if (rad > radius) return false;
            disk_local_sp(p, rad, h);
            h->t = tv; h->prim = 0; h->bu = 0; h->bv = 0;
            sp_to_world(h);
            h->wr = l2w.vec(-lr.d);
            return true;
        }
        default: {
            T t, u, v; uint32_t slot;
            if (!mesh->closest(r, &t, &u, &v, &slot, c)) return false;
            h->pos = r.at(t);
            mesh->surface(slot, u, v, h);
            h->t = t; h->wr = -r.d; h->prim = int32_t(mesh->orig_id[slot]); h->bu = u; h->bv = v;
            return true;
        }
        }
    }

    // DoubleSidedGeometry::sample(ref, pdf, sam) (double_sided.cpp:66-87): the
    // first number picks the side and is rescaled for the wrapped geometry;
    // the pdf halves and a back-side sample turns its frames around
    SurfacePoint<T> sample(const V3<T> &ref, T *pdf, T su, T sv, T sw, T eps) const {
        if (!double_sided) return sample_one_sided(ref, pdf, su, sv, sw, eps);
        const bool back = su < T(0.5);
        SurfacePoint<T> sp = sample_one_sided(ref, pdf, back ? 2 * su : 2 * (su - T(0.5)), sv, sw, eps);
        *pdf = *pdf * T(0.5);
        if (back) {
            sp.geometry_coord = flip(sp.geometry_coord);
            sp.user_coord = flip(sp.user_coord);
        }
        return sp;
    }

    // Geometry::sample(ref, pdf, sam) per kind; pdf w.r.t. area
    SurfacePoint<T> sample_one_sided(const V3<T> &ref, T *pdf, T su, T sv, T sw, T eps) const {
        SurfacePoint<T> sp;
        switch (kind) {
        case AGX_GEOM_QUAD: {                                   // quad.cpp:80-105
            const V2<T> bi = uniform_on_triangle(su, sv);
            if (sw < sample_abc_prob) {
                sp.pos = a + b_a * bi.x + c_a * bi.y;
                sp.geometry_coord = Coord<T>(x_abc, cross(z, x_abc), z);
                sp.uv = t_a + t_b_a * bi.x + t_c_a * bi.y;
            } else {
                sp.pos = a + c_a * bi.x + d_a * bi.y;
                sp.geometry_coord = Coord<T>(x_acd, cross(z, x_acd), z);
                sp.uv = t_a + t_c_a * bi.x + t_d_a * bi.y;
            }
            sp.user_coord = sp.geometry_coord;
            *pdf = 1 / area;
            return sp;
        }
        case AGX_GEOM_TRIANGLE: {                               // triangle.cpp:72-87
            const V2<T> bi = uniform_on_triangle(su, sv);
            sp.pos = a + b_a * bi.x + c_a * bi.y;
            sp.geometry_coord = Coord<T>(x_abc, cross(z, x_abc), z);
            sp.uv = t_a + t_b_a * bi.x + t_c_a * bi.y;
            sp.user_coord = sp.geometry_coord;
            sp_to_world(&sp);
            *pdf = 1 / area;
            return sp;
        }
        case AGX_GEOM_SPHERE: {                                 // sphere.cpp:93-123
            const V3<T> lref = w2l.point(ref);
            const T d = lref.length();
            if (d <= radius) {
                const V3<T> up = uniform_on_sphere(su, sv);
                *pdf = uniform_on_sphere_pdf<T> / (world_radius * world_radius);
                const V3<T> p = up * radius;
                sp.pos = p;
                sphere_local_uv_coord(p, &sp, eps);
                sp_to_world(&sp);
                return sp;
            }
            const T cos_t = std::min(radius / d, T(1));
            const V3<T> dir = uniform_on_cone(cos_t, su, sv);
            const T lpdf = uniform_on_cone_pdf(cos_t);
            const V3<T> p = Coord<T>::from_z(lref).local_to_global(dir).normalize() * radius;
            sp.pos = p;
            sphere_local_uv_coord(p, &sp, eps);
            sp_to_world(&sp);
            *pdf = lpdf / (world_radius * world_radius);
            return sp;
        }
        case AGX_GEOM_DISK: {                                   // disk.cpp:100-110
            const V2<T> dp = uniform_on_unit_disk(su, sv);
            const V2<T> p2 = dp * radius;
            *pdf = 1 / area;
            disk_local_sp(V3<T>(p2.x, p2.y, 0), std::sqrt(p2.x * p2.x + p2.y * p2.y), &sp);
            sp_to_world(&sp);
            return sp;
        }
        default:
            return mesh->sample(pdf, su, sv, sw);
        }
    }

    // Geometry::pdf(ref, pos), area measure (halved for a double-sided geometry,
    // double_sided.cpp:89-99)
    T pdf(const V3<T> &ref) const {
        const T p = pdf_one_sided(ref);
        return double_sided ? p * T(0.5) : p;
    }
    T pdf_one_sided(const V3<T> &ref) const {
        if (kind == AGX_GEOM_SPHERE) {                          // sphere.cpp:131-143
            const V3<T> lref = w2l.point(ref);
            const T d = lref.length();
            if (d <= radius) return 1 / area;
            const T cos_t = std::min(radius / d, T(1));
            return uniform_on_cone_pdf(cos_t) / (world_radius * world_radius);
        }
        return 1 / area;
    }
};

// ------------------------------------------------------------------------------------
// BSDFs  (core/bsdf.h:113-202, material/*)
// ------------------------------------------------------------------------------------

template<class T> struct BSDFSample {
    V3<T> dir; V3<T> f; T pdf = 0; bool is_delta = false;
};

template<class T> struct BSDF {
    int kind = AGX_MAT_IDEAL_DIFFUSE;
    Coord<T> geo, shd;
    V3<T> albedo;         // AggregateBSDF albedo_ / mirror rc / glass mean colour
    V3<T> coef;           // diffuse albedo/pi
    // metal GGX
    V3<T> color, eta2, k2;   // ConductorPoint eta_2_, eta_k_2_ (eta_out = 1)
    T ax = 1, ay = 1;
    // glass
    V3<T> c_refl, c_refr; T ior = 1.5f;
    // mirror
    V3<T> m_eta2, m_k2;      // eta_in = 1, eta_out = ior
    // disney (material/disney.cpp:18-457): C = albedo
    V3<T> d_tint, d_spec_scale;
    T d_metallic = 0, d_rough = 0, d_spec_tint = 0, d_sheen = 0, d_sheen_tint = 0, d_clearcoat = 0;
    T d_trans = 0, d_ior = T(1.5), d_tax = 1, d_tay = 1, d_ccr = T(0.01);
    T w_diffuse = T(0.25), w_spec = T(0.25), w_cc = T(0.25), w_trans = T(0.25);
    T eps_ = T(3e-4);        // EPS() of the |cos| < EPS tests in eval / pdf

    // ---- LocalBSDF black-fringe helpers (bsdf.h:122-187)
    bool fringe(const V3<T> &w) const { return shd.in_positive_z_hemisphere(w) != geo.in_positive_z_hemisphere(w); }
    V3<T> albedo_v() const { return albedo; }

    // conductor Fresnel (fresnel_point.h:44-64)
    static V3<T> conductor(T cos_i, const V3<T> &e2, const V3<T> &ek2) {
        if (cos_i <= 0) return V3<T>(0);
        const T cos2 = cos_i * cos_i;
        const T sin2 = std::max(T(0), 1 - cos2);
        V3<T> out;
        for (int ch = 0; ch < 3; ++ch) {
            const T t0 = e2[ch] - ek2[ch] - sin2;
            const T a2b2 = std::sqrt(t0 * t0 + T(4) * e2[ch] * ek2[ch]);
            const T t1 = a2b2 + cos2;
            const T aa = std::sqrt(T(0.5) * (a2b2 + t0));
            const T t2 = 2 * cos_i * aa;
            const T rs = (t1 - t2) / (t1 + t2);
            const T t3 = cos2 * a2b2 + sin2 * sin2;
            const T t4 = t2 * sin2;
            const T rp = rs * (t3 - t4) / (t3 + t4);
            out[ch] = T(0.5) * (rp + rs);
        }
        return out;
    }
    // refl_aux::dielectric_fresnel (utility/reflection.h:48-72)
    static T dielectric(T eta_i, T eta_o, T cos_i) {
        if (cos_i < 0) { std::swap(eta_i, eta_o); cos_i = -cos_i; }
        const T sin_i = std::sqrt(std::max(T(0), 1 - cos_i * cos_i));
        const T sin_t = eta_o / eta_i * sin_i;
        if (sin_t >= 1) return 1;
        const T cos_t = std::sqrt(std::max(T(0), 1 - sin_t * sin_t));
        const T para = (eta_i * cos_i - eta_o * cos_t) / (eta_i * cos_i + eta_o * cos_t);
        const T perp = (eta_o * cos_i - eta_i * cos_t) / (eta_o * cos_i + eta_i * cos_t);
        return T(0.5) * (para * para + perp * perp);
    }
    static T normal_corr(const V3<T> &g, const V3<T> &s, const V3<T> &wi, T eps) {  // common.h:239-244
        const T dem = std::abs(cos_between(g, wi));
        return dem < eps ? T(1) : std::abs(cos_between(s, wi) / dem);
    }
    static T phi_of(const V3<T> &w) {  // common.h:220-226
        if (!w.y && !w.x) return 0;
        const T r = std::atan2(w.y, w.x);
        return r < 0 ? r + 2 * PI<T> : r;
    }
    static T tan_theta(const V3<T> &w) { const T t = 1 - w.z * w.z; if (t <= 0) return 0; return std::sqrt(t) / w.z; }

    // GGX component eval / pdf (microfacet_refl_comp.cpp:8-65, utility/microfacet.cpp:44-61)
    void ggx_eval_pdf(const V3<T> &lwi, const V3<T> &lwo, V3<T> *f, T *pdf) const {
        if (lwi.z <= 0 || lwo.z <= 0) { if (f) *f = V3<T>(0); if (pdf) *pdf = 0; return; }
        const T cti = lwi.z, cto = lwo.z;
        const V3<T> lwh = (lwi + lwo).normalize();
        const T ctd = dot(lwi, lwh);
        const T phi_h = phi_of(lwh);
        const T sph = std::sin(phi_h), cph = std::cos(phi_h);
        const T cth = lwh.z;
        const T sth = std::sqrt(std::max(T(0), 1 - cth * cth));
        const T A = (cph / ax) * (cph / ax) + (sph / ay) * (sph / ay);
        const T RD = sth * sth * A + cth * cth;
        const T D = 1 / (PI<T> * ax * ay * (RD * RD));
        const T phi_i = phi_of(lwi), phi_o = phi_of(lwo);
        const T spi = std::sin(phi_i), cpi = std::cos(phi_i);
        const T spo = std::sin(phi_o), cpo = std::cos(phi_o);
        auto smith = [&](T cp, T sp, T tt) {
            const T t = (ax * cp) * (ax * cp) + (ay * sp) * (ay * sp);
            const T sv = 1 + t * (tt * tt);
            const T lambda = -T(0.5) + T(0.5) * std::sqrt(sv);
            return 1 / (1 + lambda);
        };
        const T Gi = smith(cpi, spi, tan_theta(lwi));
        const T Go = smith(cpo, spo, tan_theta(lwo));
        if (f) {
            const V3<T> F = conductor(ctd, eta2, k2) * color;
            const T G = Gi * Go;
            *f = F * D * G / std::abs(4 * cti * cto);
        }
        if (pdf) *pdf = Go * D / (4 * lwo.z);
    }
    // microfacet::sample_anisotropic_gtr2_vnor (microfacet.cpp:104-128)
    V3<T> ggx_vndf(const V3<T> &ve, T su, T sv, T eps) const {
        const V3<T> vh = V3<T>(ax * ve.x, ay * ve.y, ve.z).normalize();
        const T lensq = vh.x * vh.x + vh.y * vh.y;
        const V3<T> t1 = lensq > eps ? V3<T>(-vh.y, vh.x, 0) / std::sqrt(lensq) : V3<T>(1, 0, 0);
        const V3<T> t2 = cross(vh, t1);
        const T r = std::sqrt(su);
        const T phi = 2 * PI<T> * sv;
        const T t_1 = r * std::cos(phi);
        const T _t_2 = r * std::sin(phi);
        const T s = T(0.5) * (1 + vh.z);
        const T t_2 = (1 - s) * std::sqrt(1 - t_1 * t_1) + s * _t_2;
        const V3<T> nh = t1 * t_1 + t2 * t_2 + vh * std::sqrt(std::max(T(0), 1 - t_1 * t_1 - t_2 * t_2));
        return V3<T>(ax * nh.x, ay * nh.y, std::max(T(0), nh.z)).normalize();
    }

    // ---- microfacet distributions (material/utility/microfacet.cpp:13-105)
    static T sqr(T x) { return x * x; }
    static T aniso_gtr2(T sph, T cph, T sth, T cth, T ax_, T ay_) {
        const T A = sqr(cph / ax_) + sqr(sph / ay_);
        const T RD = sqr(sth) * A + sqr(cth);
        return 1 / (PI<T> * ax_ * ay_ * sqr(RD));
    }
    static T smith_aniso(T cp, T sp, T ax_, T ay_, T tt) {
        const T t = sqr(ax_ * cp) + sqr(ay_ * sp);
        const T sv = 1 + t * sqr(tt);
        const T lambda = -T(0.5) + T(0.5) * std::sqrt(sv);
        return 1 / (1 + lambda);
    }
    static T gtr1(T sth, T cth, T a) {
        const T U = sqr(a) - 1;
        const T LD = 2 * PI<T> * std::log(a);
        const T RD = sqr(a * cth) + sqr(sth);
        return U / (LD * RD);
    }
    static T smith_gtr2(T tt, T a) {
        if (!tt) return 1;
        const T root = a * tt;
        return 2 / (1 + std::sqrt(1 + root * root));
    }
    static V3<T> sample_aniso_gtr2(T ax_, T ay_, T u, T v) {
        T sph = ay_ * std::sin(2 * PI<T> * u);
        T cph = ax_ * std::cos(2 * PI<T> * u);
        const T nor = 1 / std::sqrt(sqr(sph) + sqr(cph));
        sph *= nor;
        cph *= nor;
        const T A = sqr(cph / ax_) + sqr(sph / ay_);
        const T cth = std::sqrt(A * (1 - v) / ((1 - A) * v + A));
        const T sth = std::sqrt(std::max(T(0), 1 - sqr(cth)));
        return V3<T>(sth * cph, sth * sph, cth).normalize();
    }
    static V3<T> sample_gtr1(T a, T u, T v) {
        const T phi = 2 * PI<T> * u;
        const T ct = std::sqrt((std::pow(a, 2 - 2 * v) - 1) / (sqr(a) - 1));
        const T st = std::sqrt(std::max(T(0), 1 - ct * ct));
        return V3<T>(st * std::cos(phi), st * std::sin(phi), ct).normalize();
    }
    // refl_aux::refract (utility/reflection.h:31-40); false = total internal reflection
    static bool refract(const V3<T> &w, const V3<T> &n, T eta, V3<T> *out) {
        const T ci = dot(w, n);
        const T s2t = sqr(eta) * (1 - sqr(ci));
        const T c2t = 1 - s2t;
        if (c2t <= 0) return false;
        const T ct = std::sqrt(c2t);
        *out = (n * (eta * ci - ct) - w * eta).normalize();
        return true;
    }
    static T one_minus_5(T x) { const T t = 1 - x; const T t2 = t * t; return t2 * t2 * t; }
    static T cos_2_sin(T c) { return std::sqrt(std::max(T(0), 1 - c * c)); }
    static bool zero(const V3<T> &v) { return !v.x && !v.y && !v.z; }
    // math::mix (agz-utils, parity unpinned): GLSL mix, a (1 - t) + b t
    static V3<T> mixv(const V3<T> &a, const V3<T> &b, T t) { return a * (1 - t) + b * t; }

    // ---- Disney BSDF (material/disney.cpp:77-395)
    V3<T> d_diffuse(T ci, T co, T cd) const {
        const V3<T> f_lambert = albedo / PI<T>;
        const T FL = one_minus_5(ci), FV = one_minus_5(co);
        const T RR = 2 * d_rough * cd * cd;
        const V3<T> retro = albedo / PI<T> * RR * (FL + FV + FL * FV * (RR - 1));
        return f_lambert * (1 - T(0.5) * FL) * (1 - T(0.5) * FV) + retro;
    }
    V3<T> d_sheen_f(T cd) const { return mixv(V3<T>(1), d_tint, d_sheen_tint) * (4 * d_sheen) * one_minus_5(cd); }
    V3<T> d_clearcoat_f(T ci, T co, T tti, T tto, T sth, T cth, T cd) const {
        const T D = gtr1(sth, cth, d_ccr);
        const T F = T(0.04) + (1 - T(0.04)) * one_minus_5(cd);
        const T G = smith_gtr2(tti, T(0.25)) * smith_gtr2(tto, T(0.25));
        return V3<T>(d_clearcoat * D * F * G / std::abs(4 * ci * co));
    }
    void d_angles(const V3<T> &lwi, const V3<T> &lwo, T tax, T tay, const V3<T> &lwh, T *D, T *G) const {
        const T phi_h = phi_of(lwh);
        const T sph = std::sin(phi_h), cph = std::cos(phi_h);
        const T cth = lwh.z, sth = cos_2_sin(cth);
        *D = aniso_gtr2(sph, cph, sth, cth, tax, tay);
        const T phi_i = phi_of(lwi), phi_o = phi_of(lwo);
        const T spi = std::sin(phi_i), cpi = std::cos(phi_i);
        const T spo = std::sin(phi_o), cpo = std::cos(phi_o);
        *G = smith_aniso(cpi, spi, tax, tay, tan_theta(lwi)) * smith_aniso(cpo, spo, tax, tay, tan_theta(lwo));
    }
    V3<T> d_trans_f(const V3<T> &lwi, const V3<T> &lwo) const {
        const T ci = lwi.z, co = lwo.z;
        const T eta = co > 0 ? d_ior : 1 / d_ior;
        V3<T> lwh = (lwo + lwi * eta).normalize();
        if (lwh.z < 0) lwh = -lwh;
        const T cd = dot(lwo, lwh);
        const T F = dielectric(d_ior, 1, cd);
        T D, G;
        d_angles(lwi, lwo, d_tax, d_tay, lwh, &D, &G);
        const T sdem = cd + eta * dot(lwi, lwh);
        const T corr = 1 / eta;  // TransMode::Radiance
        const T val = (1 - F) * D * G * eta * eta * dot(lwi, lwh) * dot(lwo, lwh) * corr * corr /
                      (ci * co * sdem * sdem);
        const T tf = co > 0 ? d_trans : 1;
        return vsqrt(albedo) * ((1 - d_metallic) * tf) * std::abs(val);
    }
    V3<T> d_inner_refl_f(const V3<T> &lwi, const V3<T> &lwo) const {
        const V3<T> lwh = -(lwi + lwo).normalize();
        const T cd = dot(lwo, lwh);
        const T F = dielectric(d_ior, 1, cd);
        T D, G;
        d_angles(lwi, lwo, d_tax, d_tay, lwh, &D, &G);
        return albedo * d_trans * std::abs(F * D * G / (4 * lwi.z * lwo.z));
    }
    V3<T> d_specular_f(const V3<T> &lwi, const V3<T> &lwo) const {
        const T ci = lwi.z, co = lwo.z;
        const V3<T> lwh = (lwi + lwo).normalize();
        const T cd = dot(lwi, lwh);
        const V3<T> cspec = mixv(mixv(V3<T>(1), d_tint, d_spec_tint), albedo, d_metallic);
        const V3<T> diel = cspec * dielectric(d_ior, 1, cd);
        const V3<T> cond = cspec + (V3<T>(1) - cspec) * one_minus_5(cd);
        const V3<T> F = mixv(d_spec_scale * diel, cond, d_metallic);
        T D, G;
        d_angles(lwi, lwo, ax, ay, lwh, &D, &G);
        return F * D * G / std::abs(4 * ci * co);
    }
    V3<T> disney_eval(const V3<T> &wi, const V3<T> &wo) const {
        if (fringe(wi) || fringe(wo)) {
            if (!geo.in_positive_z_hemisphere(wi) || !geo.in_positive_z_hemisphere(wo)) return V3<T>(0);
            return albedo / PI<T>;
        }
        const V3<T> lwi = shd.global_to_local(wi).normalize();
        const V3<T> lwo = shd.global_to_local(wo).normalize();
        if (std::abs(lwi.z) < eps_ || std::abs(lwo.z) < eps_) return V3<T>(0);
        if (lwi.z * lwo.z < 0) {
            if (!d_trans) return V3<T>(0);
            return d_trans_f(lwi, lwo) * normal_corr(geo.z, shd.z, wi, eps_);
        }
        if (lwi.z < 0 && lwo.z < 0) {
            if (!d_trans) return V3<T>(0);
            return d_inner_refl_f(lwi, lwo) * normal_corr(geo.z, shd.z, wi, eps_);
        }
        if (lwi.z <= 0 || lwo.z <= 0) return V3<T>(0);
        const T ci = lwi.z, co = lwo.z;
        const V3<T> lwh = (lwi + lwo).normalize();
        const T cd = dot(lwi, lwh);
        V3<T> diffuse, sheen;
        if (d_metallic < 1) {
            diffuse = d_diffuse(ci, co, cd);
            if (d_sheen > 0) sheen = d_sheen_f(cd);
        }
        const V3<T> specular = d_specular_f(lwi, lwo);
        V3<T> clearcoat;
        if (d_clearcoat > 0) {
            const T cth = lwh.z;
            clearcoat = d_clearcoat_f(ci, co, tan_theta(lwi), tan_theta(lwo), cos_2_sin(cth), cth, cd);
        }
        const V3<T> value = (diffuse + sheen) * ((1 - d_metallic) * (1 - d_trans)) + specular + clearcoat;
        return value * normal_corr(geo.z, shd.z, wi, eps_);
    }
    T d_pdf_trans(const V3<T> &lwi, const V3<T> &lwo) const {
        const T eta = lwo.z > 0 ? d_ior : 1 / d_ior;
        V3<T> lwh = (lwo + lwi * eta).normalize();
        if (lwh.z < 0) lwh = -lwh;
        if (((lwo.z > 0) != (dot(lwh, lwo) > 0)) || ((lwi.z > 0) != (dot(lwh, lwi) > 0))) return 0;
        const T sdem = dot(lwo, lwh) + eta * dot(lwi, lwh);
        const T dwh_to_dwi = eta * eta * dot(lwi, lwh) / (sdem * sdem);
        const T phi_h = phi_of(lwh);
        const T cth = lwh.z;
        const T D = aniso_gtr2(std::sin(phi_h), std::cos(phi_h), cos_2_sin(cth), cth, d_tax, d_tay);
        return std::abs(cth * D * dwh_to_dwi);
    }
    T d_pdf_inner(const V3<T> &lwi, const V3<T> &lwo) const {
        const V3<T> lwh = -(lwi + lwo).normalize();
        const T phi_h = phi_of(lwh);
        const T cth = lwh.z;
        const T cd = dot(lwi, lwh);
        const T D = aniso_gtr2(std::sin(phi_h), std::cos(phi_h), cos_2_sin(cth), cth, d_tax, d_tay);
        return std::abs(cth * D / (4 * cd));
    }
    T disney_pdf(const V3<T> &wi, const V3<T> &wo) const {
        if (fringe(wi) || fringe(wo)) {
            if (geo.in_positive_z_hemisphere(wi) != geo.in_positive_z_hemisphere(wo)) return 0;
            return zweighted_on_hemisphere_pdf(geo.global_to_local(wi).normalize().z);
        }
        const V3<T> lwi = shd.global_to_local(wi).normalize();
        const V3<T> lwo = shd.global_to_local(wo).normalize();
        if (std::abs(lwi.z) < eps_ || std::abs(lwo.z) < eps_) return 0;
        if (lwo.z < 0) {
            if (!d_trans) return 0;
            const T mF = std::min(T(0.9), std::max(T(0.1), dielectric(d_ior, 1, lwo.z)));
            if (lwi.z > 0) return (1 - mF) * d_pdf_trans(lwi, lwo);
            if (!mF) return 0;
            return mF * d_pdf_inner(lwi, lwo);
        }
        if (lwi.z < 0) return w_trans * d_pdf_trans(lwi, lwo);
        const T diffuse = zweighted_on_hemisphere_pdf(lwi.z);
        // pdf_specular_clearcoat (disney.cpp:317-348): sin(phi_o) taken as |sin|, as Atrc does
        const V3<T> lwh = (lwi + lwo).normalize();
        const T phi_h = phi_of(lwh);
        const T sph = std::sin(phi_h), cph = std::cos(phi_h);
        const T cth = lwh.z, sth = cos_2_sin(cth);
        const T cd = dot(lwi, lwh);
        const T cpo = std::cos(phi_of(lwo));
        const T spo = cos_2_sin(cpo);
        const T spec_D = aniso_gtr2(sph, cph, sth, cth, ax, ay);
        const T pdf_spec = smith_aniso(cpo, spo, ax, ay, tan_theta(lwo)) * spec_D / (4 * lwo.z);
        const T cc_D = gtr1(sth, cth, d_ccr);
        const T pdf_cc = cth * cc_D / (4 * cd);
        return w_diffuse * diffuse + w_spec * pdf_spec + w_cc * pdf_cc;
    }
    V3<T> d_sample_trans(const V3<T> &lwo, T u, T v) const {
        const V3<T> lwh = sample_aniso_gtr2(d_tax, d_tay, u, v);
        if (lwh.z <= 0) return V3<T>(0);
        if ((lwo.z > 0) != (dot(lwh, lwo) > 0)) return V3<T>(0);
        const T eta = lwo.z > 0 ? 1 / d_ior : d_ior;
        const V3<T> owh = dot(lwh, lwo) > 0 ? lwh : -lwh;
        V3<T> r;
        if (!refract(lwo, owh, eta, &r)) return V3<T>(0);
        const V3<T> lwi = r.normalize();
        if (lwi.z * lwo.z > 0 || ((lwi.z > 0) != (dot(lwh, lwi) > 0))) return V3<T>(0);
        return lwi;
    }
    V3<T> d_sample_inner(const V3<T> &lwo, T u, T v) const {
        const V3<T> lwh = sample_aniso_gtr2(d_tax, d_tay, u, v);
        if (lwh.z <= 0) return V3<T>(0);
        const V3<T> lwi = lwh * (2 * dot(lwo, lwh)) - lwo;
        if (lwi.z > 0) return V3<T>(0);
        return lwi.normalize();
    }
    BSDFSample<T> disney_sample(const V3<T> &wo, T su, T sv, T sw, T eps) const {
        BSDFSample<T> res;
        if (fringe(wo)) {                                        // bsdf.h:142-159
            if (!geo.in_positive_z_hemisphere(wo)) return res;
            V3<T> lwi; T p;
            zweighted_on_hemisphere(su, sv, &lwi, &p);
            if (p < eps) return res;
            const V3<T> wi = geo.local_to_global(lwi).normalize();
            res.dir = wi; res.f = albedo / PI<T> * normal_corr(geo.z, shd.z, wi, eps); res.pdf = p;
            return res;
        }
        const V3<T> lwo = shd.global_to_local(wo).normalize();
        if (std::abs(lwo.z) < eps) return res;
        V3<T> lwi;
        if (lwo.z < 0) {                                         // disney.cpp:536-560
            if (!d_trans) return res;
            const T mF = std::min(T(0.9), std::max(T(0.1), dielectric(d_ior, 1, lwo.z)));
            lwi = su >= mF ? d_sample_trans(lwo, sv, sw) : d_sample_inner(lwo, sv, sw);
        } else {                                                 // disney.cpp:562-590
            // a diffuse pick (sel < w_diffuse) samples the cosine lobe, but the
            // if/else chain that follows always overwrites that direction
            T sel = su;
            if (!(sel < w_diffuse)) sel -= w_diffuse;
            if (sel < w_trans) {
                lwi = d_sample_trans(lwo, sv, sw);
            } else {
                sel -= w_trans;
                if (sel < w_spec) {
                    const V3<T> lwh = ggx_vndf(lwo, sv, sw, eps).normalize();
                    lwi = V3<T>(0);
                    if (lwh.z > 0) {
                        const V3<T> c = (lwh * (2 * dot(lwo, lwh)) - lwo).normalize();
                        if (c.z > 0) lwi = c;
                    }
                } else {
                    const V3<T> lwh = sample_gtr1(d_ccr, sv, sw);
                    lwi = V3<T>(0);
                    if (lwh.z > 0) {
                        const V3<T> c = (lwh * (2 * dot(lwo, lwh)) - lwo).normalize();
                        if (c.z > 0) lwi = c;
                    }
                }
            }
        }
        if (zero(lwi)) return res;
        const V3<T> wi = shd.local_to_global(lwi);
        res.dir = wi;
        res.f = disney_eval(wi, wo);
        res.pdf = disney_pdf(wi, wo);
        return res;
    }

    V3<T> eval(const V3<T> &wi, const V3<T> &wo) const {
        if (kind == AGX_MAT_DISNEY) return disney_eval(wi, wo);
        if (kind == AGX_MAT_GLASS || kind == AGX_MAT_MIRROR || kind == AGX_MAT_IDEAL_BLACK ||
            kind == AGX_MAT_INVISIBLE_SURFACE) return V3<T>(0);
        // AggregateBSDF::eval (component/aggregate.h:39-64)
        if (fringe(wi) || fringe(wo)) {
            if (!geo.in_positive_z_hemisphere(wi) || !geo.in_positive_z_hemisphere(wo)) return V3<T>(0);
            return albedo / PI<T>;
        }
        const V3<T> lwi = shd.global_to_local(wi).normalize();
        const V3<T> lwo = shd.global_to_local(wo).normalize();
        if (!lwi.z || !lwo.z) return V3<T>(0);
        if (kind == AGX_MAT_IDEAL_DIFFUSE) {                    // diffuse_comp.cpp:10-15
            if (lwi.z <= 0 || lwo.z <= 0) return V3<T>(0);
            return coef;
        }
        V3<T> f; ggx_eval_pdf(lwi, lwo, &f, nullptr);
        return f;
    }

    T pdf(const V3<T> &wi, const V3<T> &wo) const {
        if (kind == AGX_MAT_DISNEY) return disney_pdf(wi, wo);
        if (kind == AGX_MAT_GLASS || kind == AGX_MAT_MIRROR || kind == AGX_MAT_IDEAL_BLACK ||
            kind == AGX_MAT_INVISIBLE_SURFACE) return 0;
        // AggregateBSDF::pdf (aggregate.h:218-246)
        if (fringe(wi) || fringe(wo)) {                         // bsdf.h:178-185
            if (geo.in_positive_z_hemisphere(wi) != geo.in_positive_z_hemisphere(wo)) return 0;
            const V3<T> l = geo.global_to_local(wi).normalize();
            return zweighted_on_hemisphere_pdf(l.z);
        }
        const V3<T> lwi = shd.global_to_local(wi).normalize();
        const V3<T> lwo = shd.global_to_local(wo).normalize();
        if (!lwi.z || !lwo.z) return 0;
        T p;
        if (kind == AGX_MAT_IDEAL_DIFFUSE) {
            p = (lwi.z <= 0 || lwo.z <= 0) ? T(0) : zweighted_on_hemisphere_pdf(lwi.z);
        } else {
            ggx_eval_pdf(lwi, lwo, nullptr, &p);
        }
        return T(1) * p / T(1);
    }

    BSDFSample<T> sample(const V3<T> &wo, T su, T sv, T sw, T eps) const {
        if (kind == AGX_MAT_DISNEY) return disney_sample(wo, su, sv, sw, eps);
        BSDFSample<T> res;
        if (kind == AGX_MAT_IDEAL_BLACK) return res;
        if (kind == AGX_MAT_INVISIBLE_SURFACE) {                // invisible_surface.cpp:26-31
            const T cv = std::abs(cos_between(geo.z, wo));
            res.dir = -wo; res.f = V3<T>(1) / (cv < eps ? T(1) : cv); res.pdf = 1; res.is_delta = true;
            return res;
        }
        if (kind == AGX_MAT_MIRROR) {                           // mirror.cpp:32-50
            const V3<T> lo = shd.global_to_local(wo);
            if (lo.z <= 0) return res;
            const V3<T> nwo = lo.normalize();
            const V3<T> dir = shd.local_to_global(V3<T>(0, 0, 2 * nwo.z) - nwo);
            const V3<T> f = conductor(nwo.z, m_eta2, m_k2) * albedo / std::abs(nwo.z);
            res.dir = dir; res.f = f * normal_corr(geo.z, shd.z, dir, eps); res.pdf = 1; res.is_delta = true;
            return res;
        }
        if (kind == AGX_MAT_GLASS) {                            // glass.cpp:52-91
            const V3<T> nwo = shd.global_to_local(wo).normalize();
            const V3<T> nor = nwo.z > 0 ? V3<T>(0, 0, 1) : V3<T>(0, 0, -1);
            const T fr = dielectric(ior, T(1), nwo.z);
            if (su < fr) {
                const V3<T> lwi(-nwo.x, -nwo.y, nwo.z);
                const V3<T> wi = shd.local_to_global(lwi);
                const V3<T> f = c_refl * fr / std::abs(lwi.z);
                res.dir = wi; res.f = f * normal_corr(geo.z, shd.z, wi, eps); res.pdf = fr; res.is_delta = true;
                return res;
            }
            const T eta_i = nwo.z > 0 ? T(1) : ior;
            const T eta_t = nwo.z > 0 ? ior : T(1);
            const T eta = eta_i / eta_t;
            const T cti = std::abs(nwo.z);
            const T st2 = eta * eta * std::max(T(0), 1 - cti * cti);
            if (st2 >= 1) return res;
            const T ctt = std::sqrt(1 - st2);
            const V3<T> nwi = (nor * (eta * cti - ctt) - nwo * eta).normalize();
            const T corr = eta * eta;
            const V3<T> dir = shd.local_to_global(nwi);
            const V3<T> f = c_refr * corr * (1 - fr) / std::abs(nwi.z);
            res.dir = dir; res.f = f * normal_corr(geo.z, shd.z, dir, eps); res.pdf = 1 - fr; res.is_delta = true;
            return res;
        }
        // AggregateBSDF<1>::sample (aggregate.h:66-139)
        if (fringe(wo)) {                                        // bsdf.h:142-159
            if (!geo.in_positive_z_hemisphere(wo)) return res;
            V3<T> lwi; T p;
            zweighted_on_hemisphere(su, sv, &lwi, &p);
            if (p < eps) return res;
            const V3<T> wi = geo.local_to_global(lwi).normalize();
            res.dir = wi; res.f = albedo / PI<T> * normal_corr(geo.z, shd.z, wi, eps); res.pdf = p;
            return res;
        }
        const V3<T> lwo = shd.global_to_local(wo).normalize();
        if (!lwo.z) return res;
        // single component of weight 1: (u - 0.001) * 1 <= 1 always selects it
        if (!((su - T(0.001)) * T(1) <= T(1))) return res;
        V3<T> lwi, f; T p;
        if (kind == AGX_MAT_IDEAL_DIFFUSE) {                    // diffuse_comp.cpp:29-45
            if (lwo.z <= 0) return res;
            zweighted_on_hemisphere(sv, sw, &lwi, &p);
            if (p < T(0.001)) return res;
            f = coef;
        } else {                                                 // microfacet_refl_comp.cpp:85-100
            if (lwo.z <= 0) return res;
            const V3<T> lwh = ggx_vndf(lwo, sv, sw, eps).normalize();
            if (lwh.z <= 0) return res;
            lwi = (lwh * (2 * dot(lwo, lwh)) - lwo).normalize();
            if (lwi.z <= 0) return res;
            ggx_eval_pdf(lwi, lwo, &f, &p);
        }
        lwi = lwi.normalize();
        p *= T(1);
        const V3<T> wi = shd.local_to_global(lwi);
        res.dir = wi; res.f = f * normal_corr(geo.z, shd.z, wi, eps); res.pdf = p;
        return res;
    }
};

// Spectrum::lum (agz-utils, parity unpinned): Rec. 709 luminance
template<class T> inline T spectrum_lum(const V3<T> &c) { return T(0.2126) * c.x + T(0.7152) * c.y + T(0.0722) * c.z; }
// DisneyBSDF::to_tint (disney.cpp:72-76)
template<class T> inline V3<T> disney_tint(const V3<T> &c) {
    const T l = spectrum_lum(c);
    return l > 0 ? c / l : V3<T>(1);
}

template<class T> struct Material {
    agx_material_desc d;
    T eps = T(3e-4);
    // Material::shade -> BSDF (ideal_diffuse.cpp:25-39, metal.cpp:39-62, glass.cpp:180-206, mirror.cpp:98-115)
    BSDF<T> shade(const SurfacePoint<T> &sp, V3<T> *shading_normal) const {
        BSDF<T> b;
        b.kind = d.kind;
        b.eps_ = eps;
        b.geo = sp.geometry_coord;
        b.shd = sp.user_coord;
        const V3<T> col(T(d.color[0]), T(d.color[1]), T(d.color[2]));
        switch (d.kind) {
        case AGX_MAT_IDEAL_DIFFUSE:
            b.albedo = col; b.coef = col / PI<T>;
            break;
        case AGX_MAT_METAL: {
            // create_metal forwards (eta, k) into Metal(color, k, eta): the JSON "k"
            // becomes eta_in and the JSON "eta" becomes k (material_creators.cpp:252-266,
            // metal.cpp:23-29, 65-78).
            b.albedo = col; b.color = col;
            const V3<T> je(T(d.eta[0]), T(d.eta[1]), T(d.eta[2]));
            const V3<T> jk(T(d.k[0]), T(d.k[1]), T(d.k[2]));
            b.eta2 = jk * jk;      // (eta_in / eta_out)^2, eta_out = 1
            b.k2 = je * je;        // (k / eta_out)^2
            const T rough = T(d.roughness), an = T(d.anisotropic);
            const T aspect = an > 0 ? std::sqrt(1 - T(0.9) * an) : T(1);   // microfacet_refl_comp.cpp:72-75
            b.ax = std::max(T(0.001), rough * rough / aspect);
            b.ay = std::max(T(0.001), rough * rough * aspect);
            break;
        }
        case AGX_MAT_GLASS: {
            b.ior = T(d.ior);
            b.c_refl = col;
            const V3<T> c2(T(d.color2[0]), T(d.color2[1]), T(d.color2[2]));
            b.c_refr = d.glass_separate_refraction ? vsqrt(c2) : vsqrt(col);
            b.albedo = (b.c_refl + b.c_refr) * T(0.5);
            break;
        }
        case AGX_MAT_MIRROR: {
            // ConductorPoint(eta_out = eta, eta_in = 1, k) (mirror.cpp:104-105)
            b.albedo = col;
            for (int ch = 0; ch < 3; ++ch) {
                const T eo = T(d.eta[ch]), kk = T(d.k[ch]);
                b.m_eta2[ch] = (1 / eo) * (1 / eo);
                b.m_k2[ch] = (kk / eo) * (kk / eo);
            }
            break;
        }
        case AGX_MAT_INVISIBLE_SURFACE:                     // invisible_surface.cpp:47-50
            b.albedo = V3<T>(1);
            break;
        case AGX_MAT_DISNEY: {                              // disney.cpp:405-457, 732-768
            b.albedo = col;
            b.d_tint = disney_tint(col);
            b.d_metallic = T(d.metallic);
            b.d_rough = T(d.roughness);
            b.d_spec_scale = V3<T>(T(d.color2[0]), T(d.color2[1]), T(d.color2[2]));
            b.d_spec_tint = T(d.specular_tint);
            b.d_sheen = T(d.sheen);
            b.d_sheen_tint = T(d.sheen_tint);
            b.d_trans = T(d.transmission);
            b.d_ior = std::max(T(1.01), T(d.ior));
            const T an = T(d.anisotropic);
            const T aspect = an > 0 ? std::sqrt(1 - T(0.9) * an) : T(1);
            const T r = T(d.roughness), tr = T(d.transmission_roughness);
            b.ax = std::max(T(0.001), r * r / aspect);
            b.ay = std::max(T(0.001), r * r * aspect);
            b.d_tax = std::max(T(0.001), tr * tr / aspect);
            b.d_tay = std::max(T(0.001), tr * tr * aspect);
            b.d_clearcoat = T(d.clearcoat);
            const T g = T(d.clearcoat_gloss);
            T ccr = T(0.1) * (1 - g) + T(0) * g;             // mix(0.1, 0, gloss)
            ccr *= ccr;
            b.d_ccr = std::max(ccr, T(0.0001));
            const T A = std::min(T(0.7), std::max(T(0.3), spectrum_lum(col) * (1 - b.d_metallic)));
            const T B = 1 - A;
            b.w_diffuse = A * (1 - b.d_trans);
            b.w_trans = A * b.d_trans;
            b.w_spec = B * 2 / (2 + b.d_clearcoat);
            b.w_cc = B * b.d_clearcoat / (2 + b.d_clearcoat);
            break;
        }
        default:
            b.albedo = V3<T>(0);
            break;
        }
        *shading_normal = sp.user_coord.z;
        return b;
    }
};

// ------------------------------------------------------------------------------------
// participating media  (core/medium/{medium_void,homogeneous,heterogeneous}.cpp,
//                       utility/phase_function.h, core/texture3d.h, texture3d/image3d.cpp)
// ------------------------------------------------------------------------------------

// Texture3D lookup: transform, per-axis wrap, nearest or trilinear over texel
// centres (agz-utils linear_sample3d / nearest_sample3d: unpinned), ^inv_gamma.
// Texture2D (core/texture2d.h:90-125; texture2d/constant2d.cpp, checker_board.cpp:14-24,
// hdr.cpp:16-36): uv transform, wrap per axis, texture, ^inv_gamma.  The
// bilinear filter over texel centres stands for agz-utils linear_sample2d
// (unpinned, as its 3D counterpart).
template<class T> struct Texture2D {
    int kind = AGX_TEX2D_CONSTANT;
    bool linear = true;
    int w = 1, h = 1, grid = 1;
    std::vector<float> data;
    V3<T> value, value2;
    T xf[6] = {1, 0, 0, 0, 1, 0};
    int wrap[2] = {0, 0};
    T inv_gamma = 1;
    void init(const agx_texture2d_desc &t) {
        kind = t.kind;
        linear = t.linear != 0;
        value = V3<T>(T(t.value[0]), T(t.value[1]), T(t.value[2]));
        value2 = V3<T>(T(t.value2[0]), T(t.value2[1]), T(t.value2[2]));
        for (int i = 0; i < 6; ++i) xf[i] = T(t.transform[i]);
        wrap[0] = t.wrap[0];
        wrap[1] = t.wrap[1];
        inv_gamma = T(t.inv_gamma > 0 ? t.inv_gamma : 1);
        grid = t.grid_count;
        if (kind == AGX_TEX2D_IMAGE) {
            if (t.width <= 0 || t.height <= 0 || !t.data) throw std::runtime_error("invalid 2d texture");
            w = t.width; h = t.height;
            data.assign(t.data, t.data + size_t(w) * size_t(h) * 3);
        } else if (kind == AGX_TEX2D_CHECKER) {
            if (grid <= 0) throw std::runtime_error("invalid grid_count value");
        } else if (kind != AGX_TEX2D_CONSTANT) {
            throw std::runtime_error("unknown 2d texture");
        }
    }
    static T wrapf(int mode, T x) {
        if (mode == AGX_WRAP_REPEAT) return std::min(T(1), std::max(T(0), x - std::floor(x)));
        if (mode == AGX_WRAP_MIRROR) {
            const int ip = int(x >= 0 ? x : std::floor(x));
            const T v = (ip & 1) ? 1 - x + T(ip) : x - T(ip);
            return std::min(T(1), std::max(T(0), v));
        }
        return std::min(T(1), std::max(T(0), x));
    }
    int width() const { return kind == AGX_TEX2D_IMAGE ? w : 1; }
    int height() const { return kind == AGX_TEX2D_IMAGE ? h : 1; }
    V3<T> texel(int x, int y) const {
        const size_t o = (size_t(y) * size_t(w) + size_t(x)) * 3;
        return V3<T>(T(data[o]), T(data[o + 1]), T(data[o + 2]));
    }
    V3<T> sample(T u0, T v0) const {
        const T u = wrapf(wrap[0], xf[0] * u0 + xf[1] * v0 + xf[2]);
        const T v = wrapf(wrap[1], xf[3] * u0 + xf[4] * v0 + xf[5]);
        V3<T> r;
        if (kind == AGX_TEX2D_CONSTANT) {
            r = value;
        } else if (kind == AGX_TEX2D_CHECKER) {
            const T gs = T(1) / T(grid);
            const int ui = int(std::min(T(1), std::max(T(0), u)) / gs), vi = int(std::min(T(1), std::max(T(0), v)) / gs);
            r = ((ui + vi) & 1) ? value : value2;
        } else if (!linear) {
            r = texel(std::min(w - 1, int(u * T(w))), std::min(h - 1, int(v * T(h))));
        } else {
            const T fu = u * T(w) - T(0.5), fv = v * T(h) - T(0.5);
            const T bu = std::floor(fu), bv = std::floor(fv);
            const T du = fu - bu, dv = fv - bv;
            const int x0 = std::min(w - 1, std::max(0, int(bu))), x1 = std::min(w - 1, std::max(0, int(bu) + 1));
            const int y0 = std::min(h - 1, std::max(0, int(bv))), y1 = std::min(h - 1, std::max(0, int(bv) + 1));
            auto lerp = [](const V3<T> &a, const V3<T> &b, T k) { return a + (b - a) * k; };
            r = lerp(lerp(texel(x0, y0), texel(x1, y0), du), lerp(texel(x0, y1), texel(x1, y1), du), dv);
        }
        if (inv_gamma != 1) r = V3<T>(std::pow(r.x, inv_gamma), std::pow(r.y, inv_gamma), std::pow(r.z, inv_gamma));
        return r;
    }
};

// IBL + EnvironmentLightSampler (envir_light/ibl.cpp:24-60, 64-80, 160-170;
// env_sampler.h:17-183): radiance by spherical (phi, theta) -> (u, v); patches
// of at most 200 x 200 weighted by luminance (one texel margin) x solid angle,
// an alias table over them, then uniform cos(theta) and phi inside the patch.
// no_importance_sampling samples a constant 1x1 texture: one patch = uniform sphere.
template<class T> struct IblLight {
    Texture2D<T> tex;
    int pw = 1, ph = 1;
    std::vector<T> probs;
    AliasTable<T> alias;
    static T phi_of(const V3<T> &d) {  // local_angle::phi: atan2(y, x) in [0, 2 pi)
        const T p = std::atan2(d.y, d.x);
        return p < 0 ? p + 2 * PI<T> : p;
    }
    void init(const agx_env_desc &e) {
        tex.init(e.tex);
        const bool uniform = e.no_importance_sampling != 0;
        const int width = uniform ? 1 : tex.width(), height = uniform ? 1 : tex.height();
        pw = std::min(width, 200);
        ph = std::min(height, 200);
        probs.assign(size_t(pw) * ph, T(0));
        T lum_sum = 0;
        for (int y = 0; y < ph; ++y) {
            const T y0 = T(y) / T(ph), y1 = T(y + 1) / T(ph);
            const int ys = std::max(0, int(std::floor(y0 * T(height)) - 1));
            const int yl = std::min(height - 1, int(std::floor(y1 * T(height)) + 1));
            for (int x = 0; x < pw; ++x) {
                const T x0 = T(x) / T(pw), x1 = T(x + 1) / T(pw);
                const int xs = std::max(0, int(std::floor(x0 * T(width)) - 1));
                const int xl = std::min(width - 1, int(std::floor(x1 * T(width)) + 1));
                T lum = 0;
                for (int yy = ys; yy <= yl; ++yy)
                    for (int xx = xs; xx <= xl; ++xx) {
                        const V3<T> c = uniform ? V3<T>(1)
                                                : tex.sample((T(xx) + T(0.5)) / T(width), (T(yy) + T(0.5)) / T(height));
                        lum += T(0.2126) * c.x + T(0.7152) * c.y + T(0.0722) * c.z;
                    }
                const T area = std::abs(2 * PI<T> * (x1 - x0) * (std::cos(PI<T> * y1) - std::cos(PI<T> * y0)));
                probs[size_t(y) * pw + x] = lum * area;
                lum_sum += lum * area;
            }
        }
        if (lum_sum > 0) for (T &p : probs) p *= 1 / lum_sum;
        else for (T &p : probs) p = T(1) / T(probs.size());
        alias.build(probs);
    }
    V3<T> radiance(const V3<T> &d) const {
        const V3<T> dir = d.normalize();
        const T theta = std::acos(std::min(T(1), std::max(T(-1), dir.z)));
        const T u = std::min(T(1), std::max(T(0), phi_of(dir) / (2 * PI<T>)));
        const T v = std::min(T(1), std::max(T(0), theta / PI<T>));
        return tex.sample(u, v);
    }
    T patch_pdf(int px, int py, T *cvmin, T *cvmax, T *u0, T *u1) const {
        *u0 = T(px) / T(pw);
        *u1 = T(px + 1) / T(pw);
        const T v0 = T(py) / T(ph), v1 = T(py + 1) / T(ph);
        const T c1 = std::cos(PI<T> * v1), c0 = std::cos(PI<T> * v0);
        *cvmin = std::min(c1, c0);
        *cvmax = std::max(c1, c0);
        return probs[size_t(py) * pw + px] / (2 * PI<T> * ((*u1 - *u0) * (*cvmax - *cvmin)));
    }
    T pdf(const V3<T> &d) const {
        const V3<T> dir = d.normalize();
        const T theta = std::acos(std::min(T(1), std::max(T(-1), dir.z)));
        const T u = phi_of(dir) / (2 * PI<T>), v = theta / PI<T>;
        const int px = std::min(pw - 1, std::max(0, int(std::floor(u * T(pw)))));
        const int py = std::min(ph - 1, std::max(0, int(std::floor(v * T(ph)))));
        T a, b, c, e;
        return patch_pdf(px, py, &a, &b, &c, &e);
    }
    V3<T> sample(T su, T sv, T sw, T *pdf) const {
        const int32_t pick = alias.sample(su);
        const int px = pick % pw, py = pick / pw;
        T cvmin, cvmax, u0, u1;
        *pdf = patch_pdf(px, py, &cvmin, &cvmax, &u0, &u1);
        const T ct = cvmin + sw * (cvmax - cvmin);
        const T st = std::sqrt(std::max(T(0), 1 - ct * ct));
        const T phi = 2 * PI<T> * (u0 * (1 - sv) + u1 * sv);
        return V3<T>(st * std::cos(phi), st * std::sin(phi), ct);
    }
};

template<class T> struct Texture3D {
    int kind = AGX_TEX3D_CONSTANT;
    bool linear = true;
    int w = 1, h = 1, d = 1, ch = 1;
    std::vector<float> data;
    V3<T> value;
    Affine<T> xf;
    int wrap[3] = {0, 0, 0};
    T inv_gamma = 1;
    void init(const agx_texture3d_desc &t) {
        kind = t.kind;
        linear = t.linear != 0;
        value = V3<T>(T(t.value[0]), T(t.value[1]), T(t.value[2]));
        xf = Affine<T>::from(t.transform);
        for (int i = 0; i < 3; ++i) wrap[i] = t.wrap[i];
        inv_gamma = T(t.inv_gamma > 0 ? t.inv_gamma : 1);
        if (kind == AGX_TEX3D_GRID) {
            w = t.width; h = t.height; d = t.depth; ch = t.channels;
            if (w <= 0 || h <= 0 || d <= 0 || (ch != 1 && ch != 3) || !t.data)
                throw std::runtime_error("invalid 3d texture grid");
            data.assign(t.data, t.data + size_t(w) * size_t(h) * size_t(d) * size_t(ch));
        }
    }
    static T wrapf(int mode, T x) {  // texture3d.h: wrap_clamp / wrap_repeat / wrap_mirror
        if (mode == AGX_WRAP_REPEAT) return std::min(T(1), std::max(T(0), x - std::floor(x)));
        if (mode == AGX_WRAP_MIRROR) {
            const int ip = int(x >= 0 ? x : std::floor(x));
            const T v = (ip & 1) ? 1 - x + T(ip) : x - T(ip);
            return std::min(T(1), std::max(T(0), v));
        }
        return std::min(T(1), std::max(T(0), x));
    }
    V3<T> texel(int x, int y, int z) const {
        const size_t o = ((size_t(z) * size_t(h) + size_t(y)) * size_t(w) + size_t(x)) * size_t(ch);
        if (ch == 1) return V3<T>(T(data[o]));
        return V3<T>(T(data[o]), T(data[o + 1]), T(data[o + 2]));
    }
    // constant3d overrides sample_spectrum/sample_real: no transform, no wrap, no
    // inv_gamma (constant3d.cpp:43-58); image3d goes through texture3d.h:208-233
    V3<T> sample(const V3<T> &uvw) const {
        if (kind == AGX_TEX3D_CONSTANT) return value;
        const V3<T> t = xf.point(uvw);
        const T u = wrapf(wrap[0], t.x), v = wrapf(wrap[1], t.y), q = wrapf(wrap[2], t.z);
        V3<T> r;
        if (!linear) {
            const int x = std::min(w - 1, int(u * T(w))), y = std::min(h - 1, int(v * T(h))), z = std::min(d - 1, int(q * T(d)));
            r = texel(x, y, z);
        } else {
            const T fu = u * T(w) - T(0.5), fv = v * T(h) - T(0.5), fq = q * T(d) - T(0.5);
            const T bu = std::floor(fu), bv = std::floor(fv), bq = std::floor(fq);
            const T du = fu - bu, dv = fv - bv, dq = fq - bq;
            auto cl = [](int i, int n) { return std::min(n - 1, std::max(0, i)); };
            const int x0 = cl(int(bu), w), x1 = cl(int(bu) + 1, w);
            const int y0 = cl(int(bv), h), y1 = cl(int(bv) + 1, h);
            const int z0 = cl(int(bq), d), z1 = cl(int(bq) + 1, d);
            auto lerp = [](const V3<T> &a, const V3<T> &b, T t_) { return a + (b - a) * t_; };
            const V3<T> c00 = lerp(texel(x0, y0, z0), texel(x1, y0, z0), du);
            const V3<T> c10 = lerp(texel(x0, y1, z0), texel(x1, y1, z0), du);
            const V3<T> c01 = lerp(texel(x0, y0, z1), texel(x1, y0, z1), du);
            const V3<T> c11 = lerp(texel(x0, y1, z1), texel(x1, y1, z1), du);
            r = lerp(lerp(c00, c10, dv), lerp(c01, c11, dv), dq);
        }
        return gamma(r);
    }
    V3<T> gamma(const V3<T> &c) const {
        if (inv_gamma == 1) return c;
        return V3<T>(std::pow(c.x, inv_gamma), std::pow(c.y, inv_gamma), std::pow(c.z, inv_gamma));
    }
    // max_real of the stored texels, before inv_gamma (image3d.cpp:108-147, constant3d.cpp:43-46)
    T max_real() const {
        if (kind == AGX_TEX3D_CONSTANT) return value.x;
        T m = -std::numeric_limits<T>::max();
        for (size_t i = 0; i < data.size(); i += size_t(ch)) m = std::max(m, T(data[i]));
        return m;
    }
};

// Henyey-Greenstein phase function as a BSDF (utility/phase_function.h:8-90)
template<class T> struct HG {
    T g = 0;
    T phase(T u) const {
        const T g2 = g * g;
        const T dem = 1 + g2 - 2 * g * u;
        return (1 - g2) / (4 * PI<T> * dem * std::sqrt(dem));
    }
    T eval(const V3<T> &wi, const V3<T> &wo) const { return phase(-cos_between(wi, wo)); }
    void sample(const V3<T> &wo, T su, T sv, V3<T> *dir, T *f) const {
        const T s = 2 * su - 1;
        T u;
        if (std::abs(g) < T(0.001)) u = s;
        else {
            const T g2 = g * g;
            const T q = (1 - g2) / (1 + g * s);
            u = (1 + g2 - q * q) / (2 * g);
        }
        const T ct = -u;
        const T st = std::sqrt(std::max(T(0), 1 - ct * ct));
        const T phi = 2 * PI<T> * sv;
        const V3<T> local(st * std::sin(phi), st * std::cos(phi), ct);
        *dir = Coord<T>::from_z(wo).local_to_global(local);
        *f = phase(u);
    }
};

template<class T> struct MediumScatter {
    bool happened = false;
    V3<T> pos, wr, throughput{1};
    HG<T> phase;
};

// Tracking loops stop after this many steps and on a NaN segment length; for
// finite segments below that count they are Atrc's loops (heterogeneous.cpp:58-71, 88-120).
constexpr int MAX_TRACKING_STEPS = 1 << 20;

template<class T> struct Medium {
    int kind = AGX_MEDIUM_VOID;
    int max_scat = std::numeric_limits<int>::max();
    V3<T> sa, ss, st;
    T g = 0;
    Affine<T> l2w, w2l;
    Texture3D<T> density, albedo, gtex;
    T max_density = 1, inv_max = 1;
    bool white_for_indirect = false;
    T eps = T(3e-4);

    void init(const agx_medium_desc &m, T eps_) {
        eps = eps_;
        kind = m.kind;
        max_scat = m.max_scattering_count;
        if (kind == AGX_MEDIUM_HOMOGENEOUS) {               // homogeneous.cpp:23-41
            sa = V3<T>(T(m.sigma_a[0]), T(m.sigma_a[1]), T(m.sigma_a[2]));
            ss = V3<T>(T(m.sigma_s[0]), T(m.sigma_s[1]), T(m.sigma_s[2]));
            st = sa + ss;
            g = T(m.g);
            if (g <= -1 || g >= 1) throw std::runtime_error("invalid g value");
        } else if (kind == AGX_MEDIUM_HETEROGENEOUS) {      // heterogeneous.cpp:27-45
            l2w = Affine<T>::from(m.local_to_world);
            w2l = Affine<T>::from(m.world_to_local);
            density.init(m.density);
            albedo.init(m.albedo);
            gtex.init(m.g3);
            max_density = std::max(density.max_real(), eps);
            inv_max = 1 / max_density;
            white_for_indirect = m.white_for_indirect != 0;
        } else if (kind != AGX_MEDIUM_VOID) {
            throw std::runtime_error("unknown medium kind");
        }
    }
    static T dist(const V3<T> &a, const V3<T> &b) { return (a - b).length(); }
    static V3<T> lerp(const V3<T> &a, const V3<T> &b, T t) { return a + (b - a) * t; }

    // Medium::tr; `s` is the stream the caller designates (ratio tracking draws)
    V3<T> tr(const V3<T> &a, const V3<T> &b, Sampler<T> &s) const {
        if (kind == AGX_MEDIUM_HOMOGENEOUS) {               // homogeneous.cpp:48-57
            const V3<T> e = -st * dist(a, b);
            return V3<T>(std::exp(e.x), std::exp(e.y), std::exp(e.z));
        }
        if (kind == AGX_MEDIUM_HETEROGENEOUS) {             // heterogeneous.cpp:52-72
            T result = 1, t = 0;
            const T t_max = dist(a, b);
            for (int step = 0; step < MAX_TRACKING_STEPS; ++step) {
                t += -std::log(1 - s.u1()) * inv_max;
                if (!(t < t_max)) break;
                const V3<T> unit = w2l.point(lerp(a, b, t / t_max));
                result *= 1 - density.sample(unit).x / max_density;
            }
            return V3<T>(result);
        }
        return V3<T>(1);
    }
    V3<T> ab(const V3<T> &a, const V3<T> &b, Sampler<T> &s) const {
        if (kind == AGX_MEDIUM_HOMOGENEOUS) {
            const V3<T> e = -sa * dist(a, b);
            return V3<T>(std::exp(e.x), std::exp(e.y), std::exp(e.z));
        }
        if (kind == AGX_MEDIUM_HETEROGENEOUS) return tr(a, b, s);  // heterogeneous.cpp:74-79
        return V3<T>(1);
    }
    // Medium::sample_scattering (homogeneous.cpp:62-110, heterogeneous.cpp:81-123)
    MediumScatter<T> sample_scattering(const V3<T> &a, const V3<T> &b, Sampler<T> &s, bool indirect) const {
        MediumScatter<T> r;
        if (kind == AGX_MEDIUM_HOMOGENEOUS) {
            const T sam = s.u1();
            if (ss.is_black()) { r.throughput = tr(a, b, s); return r; }
            const T t_max = dist(a, b);
            // extract_uniform_int(u, 0, 3): channel and the re-scaled remainder
            int chn = int(sam * T(3));
            if (chn > 2) chn = 2;
            const T nsam = sam * T(3) - T(chn);
            const T stv = -std::log(nsam) / st[chn];
            const bool in_medium = stv < t_max;
            V3<T> trv;
            for (int i = 0; i < 3; ++i) trv[i] = std::exp(-st[i] * std::min(stv, t_max));
            const V3<T> dens = in_medium ? st * trv : trv;
            T pdf = 0;
            for (int i = 0; i < 3; ++i) pdf += dens[i];
            pdf /= 3;
            pdf = std::max(pdf, eps);
            V3<T> thr = trv / pdf;
            if (in_medium) thr *= ss;
            r.throughput = thr;
            if (in_medium) {
                r.happened = true;
                r.pos = lerp(a, b, stv / t_max);
                r.wr = (a - b) / t_max;
                r.phase.g = g;
            }
            return r;
        }
        if (kind == AGX_MEDIUM_HETEROGENEOUS) {
            const T t_max = dist(a, b);
            T t = 0;
            for (int step = 0; step < MAX_TRACKING_STEPS; ++step) {
                t += -std::log(1 - s.u1()) * inv_max;
                if (!(t < t_max)) break;
                const V3<T> pos = lerp(a, b, t / t_max);
                const V3<T> unit = w2l.point(pos);
                const T dn = density.sample(unit).x;
                if (s.u1() < dn / max_density) {
                    const V3<T> alb = white_for_indirect && indirect ? V3<T>(1) : albedo.sample(unit);
                    r.happened = true;
                    r.pos = pos;
                    r.wr = (a - b) / t_max;
                    r.throughput = alb;
                    r.phase.g = gtex.sample(unit).x;
                    return r;
                }
            }
            return r;
        }
        return r;  // void: no scattering, throughput 1
    }
};

// ------------------------------------------------------------------------------------
// scene, lights  (core/scene/scene_default.cpp, entity/area_light, envir_light/native_sky.cpp)
// ------------------------------------------------------------------------------------

template<class T> struct Entity {
    Geometry<T> geom;
    int32_t material = 0;
    V3<T> emit;
    bool is_light = false;
    bool no_denoise = false;
    int32_t med_in = -1, med_out = -1;  // index into Scene::media, -1 = void
};

template<class T> struct LightSample {
    V3<T> pos, nor, radiance; T pdf = 0;
};

template<class T> struct Scene {
    using Real = T;
    std::vector<Entity<T>> entities;
    std::vector<Material<T>> materials;
    std::vector<Medium<T>> media;
    Medium<T> void_medium;
    // EntityIntersection::medium(d) (intersection.h:52-65): the side of the surface d points to
    const Medium<T> &medium_of(const EntityHit<T> &h, const V3<T> &d) const {
        const Entity<T> &e = entities[size_t(h.entity)];
        const int32_t m = dot(d, h.geometry_coord.z) >= 0 ? e.med_out : e.med_in;
        return m < 0 ? void_medium : media[size_t(m)];
    }
    int env_kind = AGX_ENV_NONE;
    V3<T> sky_top, sky_bottom;
    IblLight<T> ibl;
    T world_radius = 1;
    V3<T> world_centre;
    std::vector<int32_t> lights;  // -1 = environment, else entity index (scene_default.cpp:66-82)
    int aggregate = AGX_AGGREGATE_NATIVE;
    // entity BVH (core/aggregate/entity_bvh.cpp)
    struct ENode { AABB<T> bound; bool leaf; size_t a, b; };  // leaf: [a, b) into eprims; else children a, b
    std::vector<ENode> enodes;
    std::vector<int32_t> eprims;
    int max_leaf = 5;
    T eps = T(3e-4);

    void build(const agx_scene_desc &d) {
        if (d.abi_version != AGX_ABI_VERSION) throw std::runtime_error("abi version mismatch");
        entities.clear(); entities.resize(d.entity_count);
        for (uint32_t i = 0; i < d.entity_count; ++i) {
            const agx_entity_desc &e = d.entities[i];
            Entity<T> &en = entities[i];
            en.geom.init(e);
            if (e.material < 0 || uint32_t(e.material) >= d.material_count) throw std::runtime_error("bad material index");
            en.material = e.material;
            en.emit = V3<T>(T(e.emit_radiance[0]), T(e.emit_radiance[1]), T(e.emit_radiance[2]));
            en.is_light = !en.emit.is_black();
            en.no_denoise = e.no_denoise != 0;
            en.med_in = e.medium_in;
            en.med_out = e.medium_out;
            if (en.med_in >= int32_t(d.medium_count) || en.med_out >= int32_t(d.medium_count))
                throw std::runtime_error("bad medium index");
        }
        media.clear();
        media.resize(d.medium_count);
        for (uint32_t i = 0; i < d.medium_count; ++i) media[i].init(d.media[i], eps);
        materials.resize(d.material_count);
        for (uint32_t i = 0; i < d.material_count; ++i) {
            materials[i].d = d.materials[i];
            materials[i].eps = eps;
        }
        env_kind = d.env.kind;
        sky_top = V3<T>(T(d.env.top[0]), T(d.env.top[1]), T(d.env.top[2]));
        sky_bottom = V3<T>(T(d.env.bottom[0]), T(d.env.bottom[1]), T(d.env.bottom[2]));
        if (env_kind == AGX_ENV_IBL) ibl.init(d.env);
        else if (env_kind != AGX_ENV_NONE && env_kind != AGX_ENV_NATIVE_SKY)
            throw std::runtime_error("unknown environment light");
        lights.clear();
        if (env_kind != AGX_ENV_NONE) lights.push_back(-1);
        for (uint32_t i = 0; i < d.entity_count; ++i) if (entities[i].is_light) lights.push_back(int32_t(i));
        // Scene::start_rendering -> EnvirLight::preprocess (scene_default.cpp:162-172, core/core/light.cpp:6-14)
        AABB<T> wb;
        for (auto &en : entities) wb.grow(en.geom.bound);
        const T diag = (wb.high - wb.low).length();
        world_radius = T(1.1) * diag / 2;
        world_centre = (wb.low + wb.high) * T(0.5);
        aggregate = d.aggregate;
        max_leaf = d.max_leaf_size > 0 ? d.max_leaf_size : 5;
        enodes.clear(); eprims.clear();
        if (aggregate == AGX_AGGREGATE_BVH) {
            if (entities.empty()) { enodes.push_back({AABB<T>{V3<T>(0), V3<T>(1)}, true, 0, 0}); }
            else {
                std::vector<std::pair<int32_t, AABB<T>>> recs;
                for (size_t i = 0; i < entities.size(); ++i) recs.push_back({int32_t(i), entities[i].geom.bound});
                build_ebvh(recs.data(), recs.size());
            }
        }
    }
    // EntityBVH::build_aux (entity_bvh.cpp:46-116)
    size_t build_ebvh(std::pair<int32_t, AABB<T>> *recs, size_t count) {
        if (count <= size_t(max_leaf) || count < 2) {
            const size_t s = eprims.size();
            AABB<T> b;
            for (size_t i = 0; i < count; ++i) { eprims.push_back(recs[i].first); b.grow(recs[i].second); }
            enodes.push_back({b, true, s, s + count});
            return enodes.size() - 1;
        }
        AABB<T> all;
        for (size_t i = 0; i < count; ++i) all.grow(recs[i].second);
        T best = -1; int axis = 0;
        for (int i = 0; i < 3; ++i) { const T l = all.high[i] - all.low[i]; if (l > best) { best = l; axis = i; } }
        std::sort(recs, recs + count, [axis](const std::pair<int32_t, AABB<T>> &L, const std::pair<int32_t, AABB<T>> &R) {
            return L.second.low[axis] + L.second.high[axis] < R.second.low[axis] + R.second.high[axis];
        });
        const size_t idx = enodes.size();
        enodes.push_back({all, false, 0, 0});
        const size_t half = count / 2;
        const size_t l = build_ebvh(recs, half);
        const size_t r = build_ebvh(recs + half, count - half);
        enodes[idx].a = l; enodes[idx].b = r;
        return idx;
    }

    bool entity_closest(int32_t ei, Ray<T> &r, EntityHit<T> *h, Counters *c) const {
        GeometryHit<T> g;
        if (!entities[size_t(ei)].geom.closest(r, &g, eps, c)) return false;
        static_cast<GeometryHit<T> &>(*h) = g;
        h->entity = ei;
        r.t_max = g.t;
        return true;
    }
    bool ebvh_closest(size_t ni, const V3<T> &inv, Ray<T> &r, EntityHit<T> *h, Counters *c) const {
        const ENode &n = enodes[ni];
        if (c) c->node_visits++;
        if (!n.bound.intersect(r.o, inv, r.t_min, r.t_max)) return false;
        if (n.leaf) {
            bool ret = false;
            for (size_t i = n.a; i < n.b; ++i) ret = entity_closest(eprims[i], r, h, c) || ret;
            return ret;
        }
        const bool L = ebvh_closest(n.a, inv, r, h, c);
        const bool R = ebvh_closest(n.b, inv, r, h, c);
        return L || R;
    }
    bool ebvh_any(size_t ni, const V3<T> &inv, const Ray<T> &r, Counters *c) const {
        const ENode &n = enodes[ni];
        if (c) c->node_visits++;
        if (!n.bound.intersect(r.o, inv, r.t_min, r.t_max)) return false;
        if (n.leaf) {
            for (size_t i = n.a; i < n.b; ++i) if (entities[size_t(eprims[i])].geom.any_hit(r, c)) return true;
            return false;
        }
        return ebvh_any(n.a, inv, r, c) || ebvh_any(n.b, inv, r, c);
    }

    // DefaultScene::closest_intersection -> Aggregate (native_aggregate.cpp:31-45, entity_bvh.cpp:209-215)
    bool closest(const Ray<T> &r0, EntityHit<T> *h, Counters *c) const {
        if (c) c->closest++;
        Ray<T> r = r0;
        bool ret = false;
        if (aggregate == AGX_AGGREGATE_BVH) {
            const V3<T> inv(1 / r.d.x, 1 / r.d.y, 1 / r.d.z);
            ret = ebvh_closest(0, inv, r, h, c);
        } else {
            for (size_t i = 0; i < entities.size(); ++i) ret = entity_closest(int32_t(i), r, h, c) || ret;
        }
        if (ret && c) c->hits++;
        return ret;
    }
    bool any(const Ray<T> &r, Counters *c) const {
        if (c) c->shadow++;
        if (aggregate == AGX_AGGREGATE_BVH) {
            const V3<T> inv(1 / r.d.x, 1 / r.d.y, 1 / r.d.z);
            return ebvh_any(0, inv, r, c);
        }
        for (auto &e : entities) if (e.geom.any_hit(r, c)) return true;
        return false;
    }
    bool visible(const V3<T> &A, const V3<T> &B, Counters *c) const {  // scene_default.cpp:136-141
        const T dis = (A - B).length();
        return !any(Ray<T>(A, (B - A).normalize(), eps, dis - eps), c);
    }

    // NativeSky (native_sky.cpp:13-20, 71-105) or IBL
    V3<T> sky_radiance(const V3<T> &d) const {
        if (env_kind == AGX_ENV_IBL) return ibl.radiance(d);
        const T ct = std::min(T(1), std::max(T(-1), d.normalize().z));
        const T s = (ct + 1) / 2;
        return sky_top * s + sky_bottom * (1 - s);
    }
    T env_pdf(const V3<T> &d) const {
        return env_kind == AGX_ENV_IBL ? ibl.pdf(d) : uniform_on_sphere_pdf<T>;
    }
    V3<T> env_sample(T su, T sv, T sw, T *pdf) const {
        if (env_kind == AGX_ENV_IBL) return ibl.sample(su, sv, sw, pdf);
        *pdf = uniform_on_sphere_pdf<T>;
        return uniform_on_sphere(su, sv);
    }
    V3<T> env_emit_pos(const V3<T> &ref, const V3<T> &ref_to_light) const {
        // point where the ray ref + t d meets the plane through centre + R d, normal d
        const V3<T> d = ref_to_light.normalize();
        const T t = world_radius + dot(world_centre - ref, d);
        return ref + d * t;
    }

    // GeometryToDiffuseLight::sample (geometry_to_diffuse_light.cpp:14-31)
    LightSample<T> area_sample(int32_t ei, const V3<T> &ref, const T s[5]) const {
        LightSample<T> ls;
        const Entity<T> &en = entities[size_t(ei)];
        T pdf_area;
        const SurfacePoint<T> sp = en.geom.sample(ref, &pdf_area, s[0], s[1], s[3], eps);
        if (dot(sp.geometry_coord.z, ref - sp.pos) <= 0) return ls;
        const V3<T> spt_to_ref = ref - sp.pos;
        const T dist2 = spt_to_ref.length_square();
        ls.pdf = pdf_area * dist2 / std::abs(cos_between(sp.geometry_coord.z, spt_to_ref));
        ls.pos = sp.pos; ls.nor = sp.geometry_coord.z; ls.radiance = en.emit;
        return ls;
    }
    // GeometryToDiffuseLight::radiance / pdf (:68-90)
    V3<T> area_radiance(int32_t ei, const V3<T> &nor, const V3<T> &light_to_out) const {
        return dot(nor, light_to_out) > 0 ? entities[size_t(ei)].emit : V3<T>(0);
    }
    T area_pdf(int32_t ei, const V3<T> &ref, const V3<T> &pos, const V3<T> &nor) const {
        if (dot(nor, ref - pos) <= 0) return 0;
        const T area_pdf = entities[size_t(ei)].geom.pdf(ref);
        const V3<T> spt_to_ref = ref - pos;
        const T dist2 = spt_to_ref.length_square();
        const T abscos = std::abs(cos_between(nor, spt_to_ref));
        return area_pdf * (dist2 / abscos);
    }
};

// ------------------------------------------------------------------------------------
// trace_std  (render/path_tracing.cpp:15-225, render/direct_illum.cpp:11-192)
// ------------------------------------------------------------------------------------

template<class T> struct PixelResult {
    V3<T> value, albedo, normal;
    T denoise = 1;
};

// PTRendererParams that reach the trace functions; use_mis selects trace_std
// (1) or trace_nomis (0) as PathTracingRenderer does (pt.cpp:29-31)
struct TraceParams {
    int min_depth = 5, max_depth = 10, specular_depth = 20;
    float cont_prob = 0.9f;
    bool use_mis = true;
};

template<class T>
V3<T> mis_sample_light(const Scene<T> &scene, int32_t light, const EntityHit<T> &inct, const BSDF<T> &bsdf,
                       const T s5[5], Sampler<T> &trs, Counters *c);
template<class T>
V3<T> mis_sample_bsdf(const Scene<T> &scene, const EntityHit<T> &inct, const BSDF<T> &bsdf, const T s3[3],
                      Sampler<T> &trs, Counters *c);
template<class T>
V3<T> mis_sample_light_medium(const Scene<T> &scene, int32_t light, const MediumScatter<T> &ms,
                              const Medium<T> &med, Sampler<T> &sam, Sampler<T> &trs, Counters *c);
template<class T>
V3<T> mis_sample_bsdf_medium(const Scene<T> &scene, const MediumScatter<T> &ms, const Medium<T> &med,
                             const T s3[3], Sampler<T> &trs, Counters *c);

// `sam` is the path's stream in Atrc's draw order; `trs` feeds the ratio-tracking
// transmittance estimates of NEE and MIS-BSDF rays (a separate stream: the GPU
// evaluates them at another point of the pipeline than Atrc's call order).
template<class T>
PixelResult<T> trace_std(const TraceParams &p, const Scene<T> &scene, Ray<T> r, Sampler<T> &sam, Sampler<T> &trs,
                         Counters *c) {
    const T eps = scene.eps;
    V3<T> coef(1);
    PixelResult<T> pixel;
    int scattering_count = 0;
    for (int depth = 1, s_depth = 1; depth <= p.max_depth; ++depth) {
        if (depth > p.min_depth) {                                          // :30-35
            if (sam.u1() > T(p.cont_prob)) return pixel;
            coef /= T(p.cont_prob);
        }
        EntityHit<T> inct;
        if (!scene.closest(r, &inct, c)) {                                  // :39-49
            if (depth == 1 && scene.env_kind != AGX_ENV_NONE) pixel.value += coef * scene.sky_radiance(r.d);
            return pixel;
        }
        const Entity<T> &ent = scene.entities[size_t(inct.entity)];
        V3<T> shading_normal;
        const BSDF<T> bsdf = scene.materials[size_t(ent.material)].shade(inct, &shading_normal);
        if (depth == 1) {                                                   // :54-60
            pixel.normal = shading_normal;
            pixel.albedo = bsdf.albedo;
            if (ent.no_denoise) pixel.denoise = 0;
        }
        // medium between r.o and the hit (:64-117)
        const Medium<T> &med = scene.medium_of(inct, inct.wr);
        if (scattering_count < med.max_scat) {
            const MediumScatter<T> ms = med.sample_scattering(r.o, inct.pos, sam, scattering_count > 0);
            coef *= ms.throughput;
            if (ms.happened) {
                ++scattering_count;
                V3<T> direct;
                for (int32_t light : scene.lights) direct += coef * mis_sample_light_medium(scene, light, ms, med, sam, trs, c);
                {
                    T s3[3];
                    for (int k = 0; k < 3; ++k) s3[k] = sam.u1();
                    direct += coef * mis_sample_bsdf_medium(scene, ms, med, s3, trs, c);
                }
                pixel.value += direct / T(1);
                const T u0 = sam.u1(), u1 = sam.u1();
                (void)sam.u1();
                V3<T> dir;
                T f;
                ms.phase.sample(ms.wr, u0, u1, &dir, &f);
                if (f == 0 || f < eps) return pixel;
                r = Ray<T>(ms.pos, dir.normalize());
                coef *= V3<T>(f) / f;
                continue;
            }
        } else {
            coef *= med.ab(r.o, inct.pos, sam);
        }
        scattering_count = 0;
        if (depth == 1 && ent.is_light)                                     // :123-130
            pixel.value += coef * scene.area_radiance(inct.entity, inct.geometry_coord.z, inct.wr);

        V3<T> direct;                                                       // :134-146
        for (int32_t light : scene.lights) {
            T s5[5];
            for (int k = 0; k < 5; ++k) s5[k] = sam.u1();
            direct += coef * mis_sample_light(scene, light, inct, bsdf, s5, trs, c);
        }
        {
            T s3[3];
            for (int k = 0; k < 3; ++k) s3[k] = sam.u1();
            direct += coef * mis_sample_bsdf(scene, inct, bsdf, s3, trs, c);
        }
        pixel.value += direct * (T(1) / T(1));

        T s3[3];                                                            // :150-169
        for (int k = 0; k < 3; ++k) s3[k] = sam.u1();
        const BSDFSample<T> bs = bsdf.sample(inct.wr, s3[0], s3[1], s3[2], eps);
        if (bs.f.is_black() || bs.pdf < eps) return pixel;
        const T abscos = std::abs(cos_between(inct.geometry_coord.z, bs.dir));
        coef *= bs.f * abscos / bs.pdf;
        r = Ray<T>(inct.eps_offset(bs.dir, eps), bs.dir.normalize());
        if (bs.is_delta && depth >= 2 && s_depth <= p.specular_depth) { --depth; ++s_depth; }
    }
    return pixel;
}

// render::trace_nomis (path_tracing.cpp:227-381): no light sampling; the
// environment is added wherever a path escapes and an area light's emission
// wherever a path meets it.  Atrc's differences from trace_std are kept: the
// phase function samples around the surface hit's wr (path_tracing.cpp:290-291),
// a BSDF or phase sample is rejected only when f is black (no pdf < EPS test),
// and the next ray keeps the sampled direction as it is (Ray does not normalise).
template<class T>
PixelResult<T> trace_nomis(const TraceParams &p, const Scene<T> &scene, Ray<T> r, Sampler<T> &sam, Counters *c) {
    const T eps = scene.eps;
    V3<T> coef(1);
    PixelResult<T> pixel;
    int scattering_count = 0;
    for (int depth = 1, s_depth = 1; depth <= p.max_depth; ++depth) {
        if (depth > p.min_depth) {                                          // :241-246
            if (sam.u1() > T(p.cont_prob)) return pixel;
            coef /= T(p.cont_prob);
        }
        EntityHit<T> inct;
        if (!scene.closest(r, &inct, c)) {                                  // :250-256
            if (scene.env_kind != AGX_ENV_NONE) pixel.value += coef * scene.sky_radiance(r.d);
            return pixel;
        }
        const Entity<T> &ent = scene.entities[size_t(inct.entity)];
        V3<T> shading_normal;
        const BSDF<T> bsdf = scene.materials[size_t(ent.material)].shade(inct, &shading_normal);
        if (depth == 1) {                                                   // :260-266
            pixel.normal = shading_normal;
            pixel.albedo = bsdf.albedo;
            if (ent.no_denoise) pixel.denoise = 0;
        }
        const Medium<T> &med = scene.medium_of(inct, inct.wr);              // :270-307
        if (scattering_count < med.max_scat) {
            const MediumScatter<T> ms = med.sample_scattering(r.o, inct.pos, sam, scattering_count > 0);
            coef *= ms.throughput;
            if (ms.happened) {
                ++scattering_count;
                const T u0 = sam.u1(), u1 = sam.u1();
                (void)sam.u1();
                V3<T> dir;
                T f;
                ms.phase.sample(inct.wr, u0, u1, &dir, &f);
                if (f == 0) return pixel;
                coef *= V3<T>(f) / f;
                r = Ray<T>(ms.pos, dir);
                continue;
            }
        } else {
            coef *= med.ab(r.o, inct.pos, sam);
        }
        scattering_count = 0;
        if (ent.is_light)                                                   // :313-318
            pixel.value += coef * scene.area_radiance(inct.entity, inct.geometry_coord.z, inct.wr);
        T s3[3];                                                            // :320-337
        for (int k = 0; k < 3; ++k) s3[k] = sam.u1();
        const BSDFSample<T> bs = bsdf.sample(inct.wr, s3[0], s3[1], s3[2], eps);
        if (bs.f.is_black()) return pixel;
        const T abscos = std::abs(cos_between(inct.geometry_coord.z, bs.dir));
        coef *= bs.f * abscos / bs.pdf;
        r = Ray<T>(inct.eps_offset(bs.dir, eps), bs.dir);
        if (bs.is_delta && depth >= 2 && s_depth <= p.specular_depth) { --depth; ++s_depth; }
    }
    return pixel;
}

template<class T>
PixelResult<T> trace_path(const TraceParams &p, const Scene<T> &scene, const Ray<T> &r, Sampler<T> &sam,
                          Sampler<T> &trs, Counters *c) {
    return p.use_mis ? trace_std(p, scene, r, sam, trs, c) : trace_nomis(p, scene, r, sam, c);
}

// mis_sample_area_light / mis_sample_envir_light (direct_illum.cpp:11-42, 72-99)
template<class T>
V3<T> mis_sample_light(const Scene<T> &scene, int32_t light, const EntityHit<T> &inct, const BSDF<T> &bsdf,
                       const T s5[5], Sampler<T> &trs, Counters *c) {
    const T eps = scene.eps;
    if (light >= 0) {
        const LightSample<T> ls = scene.area_sample(light, inct.pos, s5);
        if (ls.radiance.is_black() || !ls.pdf) return V3<T>(0);
        const T len = (ls.pos - inct.pos).length() - eps;
        if (len <= eps) return V3<T>(0);
        const V3<T> to_light = (ls.pos - inct.pos).normalize();
        // Atrc tests visibility first; the factors only multiply, so the
        // visibility test is moved last (where the wavefront learns it) and the
        // transmittance draws happen in the same order on CPU and GPU
        const Medium<T> &med = scene.medium_of(inct, to_light);
        const V3<T> f = bsdf.eval(to_light, inct.wr);
        if (f.is_black()) return V3<T>(0);
        const V3<T> ff = med.tr(ls.pos, inct.pos, trs) * ls.radiance * f * std::abs(cos_between(to_light, inct.geometry_coord.z));
        const T bpdf = bsdf.pdf(to_light, inct.wr);
        if (scene.any(Ray<T>(inct.pos, to_light, eps, len), c)) return V3<T>(0);
        return ff / (ls.pdf + bpdf);
    }
    // environment: native sky (uniform sphere) or IBL (patch sampler); emit position on the bounding plane
    T epdf;
    const V3<T> dir = scene.env_sample(s5[0], s5[1], s5[2], &epdf);
    const V3<T> pos = scene.env_emit_pos(inct.pos, dir);
    const V3<T> rad = scene.sky_radiance(dir);
    if (rad.is_black()) return V3<T>(0);
    const V3<T> to_light = (pos - inct.pos).normalize();
    const V3<T> f = bsdf.eval(to_light, inct.wr);
    if (f.is_black()) return V3<T>(0);
    const V3<T> ff = rad * f * std::abs(cos_between(to_light, inct.geometry_coord.z));
    const T bpdf = bsdf.pdf(to_light, inct.wr);
    if (!scene.visible(inct.pos, pos, c)) return V3<T>(0);
    return ff / (epdf + bpdf);
}

// mis_sample_bsdf (direct_illum.cpp:130-192)
template<class T>
V3<T> mis_sample_bsdf(const Scene<T> &scene, const EntityHit<T> &inct, const BSDF<T> &bsdf, const T s3[3],
                      Sampler<T> &trs, Counters *c) {
    const T eps = scene.eps;
    BSDFSample<T> bs = bsdf.sample(inct.wr, s3[0], s3[1], s3[2], eps);
    if (bs.f.is_black()) return V3<T>(0);
    bs.dir = bs.dir.normalize();
    const Ray<T> nr(inct.eps_offset(bs.dir, eps), bs.dir);
    EntityHit<T> h;
    const bool hit = scene.closest(nr, &h, c);
    const Medium<T> &med = scene.medium_of(inct, bs.dir);
    if (!hit) {
        if (scene.env_kind == AGX_ENV_NONE) return V3<T>(0);
        const V3<T> lr = scene.sky_radiance(nr.d);
        if (lr.is_black()) return V3<T>(0);
        const V3<T> f = lr * bs.f * std::abs(dot(inct.geometry_coord.z, nr.d));
        if (bs.is_delta) return f / bs.pdf;
        return f / (bs.pdf + scene.env_pdf(nr.d));
    }
    if (!scene.entities[size_t(h.entity)].is_light) return V3<T>(0);
    const V3<T> lrad = scene.area_radiance(h.entity, h.geometry_coord.z, h.wr);
    if (lrad.is_black()) return V3<T>(0);
    const V3<T> f = med.tr(nr.o, h.pos, trs) * lrad * bs.f * std::abs(dot(inct.geometry_coord.z, nr.d));
    if (bs.is_delta) return f / bs.pdf;
    const T lpdf = scene.area_pdf(h.entity, nr.o, h.pos, h.geometry_coord.z);
    return f / (bs.pdf + lpdf);
}

// mis_sample_area_light at a medium point (direct_illum.cpp:44-70); the
// environment light contributes nothing there and draws nothing (:101-108)
template<class T>
V3<T> mis_sample_light_medium(const Scene<T> &scene, int32_t light, const MediumScatter<T> &ms,
                              const Medium<T> &med, Sampler<T> &sam, Sampler<T> &trs, Counters *c) {
    if (light < 0) return V3<T>(0);
    T s5[5];
    for (int k = 0; k < 5; ++k) s5[k] = sam.u1();
    const LightSample<T> ls = scene.area_sample(light, ms.pos, s5);
    if (ls.radiance.is_black() || !ls.pdf) return V3<T>(0);
    const V3<T> to_light = (ls.pos - ms.pos).normalize();
    const T fp = ms.phase.eval(to_light, ms.wr);
    if (fp == 0) return V3<T>(0);
    const V3<T> f = med.tr(ms.pos, ls.pos, trs) * ls.radiance * V3<T>(fp);
    if (!scene.visible(ls.pos, ms.pos, c)) return V3<T>(0);  // visibility last (see mis_sample_light)
    return f / (ls.pdf + fp);
}

// mis_sample_bsdf at a medium point (direct_illum.cpp:194-253)
template<class T>
V3<T> mis_sample_bsdf_medium(const Scene<T> &scene, const MediumScatter<T> &ms, const Medium<T> &med,
                             const T s3[3], Sampler<T> &trs, Counters *c) {
    V3<T> dir;
    T f;
    ms.phase.sample(ms.wr, s3[0], s3[1], &dir, &f);
    if (f == 0) return V3<T>(0);
    dir = dir.normalize();
    const Ray<T> nr(ms.pos, dir);
    EntityHit<T> h;
    if (!scene.closest(nr, &h, c)) {
        if (scene.env_kind == AGX_ENV_NONE) return V3<T>(0);
        const V3<T> lf = scene.sky_radiance(nr.d);
        if (lf.is_black()) return V3<T>(0);
        return lf * V3<T>(f) / (f + scene.env_pdf(nr.d));
    }
    if (!scene.entities[size_t(h.entity)].is_light) return V3<T>(0);
    const V3<T> lf = scene.area_radiance(h.entity, h.geometry_coord.z, h.wr);
    if (lf.is_black()) return V3<T>(0);
    const V3<T> ff = med.tr(nr.o, h.pos, trs) * lf * V3<T>(f);
    const T lpdf = scene.area_pdf(h.entity, nr.o, h.pos, h.geometry_coord.z);
    return ff / (f + lpdf);
}

// ------------------------------------------------------------------------------------
// camera  (camera/thin_lens.cpp:34-94; look_at is agz-utils, unpinned:
//          z = normalize(dst - pos), x = normalize(up x z), y = z x x)
// ------------------------------------------------------------------------------------

template<class T> struct Camera {
    V3<T> pos, cx, cy, cz;
    T film_w = 1, film_h = 1, lens_radius = 0, focal = 1;
    void init(const agx_camera_desc &c) {
        const V3<T> p(T(c.pos[0]), T(c.pos[1]), T(c.pos[2]));
        const V3<T> d(T(c.dst[0]), T(c.dst[1]), T(c.dst[2]));
        const V3<T> u(T(c.up[0]), T(c.up[1]), T(c.up[2]));
        focal = T(c.focal_distance);
        film_h = 2 * focal * std::tan(T(c.fov_rad) / 2);
        film_w = T(c.film_aspect) * film_h;
        lens_radius = T(c.lens_radius);
        pos = p;
        cz = (d - p).normalize();
        cx = cross(u, cz).normalize();
        cy = cross(cz, cx);
    }
    Ray<T> sample(T fx, T fy, T au, T av) const {
        const V3<T> film((T(0.5) - fx) * film_w, (fy - T(0.5)) * film_h, focal);
        const V2<T> ds = uniform_on_unit_disk(au, av);
        const V3<T> lens(lens_radius * ds.x, lens_radius * ds.y, 0);
        const V3<T> o = pos + cx * lens.x + cy * lens.y + cz * lens.z;
        const V3<T> l = film - lens;
        const V3<T> dir = (cx * l.x + cy * l.y + cz * l.z).normalize();
        return Ray<T>(o, dir);
    }
};

// ------------------------------------------------------------------------------------
// per-pixel renderer  (core/renderer/perpixel_renderer.cpp:13-193, core/framebuffer.h:49-80,
//                      film_filter/box_filter.cpp:13-23)
// ------------------------------------------------------------------------------------

template<class T> struct Kahan {
    T sum = 0, c = 0;
    void add(T x) {  // Neumaier
        const T t = sum + x;
        if (std::abs(sum) >= std::abs(x)) c += (sum - t) + x; else c += (x - t) + sum;
        sum = t;
    }
    T get() const { return sum + c; }
};

struct RenderOut {
    double *rgb = nullptr, *albedo = nullptr, *normal = nullptr, *denoise = nullptr;
};

// One camera sample of pixel (qx, qy) with stream key `key` (the sampled-grid
// index): its film position and PixelResult (the per-sample unit of render()).
template<class T>
PixelResult<T> trace_sample(const Scene<T> &scene, const agx_render_params &rp, int qx, int qy, uint64_t key,
                            uint64_t sample, float *sx, float *sy) {
    Camera<T> cam;
    cam.init(rp.camera);
    TraceParams tp;
    tp.min_depth = rp.min_depth; tp.max_depth = rp.max_depth; tp.specular_depth = rp.specular_depth;
    tp.cont_prob = rp.cont_prob;
    tp.use_mis = rp.use_mis != 0;
    Sampler<T> sam, trs;
    sam.rng.seed_path(rp.seed, key, sample);
    trs.rng.seed_path(rp.seed ^ TR_STREAM_KEY, key, sample);
    const float fu = sam.rng.uniform(), fv = sam.rng.uniform();
    *sx = float(qx) + fu;
    *sy = float(qy) + fv;
    const T au = sam.u1(), av = sam.u1();
    const Ray<T> ray = cam.sample(T(*sx / float(rp.width)), T(*sy / float(rp.height)), au, av);
    Counters c;
    return trace_path(tp, scene, ray, sam, trs, &c);
}

// GaussianFilter::eval without its normalisation factor, which cancels in
// value / weight (film_filter/gaussian_filter.cpp:60-68)
inline float gaussian_1d(float d, float alpha, float expv) {
    return std::max(0.0f, float(std::exp(-alpha * d * d)) - expv);
}

// One axis of apply_image_filter (framebuffer.h:49-80) for a sample at s and a
// film grid whose pixel range on this axis is [lo, hi] (a render tile, Atrc's
// task_grid_size): the first pixel is max(lo, ceil(s - r - 0.5)), the last
// min(hi, floor(s + r - 0.5)), and the distance handed to the filter starts at
// |first + 0.5 - s| and grows by 1 per pixel (AGZ_SCOPE_EXIT{ rel += 1; }),
// pixels whose running distance exceeds r being skipped.  (The running
// distance is not the distance to the pixel's centre; Atrc's wide filters
// weigh samples by it, and so does this restatement.)  Writes the pixels and
// distances, returns their count.
constexpr int AXIS_CAP = 160;  // > 2 r + 1 for the largest radius accepted (64)
inline int atrc_filter_axis(int lo, int hi, float r, float s, int *pix, float *rel, int cap) {
    const int a = std::max(lo, int(std::ceil(s - r - 0.5f)));
    const int b = std::min(hi, int(std::floor(s + r - 0.5f)));
    float d = std::abs(float(a) + 0.5f - s);
    int n = 0;
    for (int x = a; x <= b; ++x) {
        if (!(d > r) && n < cap) { pix[n] = x; rel[n] = d; ++n; }
        d += 1.0f;
    }
    return n;
}

template<class T>
void render(const Scene<T> &scene, const agx_render_params &rp, int threads, const RenderOut &out,
            Counters *total, int row_begin = 0, int row_end = -1) {
    Camera<T> cam;
    cam.init(rp.camera);
    TraceParams tp;
    tp.min_depth = rp.min_depth; tp.max_depth = rp.max_depth; tp.specular_depth = rp.specular_depth;
    tp.cont_prob = rp.cont_prob;
    tp.use_mis = rp.use_mis != 0;
    const int W = rp.width, H = rp.height;
    if (row_end < 0) row_end = H;
    // BoxFilter radius_ = r - EPS (box_filter.cpp:15-21); GaussianFilter keeps r
    const bool gaussian = rp.filter == AGX_FILTER_GAUSSIAN;
    const float radius_f = gaussian ? rp.filter_radius : rp.filter_radius - float(scene.eps);
    const T radius = T(radius_f);
    if (!(radius > 0)) throw std::runtime_error("invalid film filter radius");
    const float g_exp = gaussian ? float(std::exp(-rp.filter_alpha * radius_f * radius_f)) : 0.0f;
    // Filters wider than half a pixel splat every sample onto the pixels of its
    // tile within the radius, and each tile samples the pixels of
    // sample_bound_of(radius, tile) (framebuffer.h:36-80; render_target.h:384-395).
    // A sample's stream is keyed by its pixel in the whole sampled grid, so the
    // samples that two tiles both take are the same paths.
    const bool splat = gaussian || radius_f > 0.5f;
    const int b0 = splat ? -int(std::floor(0.5f - radius_f)) : 0;
    const int b1 = splat ? int(std::floor(radius_f + 0.49999f)) : 0;
    const int EW = W + b0 + b1;
    // Atrc's render tiles (task_grid_size), aligned to the image: a wide filter
    // splats a sample only within the film grid of the tile that drew it
    const int tile = rp.task_grid_size > 0 ? rp.task_grid_size : 32;
    const int ty_first = row_begin / tile;
    const int tiles_x = (W + tile - 1) / tile, tiles_y = (row_end - 1) / tile - ty_first + 1;
    std::atomic<int> next{0};
    std::vector<Counters> cs(size_t(std::max(1, threads)));
    auto write = [&](int px, int py, Kahan<T> *acc, T weight) {
        const size_t o = size_t(py) * size_t(W) + size_t(px);
        const double inv = weight > 0 ? 1.0 / double(weight) : 1.0;
        for (int k = 0; k < 3; ++k) {
            if (out.rgb) out.rgb[3 * o + k] = double(acc[k].get()) * inv;
            if (out.albedo) out.albedo[3 * o + k] = double(acc[3 + k].get()) * inv;
            if (out.normal) out.normal[3 * o + k] = double(acc[6 + k].get()) * inv;
        }
        if (out.denoise) out.denoise[o] = double(acc[9].get()) * inv;
    };
    auto worker = [&](int tid) {
        Counters &c = cs[size_t(tid)];
        for (;;) {
            const int task = next.fetch_add(1);
            if (task >= tiles_x * tiles_y) break;
            const int tx = task % tiles_x, ty = ty_first + task / tiles_x;
            // the tile's pixels [x0, x1) x [ty0, ty1); rows [y0, y1) of it are rendered
            const int x0 = tx * tile, ty0 = ty * tile;
            const int x1 = std::min(x0 + tile, W), ty1 = std::min(ty0 + tile, H);
            const int y0 = std::max(ty0, row_begin), y1 = std::min(ty1, row_end);
            if (splat) {
                const int tw = x1 - x0, th = ty1 - ty0;
                std::vector<Kahan<T>> acc(size_t(tw) * th * 10);
                std::vector<T> wsum(size_t(tw) * th, T(0));
                const int lx = int(std::floor(float(x0) + 0.5f - radius_f));
                const int ly = int(std::floor(float(ty0) + 0.5f - radius_f));
                const int hx = int(std::floor(float(x1 - 1) + 0.49999f + radius_f));
                const int hy = int(std::floor(float(ty1 - 1) + 0.49999f + radius_f));
                for (int qy = ly; qy <= hy; ++qy)
                    for (int qx = lx; qx <= hx; ++qx) {
                        const uint64_t key = uint64_t(qy + b0) * uint64_t(EW) + uint64_t(qx + b0);
                        for (int s = 0; s < rp.spp; ++s) {
                            Sampler<T> sam, trs;
                            sam.rng.seed_path(rp.seed, key, uint64_t(rp.sample_begin + s));
                            trs.rng.seed_path(rp.seed ^ TR_STREAM_KEY, key, uint64_t(rp.sample_begin + s));
                            const float fu = sam.rng.uniform(), fv = sam.rng.uniform();
                            const float sx = float(qx) + fu, sy = float(qy) + fv;
                            const T au = sam.u1(), av = sam.u1();
                            const Ray<T> ray = cam.sample(T(sx / float(W)), T(sy / float(H)), au, av);
                            const PixelResult<T> pr = trace_path(tp, scene, ray, sam, trs, &c);
                            if (!pr.value.is_finite()) continue;
                            // apply_image_filter over this tile's film grid
                            int px_[AXIS_CAP], py_[AXIS_CAP];
                            float rx_[AXIS_CAP], ry_[AXIS_CAP];
                            const int nx = atrc_filter_axis(x0, x1 - 1, radius_f, sx, px_, rx_, AXIS_CAP);
                            const int ny = atrc_filter_axis(ty0, ty1 - 1, radius_f, sy, py_, ry_, AXIS_CAP);
                            for (int jy = 0; jy < ny; ++jy) {
                                const int y = py_[jy];
                                for (int jx = 0; jx < nx; ++jx) {
                                    const int x = px_[jx];
                                    const T w = gaussian ? T(gaussian_1d(rx_[jx], rp.filter_alpha, g_exp) *
                                                             gaussian_1d(ry_[jy], rp.filter_alpha, g_exp))
                                                         : T(1);
                                    const size_t o = size_t(y - ty0) * size_t(tw) + size_t(x - x0);
                                    Kahan<T> *a = &acc[10 * o];
                                    for (int k = 0; k < 3; ++k) {
                                        a[k].add(w * pr.value[k]); a[3 + k].add(w * pr.albedo[k]);
                                        a[6 + k].add(w * pr.normal[k]);
                                    }
                                    a[9].add(w * pr.denoise);
                                    wsum[o] += w;
                                }
                            }
                        }
                    }
                for (int y = y0; y < y1; ++y)
                    for (int x = x0; x < x1; ++x) {
                        const size_t o = size_t(y - ty0) * size_t(tw) + size_t(x - x0);
                        write(x, y, &acc[10 * o], wsum[o]);
                    }
                continue;
            }
            for (int py = y0; py < y1; ++py)
                for (int px = x0; px < x1; ++px) {
                    Kahan<T> acc[10];
                    T weight = 0;
                    const uint64_t pix = uint64_t(py) * uint64_t(W) + uint64_t(px);
                    for (int s = 0; s < rp.spp; ++s) {
                        Sampler<T> sam, trs;
                        sam.rng.seed_path(rp.seed, pix, uint64_t(rp.sample_begin + s));
                        trs.rng.seed_path(rp.seed ^ TR_STREAM_KEY, pix, uint64_t(rp.sample_begin + s));
                        const float fu = sam.rng.uniform(), fv = sam.rng.uniform();
                        // pixel coordinates are float in Atrc (real = float)
                        const float pxf = float(px) + fu, pyf = float(py) + fv;
                        const T au = sam.u1(), av = sam.u1();
                        const Ray<T> ray = cam.sample(T(pxf / float(W)), T(pyf / float(H)), au, av);
                        const PixelResult<T> pr = trace_path(tp, scene, ray, sam, trs, &c);
                        if (!pr.value.is_finite()) continue;
                        // apply_image_filter with box radius r: a sample reaches only its
                        // own pixel, and only if |px + 0.5 - sample| <= r on both axes
                        const T xr = std::abs(T(float(px) + 0.5f - pxf));
                        const T yr = std::abs(T(float(py) + 0.5f - pyf));
                        if (xr > radius || yr > radius) continue;
                        weight += 1;
                        for (int k = 0; k < 3; ++k) {
                            acc[k].add(pr.value[k]); acc[3 + k].add(pr.albedo[k]); acc[6 + k].add(pr.normal[k]);
                        }
                        acc[9].add(pr.denoise);
                    }
                    write(px, py, acc, weight);
                }
        }
    };
    std::vector<std::thread> pool;
    const int nt = std::max(1, threads);
    for (int i = 1; i < nt; ++i) pool.emplace_back(worker, i);
    worker(0);
    for (auto &t : pool) t.join();
    if (total)
        for (auto &c : cs) {
            total->closest += c.closest; total->shadow += c.shadow; total->node_visits += c.node_visits;
            total->tri_tests += c.tri_tests; total->hits += c.hits;
        }
}

}  // namespace oracle
