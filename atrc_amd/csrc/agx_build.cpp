// Host-side scene construction for the gfx950 path tracer.
//
// The triangle BVH and the entity BVH are built exactly as Atrc builds them
// (same split rules, same node order, same float arithmetic), so that in the
// default mode the GPU traverses the very tree Atrc's CPU `pt` would traverse
// and ray-query results agree bit for bit:
//   triangle BVH: src/tracer/src/core/geometry/triangle_bvh.cpp:122-304, 320-361
//   entity BVH:   src/tracer/src/core/aggregate/entity_bvh.cpp:46-116
// This file is compiled with -ffp-contract=off.
#include "agx_build.h"
#include "agx_bvh4.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <limits>

namespace agx {

// IEEE binary16 bits of f when f converts to FP16 without rounding (zeros,
// normal and subnormal FP16 values), else false.
bool half_exact(float f, uint16_t *out) {
    uint32_t b;
    std::memcpy(&b, &f, 4);
    const uint32_t sign = (b >> 16) & 0x8000u, ex = (b >> 23) & 0xffu, man = b & 0x7fffffu;
    if (ex == 0 && man == 0) { *out = uint16_t(sign); return true; }
    if (ex == 255 || ex == 0) return false;  // inf / NaN, FP32 subnormals (far below FP16's range)
    const int e = int(ex) - 127;
    if (e > 15) return false;
    if (e >= -14) {
        if (man & 0x1fffu) return false;
        *out = uint16_t(sign | (uint32_t(e + 15) << 10) | (man >> 13));
        return true;
    }
    if (e < -24) return false;
    const uint32_t full = 0x800000u | man;
    const int shift = -1 - e;  // FP16 subnormal: value = m 2^-24
    if (full & ((1u << shift) - 1u)) return false;
    *out = uint16_t(sign | (full >> shift));
    return true;
}

namespace {

struct F3 {
    float x, y, z;
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline F3 add(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline F3 sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline F3 mul(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline F3 divs(F3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline float dot3(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline F3 cross3(F3 a, F3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float len3(F3 a) { return std::sqrt(dot3(a, a)); }
inline F3 unit(F3 a) { return divs(a, len3(a)); }
inline F3 neg(F3 a) { return {-a.x, -a.y, -a.z}; }
inline vec4 v4(F3 a, float w) { return vec4{a.x, a.y, a.z, w}; }
inline float bits_f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
inline uint32_t f_bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

struct Box {
    F3 lo{std::numeric_limits<float>::max(), std::numeric_limits<float>::max(), std::numeric_limits<float>::max()};
    F3 hi{std::numeric_limits<float>::lowest(), std::numeric_limits<float>::lowest(), std::numeric_limits<float>::lowest()};
    void add_point(F3 p) {
        lo = {std::min(lo.x, p.x), std::min(lo.y, p.y), std::min(lo.z, p.z)};
        hi = {std::max(hi.x, p.x), std::max(hi.y, p.y), std::max(hi.z, p.z)};
    }
    void add_box(const Box &b) {
        lo = {std::min(lo.x, b.lo.x), std::min(lo.y, b.lo.y), std::min(lo.z, b.lo.z)};
        hi = {std::max(hi.x, b.hi.x), std::max(hi.y, b.hi.y), std::max(hi.z, b.hi.z)};
    }
    // the world_bound() padding Atrc applies to flat geometry (quad.cpp:67-71)
    void pad_flat() {
        float *l = &lo.x, *h = &hi.x;
        for (int i = 0; i < 3; ++i)
            if (l[i] >= h[i]) l[i] = h[i] - 0.1f * std::abs(h[i]);
    }
};

// Duff et al. 2017 orthonormal basis; returns the x axis.
F3 basis_x_from_z(F3 zin) {
    const F3 n = unit(zin);
    const float sign = std::copysign(1.0f, n.z);
    const float a = -1.0f / (sign + n.z);
    const float b = n.x * n.y * a;
    return {1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x};
}

// dpdu_as_ex (triangle_aux.h:76-88)
F3 tangent_x(F3 B_A, F3 C_A, float bu, float bv, float cu, float cv, F3 nor) {
    const float det = bu * cv - bv * cu;
    if (!det) return basis_x_from_z(nor);
    const float inv = 1.0f / det;
    return unit(sub(mul(B_A, cv * inv), mul(C_A, bv * inv)));
}

float tri_area(F3 B_A, F3 C_A) { return len3(cross3(B_A, C_A)) / 2; }

F3 vpos(const agx_vertex &v) { return {v.pos[0], v.pos[1], v.pos[2]}; }
F3 vnor(const agx_vertex &v) { return {v.nor[0], v.nor[1], v.nor[2]}; }

// Vose's alias method: same pairing order as the CPU oracle (small/large stacks,
// last-in first-out), probabilities computed in double and stored as float.
void build_alias(const std::vector<float> &w, std::vector<float> *prob, std::vector<int32_t> *alias) {
    const size_t n = w.size();
    std::vector<float> p(n, 1.0f);
    std::vector<int32_t> al(n);
    double sum = 0;
    for (float x : w) sum += double(x);
    std::vector<double> q(n);
    std::vector<int32_t> small, large;
    for (size_t i = 0; i < n; ++i) {
        q[i] = double(w[i]) * double(n) / sum;
        al[i] = int32_t(i);
        if (q[i] < 1.0) small.push_back(int32_t(i)); else large.push_back(int32_t(i));
    }
    while (!small.empty() && !large.empty()) {
        const int32_t s = small.back();
        small.pop_back();
        const int32_t l = large.back();
        large.pop_back();
        p[size_t(s)] = float(q[size_t(s)]);
        al[size_t(s)] = l;
        q[size_t(l)] = (q[size_t(l)] + q[size_t(s)]) - 1.0;
        if (q[size_t(l)] < 1.0) small.push_back(l); else large.push_back(l);
    }
    for (int32_t i : large) p[size_t(i)] = 1.0f;
    for (int32_t i : small) p[size_t(i)] = 1.0f;
    prob->insert(prob->end(), p.begin(), p.end());
    alias->insert(alias->end(), al.begin(), al.end());
}

}  // namespace

namespace {
// Atrc's Primitive (a, b - a, c - a; the pad carries the original id) and
// PrimitiveInfo (triangle_bvh.cpp:24-35, 533-558) of triangle t
void emit_triangle(const agx_vertex *tris, uint32_t t, std::vector<vec4> *prims, std::vector<vec4> *info) {
    const agx_vertex &va = tris[3 * t], &vb = tris[3 * t + 1], &vc = tris[3 * t + 2];
    const F3 A = vpos(va), B_A = sub(vpos(vb), A), C_A = sub(vpos(vc), A);
    prims->push_back(v4(A, bits_f(t)));
    prims->push_back(v4(B_A, 0.0f));
    prims->push_back(v4(C_A, 0.0f));
    const F3 na = unit(vnor(va)), nb = unit(vnor(vb)), nc = unit(vnor(vc));
    const float tbu = vb.uv[0] - va.uv[0], tbv = vb.uv[1] - va.uv[1];
    const float tcu = vc.uv[0] - va.uv[0], tcv = vc.uv[1] - va.uv[1];
    F3 z = unit(cross3(B_A, C_A));
    if (dot3(add(add(na, nb), nc), z) < 0) z = neg(z);
    const F3 x = tangent_x(B_A, C_A, tbu, tbv, tcu, tcv, z);
    info->push_back(v4(na, va.uv[0]));
    info->push_back(v4(sub(nb, na), va.uv[1]));
    info->push_back(v4(sub(nc, na), tbu));
    info->push_back(v4(x, tbv));
    info->push_back(v4(z, tcu));
    info->push_back(vec4{tcv, 0.0f, 0.0f, 0.0f});
}

// surface area of a mesh, summed in triangle order as build_triangle_bvh does
float mesh_area(const agx_vertex *tris, uint32_t n) {
    float area = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const F3 p0 = vpos(tris[3 * i]), p1 = vpos(tris[3 * i + 1]), p2 = vpos(tris[3 * i + 2]);
        area += tri_area(sub(p1, p0), sub(p2, p0));
    }
    return area;
}
}  // namespace

int build_triangle_bvh(const agx_vertex *tris, uint32_t n, std::vector<vec4> *nodes, std::vector<vec4> *prims,
                       std::vector<vec4> *info, double *surface_area) {
    // triangle references in build order
    struct Ref { uint32_t id; F3 c; };
    std::vector<Ref> refs(n);
    float area = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const F3 p0 = vpos(tris[3 * i]), p1 = vpos(tris[3 * i + 1]), p2 = vpos(tris[3 * i + 2]);
        refs[i].id = i;
        refs[i].c = divs(add(add(p0, p1), p2), 3.0f);
        area += tri_area(sub(p1, p0), sub(p2, p0));
    }
    if (surface_area) *surface_area = double(area);

    // Breadth-first split (triangle_bvh.cpp:122-223).  Each build node records
    // its bound, its triangle range and its two children (or -1).
    struct BNode { Box box; int32_t kid[2]; uint32_t begin, end; };
    std::vector<BNode> tree;
    tree.reserve(size_t(n) / 2 + 16);
    struct Job { int32_t parent; int side; uint32_t begin, end, depth; };
    std::deque<Job> jobs;
    jobs.push_back({-1, 0, 0, n, 0});
    int max_depth = 0;
    while (!jobs.empty()) {
        const Job job = jobs.front();
        jobs.pop_front();
        Box all, cen;
        for (uint32_t i = job.begin; i < job.end; ++i) {
            const uint32_t t = refs[i].id;
            all.add_point(vpos(tris[3 * t]));
            all.add_point(vpos(tris[3 * t + 1]));
            all.add_point(vpos(tris[3 * t + 2]));
            cen.add_point(refs[i].c);
        }
        const int32_t me = int32_t(tree.size());
        tree.push_back({all, {-1, -1}, job.begin, job.end});
        if (job.parent >= 0) tree[size_t(job.parent)].kid[job.side] = me;
        max_depth = std::max(max_depth, int(job.depth));
        const uint32_t count = job.end - job.begin;
        if (count <= 5u) continue;  // leaf threshold
        const F3 ext = sub(cen.hi, cen.lo);
        const int axis = ext.x > ext.y ? (ext.x > ext.z ? 0 : 2) : (ext.y > ext.z ? 1 : 2);
        uint32_t mid = job.begin;
        if (job.depth < 64u) {
            const float split = 0.5f * (cen.hi[axis] + cen.lo[axis]);
            for (uint32_t i = job.begin; i < job.end; ++i)
                if (refs[i].c[axis] < split) std::swap(refs[i], refs[mid++]);
            if (mid == job.begin || mid == job.end) mid = job.begin + count / 2;
        } else {
            std::sort(refs.begin() + job.begin, refs.begin() + job.end,
                      [axis](const Ref &l, const Ref &r) { return l.c[axis] < r.c[axis]; });
            mid = job.begin + count / 2;
        }
        jobs.push_back({me, 0, job.begin, mid, job.depth + 1});
        jobs.push_back({me, 1, mid, job.end, job.depth + 1});
    }

    // Depth-first flattening, left child immediately after its parent
    // (compact_bvh, triangle_bvh.cpp:225-304).
    nodes->clear();
    prims->clear();
    info->clear();
    nodes->reserve(tree.size() * 2);
    prims->reserve(size_t(n) * 3);
    info->reserve(size_t(n) * 6);
    struct Visit { int32_t node; int64_t patch; };
    std::vector<Visit> todo;
    todo.push_back({0, -1});
    uint32_t next_prim = 0;
    while (!todo.empty()) {
        const Visit v = todo.back();
        todo.pop_back();
        const BNode &b = tree[size_t(v.node)];
        const uint32_t index = uint32_t(nodes->size() / 2);
        if (v.patch >= 0) (*nodes)[size_t(v.patch)].w = bits_f(index);
        if (b.kid[0] >= 0 && b.kid[1] >= 0) {
            nodes->push_back(v4(b.box.lo, bits_f(INTERIOR)));
            nodes->push_back(v4(b.box.hi, bits_f(0)));
            todo.push_back({b.kid[1], int64_t(nodes->size()) - 1});
            todo.push_back({b.kid[0], -1});
            continue;
        }
        const uint32_t first = next_prim, last = next_prim + (b.end - b.begin);
        nodes->push_back(v4(b.box.lo, bits_f(first)));
        nodes->push_back(v4(b.box.hi, bits_f(last)));
        for (uint32_t j = b.begin; j < b.end; ++j) emit_triangle(tris, refs[j].id, prims, info);
        next_prim = last;
    }
    return max_depth;
}

namespace {

struct EntRec { int32_t id; Box box; };

struct TlasBuilder {
    std::vector<vec4> *nodes;
    std::vector<int32_t> *prims;
    int max_leaf;
    int max_depth = 0;
    // EntityBVH::build_aux (entity_bvh.cpp:46-116): pre-order, median split by
    // sorting on the bound centre along the largest extent.
    uint32_t build(EntRec *r, size_t count, int depth) {
        max_depth = std::max(max_depth, depth);
        const uint32_t me = uint32_t(nodes->size() / 2);
        if (count <= size_t(max_leaf) || count < 2) {
            Box b;
            const uint32_t first = uint32_t(prims->size());
            for (size_t i = 0; i < count; ++i) { prims->push_back(r[i].id); b.add_box(r[i].box); }
            nodes->push_back(v4(b.lo, bits_f(first | TLAS_LEAF)));
            nodes->push_back(v4(b.hi, bits_f(first + uint32_t(count))));
            return me;
        }
        Box all;
        for (size_t i = 0; i < count; ++i) all.add_box(r[i].box);
        float best = -1;
        int axis = 0;
        for (int i = 0; i < 3; ++i) {
            const float l = all.hi[i] - all.lo[i];
            if (l > best) { best = l; axis = i; }
        }
        std::sort(r, r + count, [axis](const EntRec &a, const EntRec &b) {
            return a.box.lo[axis] + a.box.hi[axis] < b.box.lo[axis] + b.box.hi[axis];
        });
        nodes->push_back(v4(all.lo, 0.0f));
        nodes->push_back(v4(all.hi, 0.0f));
        const size_t half = count / 2;
        const uint32_t l = build(r, half, depth + 1);
        const uint32_t rr = build(r + half, count - half, depth + 1);
        (*nodes)[size_t(me) * 2].w = bits_f(l);
        (*nodes)[size_t(me) * 2 + 1].w = bits_f(rr);
        return me;
    }
};

void set_rows(vec4 *rows, const float *m) {
    for (int r = 0; r < 3; ++r) rows[r] = vec4{m[4 * r], m[4 * r + 1], m[4 * r + 2], m[4 * r + 3]};
}

// Texture2D lookup on the host, the same FP32 arithmetic as tex2d_env on the device
float wrap2(int mode, float x) {
    if (mode == AGX_WRAP_REPEAT) return std::min(1.0f, std::max(0.0f, x - std::floor(x)));
    if (mode == AGX_WRAP_MIRROR) {
        const int ip = int(x >= 0 ? x : std::floor(x));
        const float v = (ip & 1) ? 1 - x + float(ip) : x - float(ip);
        return std::min(1.0f, std::max(0.0f, v));
    }
    return std::min(1.0f, std::max(0.0f, x));
}
F3 tex2d_host(const DevTex2D &t, const std::vector<float> &texels, float u0, float v0) {
    const float qu = t.xf[0].x * u0 + t.xf[0].y * v0 + t.xf[0].z;
    const float qv = t.xf[1].x * u0 + t.xf[1].y * v0 + t.xf[1].z;
    const float u = wrap2(t.wrap[0], qu), v = wrap2(t.wrap[1], qv);
    auto texel = [&](int x, int y) {
        const float *g = texels.data() + (size_t(y) * size_t(t.w) + size_t(x)) * 3;
        return F3{g[0], g[1], g[2]};
    };
    F3 r;
    if (t.kind == AGX_TEX2D_CONSTANT) {
        r = F3{t.value.x, t.value.y, t.value.z};
    } else if (t.kind == AGX_TEX2D_CHECKER) {
        const float gs = 1.0f / float(t.grid_count);
        const int ui = int(std::min(1.0f, std::max(0.0f, u)) / gs), vi = int(std::min(1.0f, std::max(0.0f, v)) / gs);
        r = ((ui + vi) & 1) ? F3{t.value.x, t.value.y, t.value.z} : F3{t.value2.x, t.value2.y, t.value2.z};
    } else if (!t.linear) {
        r = texel(std::min(t.w - 1, int(u * float(t.w))), std::min(t.h - 1, int(v * float(t.h))));
    } else {
        const float fu = u * float(t.w) - 0.5f, fv = v * float(t.h) - 0.5f;
        const float bu = std::floor(fu), bv = std::floor(fv);
        const float du = fu - bu, dv = fv - bv;
        const int x0 = std::min(t.w - 1, std::max(0, int(bu))), x1 = std::min(t.w - 1, std::max(0, int(bu) + 1));
        const int y0 = std::min(t.h - 1, std::max(0, int(bv))), y1 = std::min(t.h - 1, std::max(0, int(bv) + 1));
        auto lerp = [](F3 a, F3 b, float k) { return add(a, mul(sub(b, a), k)); };
        r = lerp(lerp(texel(x0, y0), texel(x1, y0), du), lerp(texel(x0, y1), texel(x1, y1), du), dv);
    }
    if (t.inv_gamma != 1.0f) r = F3{std::pow(r.x, t.inv_gamma), std::pow(r.y, t.inv_gamma), std::pow(r.z, t.inv_gamma)};
    return r;
}

// IBL + EnvironmentLightSampler construction (ibl.cpp:24-60, env_sampler.h:17-110):
// patches of at most 200 x 200 over the texture, each weighted by the summed
// luminance of the texels around it (one texel margin) times its solid angle,
// normalised, then an alias table.  no_importance_sampling samples a constant
// 1x1 texture instead: one patch, uniform over the sphere.
int build_ibl(const agx_env_desc &env, HostScene *S, std::string *err) {
    const agx_texture2d_desc &t = env.tex;
    DevTex2D &o = S->env_tex;
    o = DevTex2D{};
    o.kind = t.kind;
    o.linear = t.linear;
    o.grid_count = t.grid_count;
    o.inv_gamma = t.inv_gamma > 0 ? t.inv_gamma : 1.0f;
    o.wrap[0] = t.wrap[0];
    o.wrap[1] = t.wrap[1];
    o.value = vec4{t.value[0], t.value[1], t.value[2], 0};
    o.value2 = vec4{t.value2[0], t.value2[1], t.value2[2], 0};
    o.xf[0] = vec4{t.transform[0], t.transform[1], t.transform[2], 0};
    o.xf[1] = vec4{t.transform[3], t.transform[4], t.transform[5], 0};
    for (int k = 0; k < 2; ++k)
        if (t.wrap[k] < AGX_WRAP_CLAMP || t.wrap[k] > AGX_WRAP_MIRROR) { *err = "ibl: invalid wrap mode"; return AGX_E_INVALID_ARG; }
    if (t.kind == AGX_TEX2D_IMAGE) {
        if (!t.data || t.width <= 0 || t.height <= 0 || int64_t(t.width) * t.height > (int64_t(1) << 28)) {
            *err = "ibl: invalid image texture";
            return AGX_E_INVALID_ARG;
        }
        o.w = t.width;
        o.h = t.height;
        S->env_texels.assign(t.data, t.data + size_t(t.width) * size_t(t.height) * 3);
    } else if (t.kind == AGX_TEX2D_CONSTANT || t.kind == AGX_TEX2D_CHECKER) {
        if (t.kind == AGX_TEX2D_CHECKER && t.grid_count <= 0) { *err = "ibl: invalid grid_count value"; return AGX_E_INVALID_ARG; }
        o.w = o.h = 1;
        S->env_texels.assign(3, 0.0f);
    } else {
        *err = "ibl: unknown texture kind";
        return AGX_E_INVALID_ARG;
    }
    const bool uniform = env.no_importance_sampling != 0;
    const int width = uniform ? 1 : o.w, height = uniform ? 1 : o.h;
    const int nw = std::min(width, 200), nh = std::min(height, 200);
    S->env_pw = nw;
    S->env_ph = nh;
    S->env_probs.assign(size_t(nw) * nh, 0.0f);
    const float pi = 3.14159265358979323846f;
    float lum_sum = 0;
    for (int y = 0; y < nh; ++y) {
        const float y0 = float(y) / float(nh), y1 = float(y + 1) / float(nh);
        const int ys = std::max(0, int(std::floor(y0 * float(height)) - 1));
        const int yl = std::min(height - 1, int(std::floor(y1 * float(height)) + 1));
        for (int x = 0; x < nw; ++x) {
            const float x0 = float(x) / float(nw), x1 = float(x + 1) / float(nw);
            const int xs = std::max(0, int(std::floor(x0 * float(width)) - 1));
            const int xl = std::min(width - 1, int(std::floor(x1 * float(width)) + 1));
            float lum = 0;
            for (int yy = ys; yy <= yl; ++yy)
                for (int xx = xs; xx <= xl; ++xx) {
                    const float su = (float(xx) + 0.5f) / float(width), sv = (float(yy) + 0.5f) / float(height);
                    const F3 c = uniform ? F3{1, 1, 1} : tex2d_host(o, S->env_texels, su, sv);
                    lum += 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z;
                }
            const float area = std::abs(2 * pi * (x1 - x0) * (std::cos(pi * y1) - std::cos(pi * y0)));
            S->env_probs[size_t(y) * nw + x] = lum * area;
            lum_sum += lum * area;
        }
    }
    if (lum_sum > 0)
        for (float &p : S->env_probs) p *= 1 / lum_sum;
    else
        std::fill(S->env_probs.begin(), S->env_probs.end(), 1.0f / float(S->env_probs.size()));
    S->env_alias_p.clear();
    S->env_alias_i.clear();
    build_alias(S->env_probs, &S->env_alias_p, &S->env_alias_i);
    return AGX_OK;
}

F3 apply_point(const float *m, F3 p) {
    return {m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3], m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
            m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]};
}
F3 apply_vec(const float *m, F3 v) {
    return {m[0] * v.x + m[1] * v.y + m[2] * v.z, m[4] * v.x + m[5] * v.y + m[6] * v.z,
            m[8] * v.x + m[9] * v.y + m[10] * v.z};
}

}  // namespace

int build_host_scene(const agx_scene_desc &desc, HostScene *out, std::string *err) {
    HostScene &S = *out;
    S = HostScene();
    if (desc.abi_version != AGX_ABI_VERSION) { *err = "scene descriptor ABI version mismatch"; return AGX_E_INVALID_ARG; }
    if (desc.entity_count && !desc.entities) { *err = "entities is NULL"; return AGX_E_INVALID_ARG; }
    if (desc.material_count && !desc.materials) { *err = "materials is NULL"; return AGX_E_INVALID_ARG; }

    // materials: derive what Material::shade derives per hit
    for (uint32_t i = 0; i < desc.material_count; ++i) {
        const agx_material_desc &m = desc.materials[i];
        DevMaterial d;
        std::memset(&d, 0, sizeof(d));
        d.kind = m.kind;
        const F3 col{m.color[0], m.color[1], m.color[2]};
        const float PI = 3.14159265358979323846f;
        switch (m.kind) {
        case AGX_MAT_IDEAL_DIFFUSE:  // ideal_diffuse.cpp:25-39, diffuse_comp.cpp:5-8
            d.albedo = v4(col, 0); d.coef = v4(divs(col, PI), 0);
            break;
        case AGX_MAT_METAL: {        // metal.cpp + material_creators.cpp:252-266 (eta/k swap)
            d.albedo = v4(col, 0); d.color = v4(col, 0);
            d.eta2 = vec4{m.k[0] * m.k[0], m.k[1] * m.k[1], m.k[2] * m.k[2], 0};
            d.k2 = vec4{m.eta[0] * m.eta[0], m.eta[1] * m.eta[1], m.eta[2] * m.eta[2], 0};
            const float an = m.anisotropic, r2 = m.roughness * m.roughness;
            const float aspect = an > 0 ? std::sqrt(1 - 0.9f * an) : 1.0f;   // microfacet_refl_comp.cpp:72-75
            d.misc = vec4{std::max(0.001f, r2 / aspect), std::max(0.001f, r2 * aspect), 0, 0};
            break;
        }
        case AGX_MAT_GLASS: {        // glass.cpp:180-206
            if (!(m.ior > 0)) { *err = "glass: invalid ior"; return AGX_E_INVALID_ARG; }
            const F3 c2 = m.glass_separate_refraction ? F3{m.color2[0], m.color2[1], m.color2[2]} : col;
            const F3 refr{std::sqrt(c2.x), std::sqrt(c2.y), std::sqrt(c2.z)};
            d.c_refl = v4(col, 0); d.c_refr = v4(refr, 0);
            d.albedo = v4(mul(add(col, refr), 0.5f), 0);
            d.misc = vec4{0, 0, m.ior, 0};
            break;
        }
        case AGX_MAT_MIRROR: {       // mirror.cpp:98-115: ConductorPoint(eta_out = eta, eta_in = 1, k)
            d.albedo = v4(col, 0);
            float e2[3], k2[3];
            for (int c = 0; c < 3; ++c) {
                const float inv = 1 / m.eta[c];
                e2[c] = inv * inv;
                const float kk = m.k[c] / m.eta[c];
                k2[c] = kk * kk;
            }
            d.eta2 = vec4{e2[0], e2[1], e2[2], 0};
            d.k2 = vec4{k2[0], k2[1], k2[2], 0};
            break;
        }
        case AGX_MAT_IDEAL_BLACK:
            break;
        case AGX_MAT_INVISIBLE_SURFACE:  // invisible_surface.cpp:47-50: albedo 1
            d.albedo = vec4{1, 1, 1, 0};
            break;
        case AGX_MAT_DISNEY: {       // DisneyBSDF::DisneyBSDF (disney.cpp:405-457)
            d.albedo = v4(col, 0);
            d.color = v4(col, 0);
            // Spectrum::lum: Rec. 709 luminance; to_tint (disney.cpp:72-76)
            const float lum = 0.2126f * col.x + 0.7152f * col.y + 0.0722f * col.z;
            d.c_refl = lum > 0 ? vec4{col.x / lum, col.y / lum, col.z / lum, 0} : vec4{1, 1, 1, 0};
            d.c_refr = vec4{m.color2[0], m.color2[1], m.color2[2], 0};
            const float ior = std::max(1.01f, m.ior);
            const float an = m.anisotropic;
            const float aspect = an > 0 ? std::sqrt(1 - 0.9f * an) : 1.0f;
            const float r = m.roughness, tr = m.transmission_roughness;
            d.misc = vec4{std::max(0.001f, r * r / aspect), std::max(0.001f, r * r * aspect), ior, r};
            const float g = m.clearcoat_gloss;
            float ccr = 0.1f * (1 - g) + 0.0f * g;  // mix(0.1, 0, gloss)
            ccr *= ccr;
            ccr = std::max(ccr, 0.0001f);
            const float A = std::min(0.7f, std::max(0.3f, lum * (1 - m.metallic)));
            const float B = 1 - A;
            d.d0 = vec4{m.metallic, m.specular_tint, m.sheen, m.sheen_tint};
            d.d1 = vec4{m.clearcoat, m.transmission, std::max(0.001f, tr * tr / aspect),
                        std::max(0.001f, tr * tr * aspect)};
            d.d2 = vec4{ccr, A * (1 - m.transmission), A * m.transmission, B * 2 / (2 + m.clearcoat)};
            d.d3 = vec4{B * m.clearcoat / (2 + m.clearcoat), 0, 0, 0};
            break;
        }
        default:
            *err = "unsupported material kind " + std::to_string(m.kind);
            return AGX_E_UNSUPPORTED;
        }
        S.mats.push_back(d);
    }

    // media (medium_creators.cpp; homogeneous.cpp:23-41; heterogeneous.cpp:27-45)
    if (desc.medium_count && !desc.media) { *err = "media is NULL"; return AGX_E_INVALID_ARG; }
    if (desc.medium_count > 4096) { *err = "at most 4096 media"; return AGX_E_LIMIT; }
    for (uint32_t i = 0; i < desc.medium_count; ++i) {
        const agx_medium_desc &m = desc.media[i];
        DevMedium dm;
        std::memset(&dm, 0, sizeof(dm));
        dm.kind = m.kind;
        dm.max_scat = m.max_scattering_count;
        dm.white_for_indirect = m.white_for_indirect;
        if (m.kind == AGX_MEDIUM_HOMOGENEOUS) {
            if (!(m.g > -1 && m.g < 1)) { *err = "homogeneous medium: invalid g value"; return AGX_E_INVALID_ARG; }
            dm.sa = vec4{m.sigma_a[0], m.sigma_a[1], m.sigma_a[2], 0};
            dm.ss = vec4{m.sigma_s[0], m.sigma_s[1], m.sigma_s[2], 0};
            dm.st = vec4{m.sigma_a[0] + m.sigma_s[0], m.sigma_a[1] + m.sigma_s[1], m.sigma_a[2] + m.sigma_s[2], 0};
            dm.misc = vec4{m.g, 0, 0, 0};
        } else if (m.kind == AGX_MEDIUM_HETEROGENEOUS) {
            set_rows(dm.w2l, m.world_to_local);
            auto tex = [&](const agx_texture3d_desc &t, DevTex3D *o, float *max_real) -> bool {
                o->kind = t.kind;
                o->linear = t.linear;
                o->inv_gamma = t.inv_gamma > 0 ? t.inv_gamma : 1.0f;
                for (int k = 0; k < 3; ++k) o->wrap[k] = t.wrap[k];
                o->value = vec4{t.value[0], t.value[1], t.value[2], 0};
                set_rows(o->xf, t.transform);
                // max_real is taken over the stored texels, before inv_gamma
                // (image3d.cpp:108-147); a constant texture ignores inv_gamma (constant3d.cpp:43-58)
                if (t.kind == AGX_TEX3D_CONSTANT) {
                    o->w = o->h = o->d = o->ch = 1;
                    if (max_real) *max_real = t.value[0];
                    return true;
                }
                if (t.kind != AGX_TEX3D_GRID || !t.data || t.width <= 0 || t.height <= 0 || t.depth <= 0 ||
                    (t.channels != 1 && t.channels != 3))
                    return false;
                o->w = t.width; o->h = t.height; o->d = t.depth; o->ch = t.channels;
                const size_t n = size_t(t.width) * size_t(t.height) * size_t(t.depth) * size_t(t.channels);
                // a grid whose every texel FP16 holds exactly is stored in half the
                // bytes; lookups widen it back to the same FP32 values
                std::vector<uint16_t> h(n);
                bool exact = true;
                for (size_t k = 0; k < n && exact; ++k) exact = half_exact(t.data[k], &h[k]);
                o->f16 = exact ? 1 : 0;
                if (exact) {
                    o->offset = uint32_t(S.grids16.size());
                    S.grids16.insert(S.grids16.end(), h.begin(), h.end());
                } else {
                    o->offset = uint32_t(S.grids.size());
                    S.grids.insert(S.grids.end(), t.data, t.data + n);
                }
                if (max_real) {
                    float mx = -std::numeric_limits<float>::max();
                    for (size_t k = 0; k < n; k += size_t(t.channels)) mx = std::max(mx, t.data[k]);
                    *max_real = mx;
                }
                return true;
            };
            float maxd = 0;
            if (!tex(m.density, &dm.density, &maxd) || !tex(m.albedo, &dm.albedo, nullptr) || !tex(m.g3, &dm.g3, nullptr)) {
                *err = "heterogeneous medium: invalid 3d texture";
                return AGX_E_INVALID_ARG;
            }
            maxd = std::max(maxd, 3e-4f);
            dm.misc = vec4{0, maxd, 1 / maxd, 0};
        } else if (m.kind != AGX_MEDIUM_VOID) {
            *err = "unknown medium kind";
            return AGX_E_INVALID_ARG;
        }
        S.media.push_back(dm);
    }
    if (S.grids.size() > size_t(0xffffffffu) || S.grids16.size() > size_t(0xffffffffu)) {
        *err = "3d grids too large";
        return AGX_E_LIMIT;
    }

    std::vector<Box> bounds(desc.entity_count);
    {
        // the triangle records of every mesh, reserved once
        size_t mesh_tris = 0;
        for (uint32_t i = 0; i < desc.entity_count; ++i)
            if (desc.entities[i].geometry == AGX_GEOM_TRIANGLE_BVH) mesh_tris += desc.entities[i].tri_count;
        S.blas_prims.reserve(3 * mesh_tris);
        S.blas_info.reserve(6 * mesh_tris);
    }
    for (uint32_t i = 0; i < desc.entity_count; ++i) {
        const agx_entity_desc &e = desc.entities[i];
        DevEntity d;
        std::memset(&d, 0, sizeof(d));
        d.kind = e.geometry;
        d.material = e.material;
        d.light = -1;
        d.flags = (e.no_denoise ? EF_NO_DENOISE : 0) | (e.double_sided ? EF_DOUBLE_SIDED : 0);
        if (e.material < 0 || uint32_t(e.material) >= desc.material_count) {
            *err = "entity " + std::to_string(i) + ": material index out of range";
            return AGX_E_INVALID_ARG;
        }
        if (e.medium_in >= int32_t(desc.medium_count) || e.medium_out >= int32_t(desc.medium_count)) {
            *err = "entity " + std::to_string(i) + ": medium index out of range";
            return AGX_E_INVALID_ARG;
        }
        d.med_in = e.medium_in < 0 ? -1 : e.medium_in;
        d.med_out = e.medium_out < 0 ? -1 : e.medium_out;
        set_rows(d.l2w, e.local_to_world);
        set_rows(d.w2l, e.world_to_local);
        const float ratio = len3(apply_vec(e.local_to_world, F3{1, 0, 0}));
        auto P = [&](int k) { return F3{e.p[k][0], e.p[k][1], e.p[k][2]}; };
        Box box;
        float area = 0;
        switch (e.geometry) {
        case AGX_GEOM_QUAD: {                               // quad.cpp:138-164
            const F3 a = P(0), ba = sub(P(1), a), ca = sub(P(2), a), da = sub(P(3), a);
            const float tbu = e.t[1][0] - e.t[0][0], tbv = e.t[1][1] - e.t[0][1];
            const float tcu = e.t[2][0] - e.t[0][0], tcv = e.t[2][1] - e.t[0][1];
            const float tdu = e.t[3][0] - e.t[0][0], tdv = e.t[3][1] - e.t[0][1];
            const F3 z = unit(cross3(ba, ca));
            d.a = v4(a, 0); d.b_a = v4(ba, 0); d.c_a = v4(ca, 0); d.d_a = v4(da, 0);
            d.z = v4(z, 0);
            d.x_abc = v4(tangent_x(ba, ca, tbu, tbv, tcu, tcv, z), 0);
            d.x_acd = v4(tangent_x(ca, da, tcu, tcv, tdu, tdv, z), 0);
            d.t_ab = vec4{e.t[0][0], e.t[0][1], tbu, tbv};
            d.t_cd = vec4{tcu, tcv, tdu, tdv};
            const float a1 = tri_area(ba, ca), a2 = tri_area(ca, da);
            area = a1 + a2;
            d.sample_prob = vec4{a1 / area, 0, 0, 0};
            box.add_point(a); box.add_point(add(a, ba)); box.add_point(add(a, ca)); box.add_point(add(a, da));
            box.pad_flat();
            break;
        }
        case AGX_GEOM_TRIANGLE: {                           // triangle.cpp:107-125
            const F3 a = P(0), ba = sub(P(1), a), ca = sub(P(2), a);
            const float tbu = e.t[1][0] - e.t[0][0], tbv = e.t[1][1] - e.t[0][1];
            const float tcu = e.t[2][0] - e.t[0][0], tcv = e.t[2][1] - e.t[0][1];
            const F3 z = unit(cross3(ba, ca));
            d.a = v4(a, 0); d.b_a = v4(ba, 0); d.c_a = v4(ca, 0);
            d.z = v4(z, 0);
            d.x_abc = v4(tangent_x(ba, ca, tbu, tbv, tcu, tcv, z), 0);
            d.t_ab = vec4{e.t[0][0], e.t[0][1], tbu, tbv};
            d.t_cd = vec4{tcu, tcv, 0, 0};
            area = tri_area(apply_vec(e.local_to_world, ba), apply_vec(e.local_to_world, ca));
            box.add_point(apply_point(e.local_to_world, a));
            box.add_point(apply_point(e.local_to_world, add(a, ba)));
            box.add_point(apply_point(e.local_to_world, add(a, ca)));
            box.pad_flat();
            break;
        }
        case AGX_GEOM_SPHERE: {                             // sphere.cpp:14-24, 60-72
            if (!(e.radius > 0)) { *err = "invalid sphere radius"; return AGX_E_INVALID_ARG; }
            const float wr = ratio * e.radius;
            d.a = vec4{e.radius, e.radius * e.radius, wr, ratio};
            area = 4 * 3.14159265358979323846f * wr * wr;
            const F3 c = apply_point(e.local_to_world, F3{0, 0, 0});
            const float pad = wr + 3e-4f;
            box.lo = sub(c, F3{pad, pad, pad});
            box.hi = add(c, F3{pad, pad, pad});
            break;
        }
        case AGX_GEOM_DISK: {                               // disk.cpp:33-47, 85-98
            if (!(e.radius > 0)) { *err = "invalid disk radius"; return AGX_E_INVALID_ARG; }
            d.a = vec4{e.radius, e.radius * e.radius, 0, ratio};
            area = ratio * ratio * 3.14159265358979323846f * (e.radius * e.radius);
            const float vs = std::min(0.1f * e.radius, 0.01f);
            for (int k = 0; k < 8; ++k)
                box.add_point(apply_point(e.local_to_world, F3{(k & 1) ? e.radius : -e.radius,
                                                               (k & 2) ? e.radius : -e.radius, (k & 4) ? vs : -vs}));
            break;
        }
        case AGX_GEOM_TRIANGLE_BVH: {
            if (!e.tris || !e.tri_count) { *err = "entity " + std::to_string(i) + ": empty triangle mesh"; return AGX_E_INVALID_ARG; }
            double sa = 0;
            const bool emissive = e.emit_radiance[0] != 0 || e.emit_radiance[1] != 0 || e.emit_radiance[2] != 0;
            const uint32_t node_base = uint32_t(S.blas_nodes.size() / 2);
            const uint32_t prim_base = uint32_t(S.blas_prims.size() / 3);
            if (desc.accel != AGX_ACCEL_ATRC && !emissive) {
                // the BVH4 modes traverse their own tree: Atrc's triangle BVH is
                // needed only for the primitive order its light sampling uses, so
                // a mesh that emits nothing keeps its triangles in input order,
                // written straight into the scene's arrays
                for (uint32_t t = 0; t < e.tri_count; ++t) emit_triangle(e.tris, t, &S.blas_prims, &S.blas_info);
                sa = double(mesh_area(e.tris, e.tri_count));
            } else {
                std::vector<vec4> nodes, prims, info;
                const int depth = build_triangle_bvh(e.tris, e.tri_count, &nodes, &prims, &info, &sa);
                S.max_blas_depth = std::max(S.max_blas_depth, depth);
                // rebase indices to the global arrays
                for (size_t k = 0; k < nodes.size(); k += 2) {
                    uint32_t s, r;
                    std::memcpy(&s, &nodes[k].w, 4);
                    std::memcpy(&r, &nodes[k + 1].w, 4);
                    if (s == INTERIOR) r += node_base;
                    else { s += prim_base; r += prim_base; }
                    nodes[k].w = bits_f(s);
                    nodes[k + 1].w = bits_f(r);
                }
                S.blas_nodes.insert(S.blas_nodes.end(), nodes.begin(), nodes.end());
                S.blas_prims.insert(S.blas_prims.end(), prims.begin(), prims.end());
                S.blas_info.insert(S.blas_info.end(), info.begin(), info.end());
            }
            d.node_base = node_base;
            d.prim_base = prim_base;
            d.prim_count = e.tri_count;
            area = float(sa);
            // the mesh's records (a, b - a, c - a) in the scene's array
            const vec4 *prims = S.blas_prims.data() + 3 * size_t(prim_base);
            // world bound of TriangleBVH (triangle_bvh.cpp:570-583)
            for (size_t k = 0; k < 3 * size_t(e.tri_count); k += 3) {
                const F3 a{prims[k].x, prims[k].y, prims[k].z};
                box.add_point(a);
                box.add_point(add(a, F3{prims[k + 1].x, prims[k + 1].y, prims[k + 1].z}));
                box.add_point(add(a, F3{prims[k + 2].x, prims[k + 2].y, prims[k + 2].z}));
            }
            box.pad_flat();
            if (emissive) {
                std::vector<float> areas(e.tri_count);
                for (uint32_t k = 0; k < e.tri_count; ++k) {
                    const vec4 &b = prims[3 * k + 1], &c = prims[3 * k + 2];
                    areas[k] = tri_area(F3{b.x, b.y, b.z}, F3{c.x, c.y, c.z});
                }
                d.alias_base = uint32_t(S.alias_prob.size());
                build_alias(areas, &S.alias_prob, &S.alias_idx);
            }
            S.total_tris += e.tri_count;
            break;
        }
        default:
            *err = "entity " + std::to_string(i) + ": unsupported geometry kind " + std::to_string(e.geometry);
            return AGX_E_UNSUPPORTED;
        }
        d.emit = vec4{e.emit_radiance[0], e.emit_radiance[1], e.emit_radiance[2], area};
        bounds[i] = box;
        S.ents.push_back(d);
    }

    // light list: environment first, then emissive entities (scene_default.cpp:66-82)
    if (desc.env.kind != AGX_ENV_NONE) {
        if (desc.env.kind != AGX_ENV_NATIVE_SKY && desc.env.kind != AGX_ENV_IBL) {
            *err = "unsupported environment light";
            return AGX_E_UNSUPPORTED;
        }
        if (desc.env.kind == AGX_ENV_IBL) {
            const int r = build_ibl(desc.env, &S, err);
            if (r != AGX_OK) return r;
        }
        S.lights.push_back(-1);
    }
    for (uint32_t i = 0; i < desc.entity_count; ++i) {
        const float *er = desc.entities[i].emit_radiance;
        if (er[0] != 0 || er[1] != 0 || er[2] != 0) {
            S.ents[i].light = int32_t(S.lights.size());
            S.lights.push_back(int32_t(i));
        }
    }
    if (S.lights.size() > size_t(MAX_LIGHTS)) {
        *err = "gpu_pt supports at most " + std::to_string(MAX_LIGHTS) + " lights (scene has " +
               std::to_string(S.lights.size()) + ")";
        return AGX_E_UNSUPPORTED;
    }
    S.env_kind = desc.env.kind;
    S.sky_top = vec4{desc.env.top[0], desc.env.top[1], desc.env.top[2], 0};
    S.sky_bottom = vec4{desc.env.bottom[0], desc.env.bottom[1], desc.env.bottom[2], 0};

    // EnvirLight::preprocess (core/core/light.cpp:6-14)
    Box world;
    for (const Box &b : bounds) world.add_box(b);
    const float diag = len3(sub(world.hi, world.lo));
    const F3 centre = mul(add(world.lo, world.hi), 0.5f);
    S.world = vec4{centre.x, centre.y, centre.z, 1.1f * diag / 2};

    S.accel = desc.accel;
    if (desc.accel == AGX_ACCEL_BVH4 || desc.accel == AGX_ACCEL_BVH4_WATERTIGHT) {
        const bool watertight = desc.accel == AGX_ACCEL_BVH4_WATERTIGHT;
        // one reference per mesh triangle (in the mesh BVH's slot order) and per analytic entity
        std::vector<PrimRef> refs;
        std::vector<uint32_t> ref_ent, ref_slot;
        refs.reserve(size_t(S.total_tris) + desc.entity_count);
        ref_ent.reserve(refs.capacity());
        ref_slot.reserve(refs.capacity());
        for (uint32_t i = 0; i < desc.entity_count; ++i) {
            const DevEntity &d = S.ents[i];
            if (d.kind == AGX_GEOM_TRIANGLE_BVH) {
                for (uint32_t k = 0; k < d.prim_count; ++k) {
                    const uint32_t slot = d.prim_base + k;
                    const vec4 &a = S.blas_prims[3 * size_t(slot)], &ba = S.blas_prims[3 * size_t(slot) + 1],
                               &ca = S.blas_prims[3 * size_t(slot) + 2];
                    const float xs[3] = {a.x, a.x + ba.x, a.x + ca.x}, ys[3] = {a.y, a.y + ba.y, a.y + ca.y},
                                zs[3] = {a.z, a.z + ba.z, a.z + ca.z};
                    PrimRef r;
                    r.lo[0] = std::min({xs[0], xs[1], xs[2]}); r.hi[0] = std::max({xs[0], xs[1], xs[2]});
                    r.lo[1] = std::min({ys[0], ys[1], ys[2]}); r.hi[1] = std::max({ys[0], ys[1], ys[2]});
                    r.lo[2] = std::min({zs[0], zs[1], zs[2]}); r.hi[2] = std::max({zs[0], zs[1], zs[2]});
                    refs.push_back(r);
                    ref_ent.push_back(i);
                    ref_slot.push_back(slot);
                }
            } else if (watertight && d.kind == AGX_GEOM_QUAD) {
                // the quad's two triangles over its own corners: quads that share
                // an edge then see bit-identical endpoints
                const float(*P)[3] = desc.entities[i].p;
                const int halves[2][3] = {{0, 1, 2}, {0, 2, 3}};
                for (int hf = 0; hf < 2; ++hf) {
                    PrimRef r;
                    for (int a = 0; a < 3; ++a) {
                        r.lo[a] = std::min({P[halves[hf][0]][a], P[halves[hf][1]][a], P[halves[hf][2]][a]});
                        r.hi[a] = std::max({P[halves[hf][0]][a], P[halves[hf][1]][a], P[halves[hf][2]][a]});
                    }
                    refs.push_back(r);
                    ref_ent.push_back(i);
                    ref_slot.push_back(hf == 0 ? BVH4_QUAD_ABC : BVH4_QUAD_ACD);
                }
            } else {
                PrimRef r;
                r.lo[0] = bounds[i].lo.x; r.lo[1] = bounds[i].lo.y; r.lo[2] = bounds[i].lo.z;
                r.hi[0] = bounds[i].hi.x; r.hi[1] = bounds[i].hi.y; r.hi[2] = bounds[i].hi.z;
                refs.push_back(r);
                ref_ent.push_back(i);
                ref_slot.push_back(BVH4_ANALYTIC);
            }
        }
        if (refs.empty()) { *err = "scene has no primitives"; return AGX_E_INVALID_ARG; }
        Bvh4 b4;
        try {
            build_bvh4(refs, 4, &b4);
        } catch (const std::exception &ex) {
            *err = std::string("BVH4 build: ") + ex.what();
            return AGX_E_LIMIT;
        }
        S.b4_nodes = std::move(b4.nodes);
        S.b4_prims.resize(3 * refs.size());
        for (size_t k = 0; k < b4.order.size(); ++k) {
            const uint32_t r = b4.order[k];
            vec4 *o = &S.b4_prims[3 * k];
            const uint32_t slot = ref_slot[r];
            if (slot == BVH4_QUAD_ABC || slot == BVH4_QUAD_ACD) {
                const float(*P)[3] = desc.entities[ref_ent[r]].p;
                const int v1 = slot == BVH4_QUAD_ABC ? 1 : 2, v2 = slot == BVH4_QUAD_ABC ? 2 : 3;
                o[0] = vec4{P[0][0], P[0][1], P[0][2], 0};
                o[1] = vec4{P[v1][0], P[v1][1], P[v1][2], 0};
                o[2] = vec4{P[v2][0], P[v2][1], P[v2][2], 0};
            } else if (slot != BVH4_ANALYTIC && watertight) {
                // the mesh's own vertices, so that triangles sharing an edge see
                // bit-identical endpoints
                const uint32_t id = f_bits(S.blas_prims[3 * size_t(slot)].w);
                const agx_vertex *t = desc.entities[ref_ent[r]].tris + 3 * size_t(id);
                o[0] = vec4{t[0].pos[0], t[0].pos[1], t[0].pos[2], 0};
                o[1] = vec4{t[1].pos[0], t[1].pos[1], t[1].pos[2], 0};
                o[2] = vec4{t[2].pos[0], t[2].pos[1], t[2].pos[2], 0};
            } else if (slot != BVH4_ANALYTIC) {
                o[0] = S.blas_prims[3 * size_t(slot)];
                o[1] = S.blas_prims[3 * size_t(slot) + 1];
                o[2] = S.blas_prims[3 * size_t(slot) + 2];
            } else {
                o[0] = o[1] = o[2] = vec4{0, 0, 0, 0};
            }
            o[0].w = bits_f(ref_ent[r]);
            o[1].w = bits_f(slot);
            o[2].w = 0;  // the entity's tie rank, set once the aggregate is built
        }
        S.b4_depth = b4.depth;
        S.b4_leaves = b4.leaves;
        S.b4_sah = b4.sah;
    } else if (desc.accel != AGX_ACCEL_ATRC) {
        *err = "unknown acceleration structure kind";
        return AGX_E_INVALID_ARG;
    }

    S.aggregate = desc.aggregate;
    if (desc.aggregate == AGX_AGGREGATE_BVH) {
        std::vector<EntRec> recs(desc.entity_count);
        for (uint32_t i = 0; i < desc.entity_count; ++i) recs[i] = {int32_t(i), bounds[i]};
        TlasBuilder tb{&S.tlas_nodes, &S.tlas_prims, desc.max_leaf_size > 0 ? desc.max_leaf_size : 5};
        if (recs.empty()) {
            S.tlas_nodes.push_back(vec4{0, 0, 0, bits_f(TLAS_LEAF)});
            S.tlas_nodes.push_back(vec4{1, 1, 1, bits_f(0)});
        } else {
            tb.build(recs.data(), recs.size(), 0);
        }
        S.max_tlas_depth = tb.max_depth;
    } else if (desc.aggregate != AGX_AGGREGATE_NATIVE) {
        *err = "unknown aggregate kind";
        return AGX_E_INVALID_ARG;
    }
    // Atrc's aggregates test entities in a fixed order (the native aggregate in
    // JSON order, native_aggregate.cpp:35-44; the entity BVH leaf by leaf, left
    // subtree first, entity_bvh.cpp:150-170) and keep the last hit whose t is
    // within the inclusive bound, so of two entities hit at exactly the same t
    // the later one wins.  The BVH4 traversal reproduces this with the rank.
    for (uint32_t i = 0; i < S.ents.size(); ++i) S.ents[i].tie_rank = int32_t(i);
    if (desc.aggregate == AGX_AGGREGATE_BVH)
        for (size_t k = 0; k < S.tlas_prims.size(); ++k) S.ents[S.tlas_prims[k]].tie_rank = int32_t(k);
    for (size_t k = 0; k + 2 < S.b4_prims.size(); k += 3)
        S.b4_prims[k + 2].w = bits_f(uint32_t(S.ents[f_bits(S.b4_prims[k].w)].tie_rank));
    if (S.max_blas_depth + 2 > BLAS_STACK || S.max_tlas_depth + 2 > TLAS_STACK) {
        *err = "BVH depth exceeds the traversal stack (blas " + std::to_string(S.max_blas_depth) + ", tlas " +
               std::to_string(S.max_tlas_depth) + ")";
        return AGX_E_LIMIT;
    }
    return AGX_OK;
}

}  // namespace agx
