// ORACLE — test infrastructure only (see atrc_oracle.hpp).
// extern "C" entry points so tests and the benchmark's CPU-baseline leg can
// drive the restatement through ctypes.
#include <chrono>
#include <cstdio>
#include <string>

#include "atrc_oracle.hpp"

namespace {

thread_local std::string g_err;

struct Holder {
    int precision = 32;
    oracle::Scene<float> sf;
    oracle::Scene<double> sd;
};

template<class T> void copy_counters(const oracle::Counters &c, uint64_t *out) {
    if (!out) return;
    out[0] = c.closest; out[1] = c.shadow; out[2] = c.node_visits; out[3] = c.tri_tests; out[4] = c.hits;
}

template<class T>
int trace(const oracle::Scene<T> &sc, const float *o, const float *d, const float *tmin, const float *tmax,
          uint32_t n, int any_hit, int32_t *ent, int32_t *prim, float *t, float *u, float *v, uint64_t *cnt) {
    oracle::Counters c;
    for (uint32_t i = 0; i < n; ++i) {
        const oracle::Ray<T> r(oracle::V3<T>(T(o[3 * i]), T(o[3 * i + 1]), T(o[3 * i + 2])),
                               oracle::V3<T>(T(d[3 * i]), T(d[3 * i + 1]), T(d[3 * i + 2])),
                               T(tmin[i]), T(tmax[i]));
        if (any_hit) {
            ent[i] = sc.any(r, &c) ? 1 : -1;
            continue;
        }
        oracle::EntityHit<T> h;
        if (sc.closest(r, &h, &c)) {
            ent[i] = h.entity;
            if (prim) prim[i] = h.prim;
            if (t) t[i] = float(h.t);
            if (u) u[i] = float(h.bu);
            if (v) v[i] = float(h.bv);
        } else {
            ent[i] = -1;
            if (prim) prim[i] = -1;
            if (t) t[i] = -1;
            if (u) u[i] = 0;
            if (v) v[i] = 0;
        }
    }
    copy_counters<T>(c, cnt);
    return 0;
}

}  // namespace

extern "C" {

const char *oracle_last_error() { return g_err.c_str(); }

void *oracle_scene_create(const agx_scene_desc *d, int precision, float eps) {
    try {
        auto *h = new Holder();
        h->precision = precision;
        if (precision == 64) { h->sd.eps = double(eps); h->sd.build(*d); }
        else { h->sf.eps = eps; h->sf.build(*d); }
        return h;
    } catch (const std::exception &e) {
        g_err = e.what();
        return nullptr;
    }
}

void oracle_scene_destroy(void *p) { delete static_cast<Holder *>(p); }

int oracle_trace_rays(void *p, const float *o, const float *d, const float *tmin, const float *tmax, uint32_t n,
                      int any_hit, int32_t *ent, int32_t *prim, float *t, float *u, float *v, uint64_t *counters) {
    auto *h = static_cast<Holder *>(p);
    if (!h || !o || !d || !tmin || !tmax || !ent) { g_err = "null argument"; return 1; }
    if (h->precision == 64) return trace(h->sd, o, d, tmin, tmax, n, any_hit, ent, prim, t, u, v, counters);
    return trace(h->sf, o, d, tmin, tmax, n, any_hit, ent, prim, t, u, v, counters);
}

// Render rows [row_begin, row_end) (row_end < 0: all) with `threads` workers.
int oracle_render(void *p, const agx_render_params *rp, int threads, double *rgb, double *albedo, double *normal,
                  double *denoise, uint64_t *counters, int row_begin, int row_end, double *seconds) {
    auto *h = static_cast<Holder *>(p);
    if (!h || !rp) { g_err = "null argument"; return 1; }
    try {
        oracle::RenderOut out{rgb, albedo, normal, denoise};
        oracle::Counters c;
        const auto t0 = std::chrono::steady_clock::now();
        if (h->precision == 64) {
            h->sd.eps = double(rp->eps);
            oracle::render(h->sd, *rp, threads, out, &c, row_begin, row_end);
        } else {
            h->sf.eps = rp->eps;
            oracle::render(h->sf, *rp, threads, out, &c, row_begin, row_end);
        }
        const auto t1 = std::chrono::steady_clock::now();
        if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
        copy_counters<float>(c, counters);
        return 0;
    } catch (const std::exception &e) {
        g_err = e.what();
        return 1;
    }
}

// One camera sample (trace_sample): out = value[3], albedo[3], normal[3], denoise, sx, sy.
int oracle_trace_sample(void *p, const agx_render_params *rp, int qx, int qy, uint64_t key, uint64_t sample,
                        double *out) {
    auto *h = static_cast<Holder *>(p);
    if (!h || !rp || !out) { g_err = "null argument"; return 1; }
    auto run = [&](auto &sc) {
        sc.eps = decltype(sc.eps)(rp->eps);
        float sx, sy;
        const auto r = oracle::trace_sample(sc, *rp, qx, qy, key, sample, &sx, &sy);
        for (int k = 0; k < 3; ++k) {
            out[k] = double(r.value[k]); out[3 + k] = double(r.albedo[k]); out[6 + k] = double(r.normal[k]);
        }
        out[9] = double(r.denoise); out[10] = sx; out[11] = sy;
    };
    if (h->precision == 64) run(h->sd); else run(h->sf);
    return 0;
}

// PCG32 uniform stream of (seed, pixel, sample): n floats.
// Closest hit over a triangle soup (9 floats per triangle: A, B, C), brute
// force, with the watertight test (mode 1) or Atrc's Moller-Trumbore test on
// (A, B - A, C - A) (mode 0).  A later triangle at an equal t replaces the
// earlier one, as in Atrc's leaf loops.
int oracle_triangle_soup_closest(const float *tris, uint32_t n_tris, const float *o, const float *d,
                                 const float *tmin, const float *tmax, uint32_t n, int mode, int32_t *prim,
                                 float *t, float *u, float *v) {
    using V = oracle::V3<float>;
    for (uint32_t i = 0; i < n; ++i) {
        const oracle::Ray<float> r(V(o[3 * i], o[3 * i + 1], o[3 * i + 2]), V(d[3 * i], d[3 * i + 1], d[3 * i + 2]),
                                   tmin[i], tmax[i]);
        const oracle::ShearedRay<float> w(r);
        float best = tmax[i];
        prim[i] = -1; t[i] = -1; u[i] = 0; v[i] = 0;
        for (uint32_t k = 0; k < n_tris; ++k) {
            const float *q = tris + 9 * size_t(k);
            const V A(q[0], q[1], q[2]), B(q[3], q[4], q[5]), C(q[6], q[7], q[8]);
            float tt, uu, vv;
            bool hit;
            if (mode == 1) {
                hit = oracle::watertight_closest(w, tmin[i], best, A, B, C, &tt, &uu, &vv);
            } else {
                const oracle::Ray<float> rr(r.o, r.d, tmin[i], best);
                hit = oracle::mt_closest(rr, A, B - A, C - A, &tt, &uu, &vv);
            }
            if (hit) { best = tt; prim[i] = int32_t(k); t[i] = tt; u[i] = uu; v[i] = vv; }
        }
    }
    return 0;
}

void oracle_pcg_stream(uint64_t seed, uint64_t pixel, uint64_t sample, uint32_t n, float *out) {
    oracle::Pcg32 r;
    r.seed_path(seed, pixel, sample);
    for (uint32_t i = 0; i < n; ++i) out[i] = r.uniform();
}

// Raw PCG32 outputs for pcg32_srandom_r(initstate, initseq).
void oracle_pcg_raw(uint64_t initstate, uint64_t initseq, uint32_t n, uint32_t *out) {
    oracle::Pcg32 r;
    r.seed(initstate, initseq);
    for (uint32_t i = 0; i < n; ++i) out[i] = r.next();
}

// Distribution samplers: kind 0 disk(u,v)->(x,y), 1 cos-hemisphere->(x,y,z,pdf),
// 2 sphere->(x,y,z), 3 cone(cos_max=w)->(x,y,z,pdf), 4 triangle->(b1,b2), 5 from_z frame (u,v,w = z) -> 9 floats
void oracle_distribution(int kind, double u, double v, double w, double *out) {
    using namespace oracle;
    switch (kind) {
    case 0: { auto d = uniform_on_unit_disk(u, v); out[0] = d.x; out[1] = d.y; break; }
    case 1: { V3<double> dd; double p; zweighted_on_hemisphere(u, v, &dd, &p); out[0] = dd.x; out[1] = dd.y; out[2] = dd.z; out[3] = p; break; }
    case 2: { auto d = uniform_on_sphere(u, v); out[0] = d.x; out[1] = d.y; out[2] = d.z; break; }
    case 3: { auto d = uniform_on_cone(w, u, v); out[0] = d.x; out[1] = d.y; out[2] = d.z; out[3] = uniform_on_cone_pdf(w); break; }
    case 4: { auto d = uniform_on_triangle(u, v); out[0] = d.x; out[1] = d.y; break; }
    case 5: {
        auto c = Coord<double>::from_z(V3<double>(u, v, w));
        const V3<double> *ax[3] = {&c.x, &c.y, &c.z};
        for (int i = 0; i < 3; ++i) { out[3 * i] = ax[i]->x; out[3 * i + 1] = ax[i]->y; out[3 * i + 2] = ax[i]->z; }
        break;
    }
    default: break;
    }
}

// BSDF hooks in the local shading frame (geometry frame = shading frame = identity
// unless geo_z is given).  mode 0: eval(wi, wo) -> f[3]; 1: pdf -> out[0];
// 2: sample(wo, u, v, w) -> dir[3], f[3], pdf, is_delta.
int oracle_bsdf(const agx_material_desc *m, int precision, const double *geo_z, const double *wi, const double *wo,
                double u, double v, double w, int mode, double *out) {
    using namespace oracle;
    auto run = [&](auto zero) -> int {
        using T = decltype(zero);
        Material<T> mat; mat.d = *m;
        SurfacePoint<T> sp;
        if (geo_z) sp.geometry_coord = Coord<T>::from_z(V3<T>(T(geo_z[0]), T(geo_z[1]), T(geo_z[2])));
        V3<T> sn;
        BSDF<T> b = mat.shade(sp, &sn);
        const V3<T> vo{T(wo[0]), T(wo[1]), T(wo[2])};
        if (mode == 0) {
            const V3<T> f = b.eval(V3<T>(T(wi[0]), T(wi[1]), T(wi[2])), vo);
            out[0] = double(f.x); out[1] = double(f.y); out[2] = double(f.z);
        } else if (mode == 1) {
            out[0] = double(b.pdf(V3<T>(T(wi[0]), T(wi[1]), T(wi[2])), vo));
        } else {
            const BSDFSample<T> s = b.sample(vo, T(u), T(v), T(w), T(3e-4));
            out[0] = double(s.dir.x); out[1] = double(s.dir.y); out[2] = double(s.dir.z);
            out[3] = double(s.f.x); out[4] = double(s.f.y); out[5] = double(s.f.z);
            out[6] = double(s.pdf); out[7] = s.is_delta ? 1.0 : 0.0;
        }
        return 0;
    };
    if (precision == 64) return run(double(0));
    return run(float(0));
}

// Medium queries on medium `mi` of a created scene, n trials along the segment
// a -> b, trial i drawing from the PCG stream (seed, pixel = i, sample = 0).
// mode 0: tr -> 3 floats; 1: ab -> 3; 2: sample_scattering(indirect) ->
// happened, pos[3], wr[3], throughput[3], g (11 floats).
int oracle_medium(void *h, int mi, int mode, const double *a, const double *b, int indirect, uint64_t seed,
                  uint32_t n, double *out) {
    using namespace oracle;
    auto *H = static_cast<Holder *>(h);
    auto run = [&](const auto &sc) -> int {
        using T = typename std::decay_t<decltype(sc)>::Real;
        if (mi < 0 || size_t(mi) >= sc.media.size()) { g_err = "medium index out of range"; return -1; }
        const Medium<T> &m = sc.media[size_t(mi)];
        const V3<T> A{T(a[0]), T(a[1]), T(a[2])}, B{T(b[0]), T(b[1]), T(b[2])};
        for (uint32_t i = 0; i < n; ++i) {
            Sampler<T> s;
            s.rng.seed_path(seed, i, 0);
            if (mode == 0 || mode == 1) {
                const V3<T> r = mode == 0 ? m.tr(A, B, s) : m.ab(A, B, s);
                out[3 * i] = double(r.x); out[3 * i + 1] = double(r.y); out[3 * i + 2] = double(r.z);
            } else {
                const MediumScatter<T> r = m.sample_scattering(A, B, s, indirect != 0);
                double *o = out + 11 * size_t(i);
                o[0] = r.happened ? 1 : 0;
                o[1] = double(r.pos.x); o[2] = double(r.pos.y); o[3] = double(r.pos.z);
                o[4] = double(r.wr.x); o[5] = double(r.wr.y); o[6] = double(r.wr.z);
                o[7] = double(r.throughput.x); o[8] = double(r.throughput.y); o[9] = double(r.throughput.z);
                o[10] = double(r.phase.g);
            }
        }
        return 0;
    };
    return H->precision == 64 ? run(H->sd) : run(H->sf);
}

// 3D texture lookups of medium `mi`: which 0 density, 1 albedo, 2 g; n points -> 3 floats each.
int oracle_tex3d(void *h, int mi, int which, const double *uvw, uint32_t n, double *out) {
    using namespace oracle;
    auto *H = static_cast<Holder *>(h);
    auto run = [&](const auto &sc) -> int {
        using T = typename std::decay_t<decltype(sc)>::Real;
        if (mi < 0 || size_t(mi) >= sc.media.size()) { g_err = "medium index out of range"; return -1; }
        const Medium<T> &m = sc.media[size_t(mi)];
        const Texture3D<T> &t = which == 0 ? m.density : (which == 1 ? m.albedo : m.gtex);
        for (uint32_t i = 0; i < n; ++i) {
            const V3<T> r = t.sample(V3<T>{T(uvw[3 * i]), T(uvw[3 * i + 1]), T(uvw[3 * i + 2])});
            out[3 * i] = double(r.x); out[3 * i + 1] = double(r.y); out[3 * i + 2] = double(r.z);
        }
        return 0;
    };
    return H->precision == 64 ? run(H->sd) : run(H->sf);
}

// Environment light of the scene: mode 0 radiance(d) -> rgb; 1 pdf(d) -> 1 value;
// 2 sample(u, v, w given in in[3i..3i+2]) -> dir[3], pdf.  4 doubles out per item.
int oracle_env(void *h, int mode, const double *in, uint32_t n, double *out) {
    using namespace oracle;
    auto *H = static_cast<Holder *>(h);
    auto run = [&](const auto &sc) -> int {
        using T = typename std::decay_t<decltype(sc)>::Real;
        if (sc.env_kind == AGX_ENV_NONE) { g_err = "scene has no environment light"; return -1; }
        for (uint32_t i = 0; i < n; ++i) {
            const V3<T> a{T(in[3 * i]), T(in[3 * i + 1]), T(in[3 * i + 2])};
            double *o = out + 4 * size_t(i);
            if (mode == 0) {
                const V3<T> r = sc.sky_radiance(a);
                o[0] = double(r.x); o[1] = double(r.y); o[2] = double(r.z); o[3] = 0;
            } else if (mode == 1) {
                o[0] = double(sc.env_pdf(a)); o[1] = o[2] = o[3] = 0;
            } else {
                T pdf;
                const V3<T> d = sc.env_sample(a.x, a.y, a.z, &pdf);
                o[0] = double(d.x); o[1] = double(d.y); o[2] = double(d.z); o[3] = double(pdf);
            }
        }
        return 0;
    };
    return H->precision == 64 ? run(H->sd) : run(H->sf);
}

// Henyey-Greenstein: mode 0 eval(wi, wo) -> 1 value; 1 sample(wo, u, v) -> dir[3], f.
int oracle_hg(double g, int mode, const double *wi, const double *wo, double u, double v, double *out) {
    using namespace oracle;
    HG<double> hg;
    hg.g = g;
    const V3<double> vo{wo[0], wo[1], wo[2]};
    if (mode == 0) {
        out[0] = hg.eval(V3<double>{wi[0], wi[1], wi[2]}, vo);
    } else {
        V3<double> d;
        double f;
        hg.sample(vo, u, v, &d, &f);
        out[0] = d.x; out[1] = d.y; out[2] = d.z; out[3] = f;
    }
    return 0;
}

// One axis of Atrc's apply_image_filter (framebuffer.h:49-80) for a sample at s
// in a film grid with pixel range [lo, hi]: the pixels reached and the
// distances passed to the filter.  Returns the count (at most cap).
int oracle_filter_axis(int lo, int hi, float r, float s, int32_t *pix, float *rel, int cap) {
    int p[oracle::AXIS_CAP];
    float d[oracle::AXIS_CAP];
    const int n = oracle::atrc_filter_axis(lo, hi, r, s, p, d, std::min(cap, oracle::AXIS_CAP));
    for (int i = 0; i < n; ++i) { pix[i] = p[i]; rel[i] = d[i]; }
    return n;
}

}  // extern "C"
