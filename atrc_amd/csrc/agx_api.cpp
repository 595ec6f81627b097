// C ABI of the gfx950 path tracer (include/atrc_gpu.h): device context, scene
// upload, the wavefront render loop, the ray-query parity hook and counters.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <condition_variable>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/atrc_gpu.h"
#include "../../include/atrc_gpu_debug.h"
#include "agx_guide.h"
#include "agx_guide_kernels.h"
#include "agx_build.h"
#include "agx_kernels.h"

using namespace agx;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) return fail(AGX_E_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

template <class T> hipError_t dalloc(T **p, size_t count) {
    *p = nullptr;
    if (!count) return hipSuccess;
    return hipMalloc(reinterpret_cast<void **>(p), count * sizeof(T));
}

template <class T> hipError_t upload(T **p, const std::vector<T> &v, hipStream_t st) {
    hipError_t e = dalloc(p, v.size());
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpyAsync(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st);
}

}  // namespace

struct agx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // scene
    bool has_scene = false;
    DevScene ds{};
    std::vector<void *> scene_allocs;
    int n_lights = 0;
    int accel = AGX_ACCEL_ATRC;
    // wavefront buffers
    PathBuffers pb{};
    std::vector<void *> wave_allocs;
    uint4 *qc_blk = nullptr;  // per-block counts / offsets of the queue compaction
    size_t cap_paths = 0;
    int cap_L = -1;
    // paths per wavefront pass, chosen once per light count (the free-memory
    // query takes milliseconds, and the choice stays stable from call to call)
    size_t pass_paths = 0;
    int pass_paths_L = -1;
    size_t pass_paths_limit = 0;  // 0: none; batch workers, which render side by side, use 2^22
    int trace_block = 128;        // threads per block of the BVH4 traversal launches (128 or 256)
    int64_t cap_wh = 0;
    uint32_t *counters = nullptr;    // device: two sets of 4 queue sizes (ping-pong)
    DevCounters *dcnt = nullptr;
    uint32_t *h_counters = nullptr;  // pinned ring: RING x 4 queue sizes read back asynchronously
    float *d_resolve = nullptr;      // normalised film planes (10 floats per pixel), kept between renders
    float *h_resolve = nullptr;      // pinned staging of the planes the caller asked for
    size_t resolve_cap = 0;          // pixels both buffers hold
    static constexpr int RING = 64;
    // worker contexts of agx_render_batch (created on first use, same device)
    std::vector<agx_ctx *> batch_workers;
    // pooled contexts of agx_render_batch_graph, kept after a successful call
    // (with their path buffers and last scenes) for the next one
    std::vector<agx_ctx *> graph_pool;
    // ray capture of agx_debug_replay_rays (null outside it): [0] closest-hit, [1] any-hit rays
    float4 *capture_o[2] = {nullptr, nullptr}, *capture_d[2] = {nullptr, nullptr};
    uint32_t *capture_count = nullptr;  // device: 2 counters
    uint32_t capture_cap = 0;
    // path guiding: network state (persists across renders) and its wave buffers
    static constexpr int GUIDE_TRAIN_BLOCKS = 256;
    bool guide_ready = false;
    float *g_params = nullptr, *g_m = nullptr, *g_v = nullptr, *g_grad = nullptr, *g_partial = nullptr;
    float *g_bias = nullptr, *g_samples = nullptr, *g_wsum = nullptr;
    __bf16 *g_packed = nullptr;
    uint32_t *g_count = nullptr;
    uint32_t *g_gather_blk = nullptr;  // per-block counts / offsets of the sample gather
    uint32_t *g_ring = nullptr;      // sample ring (next slot, valid samples) over g_samples
    uint32_t g_ring_cap = 0;         // capacity the ring state refers to (0: empty)
    uint32_t g_batch_cap = 0;
    uint64_t g_steps = 0;
    size_t g_wave_cap = 0;
    size_t g_ms_cap = 0;
    std::vector<void *> guide_allocs, guide_wave_allocs;
    hipEvent_t ev[8]{};
    std::vector<hipEvent_t> pool;    // timing / readback events, reused across renders
    hipEvent_t event(size_t i) {
        while (pool.size() <= i) {
            hipEvent_t e = nullptr;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            pool.push_back(e);
        }
        return pool[i];
    }
    agx_stats_t stats{};

    void free_scene() {
        for (void *p : scene_allocs) if (p) (void)hipFree(p);
        scene_allocs.clear();
        has_scene = false;
    }
    void free_guide() {
        for (void *p : guide_allocs) if (p) (void)hipFree(p);
        guide_allocs.clear();
        guide_ready = false;
        g_batch_cap = 0;
        g_steps = 0;
    }
    void free_guide_wave() {
        for (void *p : guide_wave_allocs) if (p) (void)hipFree(p);
        guide_wave_allocs.clear();
        pb.g_feat = pb.g_out = pb.g_rec = nullptr;
        pb.g_rec_state = nullptr;
        pb.g_ms = nullptr;
        g_wave_cap = 0;
        g_ms_cap = 0;
    }
    void free_wave() {
        for (void *p : wave_allocs) if (p) (void)hipFree(p);
        wave_allocs.clear();
        free_guide_wave();
        pb = PathBuffers{};
        counters = nullptr;
        qc_blk = nullptr;
        cap_paths = 0;
        cap_L = -1;
        cap_wh = 0;
    }
    template <class T> hipError_t walloc(T **p, size_t count) {
        hipError_t e = dalloc(p, count);
        if (e == hipSuccess) wave_allocs.push_back(*p);
        return e;
    }
    template <class T> hipError_t galloc(std::vector<void *> &list, T **p, size_t count) {
        hipError_t e = dalloc(p, count);
        if (e == hipSuccess) list.push_back(*p);
        return e;
    }
    // per-path buffers of guided rendering (allocated on the first guided render)
    int ensure_guide_wave(size_t n) {
        if (n <= g_wave_cap) return AGX_OK;
        free_guide_wave();
        const size_t nr = guide_rec_paths(n);
        HIP_TRY(galloc(guide_wave_allocs, &pb.g_feat, n * 8));
        HIP_TRY(galloc(guide_wave_allocs, &pb.g_out, n * GUIDE_ROW));
        HIP_TRY(galloc(guide_wave_allocs, &pb.g_rec, nr * GUIDE_MAX_REC * GUIDE_REC_FLOATS));
        HIP_TRY(galloc(guide_wave_allocs, &pb.g_rec_state, nr));
        g_wave_cap = n;
        return AGX_OK;
    }
    // the medium samples of guided renders in media (k_medium_pass)
    int ensure_guide_media(size_t n) {
        if (n <= g_ms_cap) return AGX_OK;
        HIP_TRY(galloc(guide_wave_allocs, &pb.g_ms, 3 * n));
        g_ms_cap = n;
        return AGX_OK;
    }
    // network state; parameters initialised uniform in +-1/sqrt(fan-in) (weights)
    // and +-0.1 (biases) from a splitmix64 stream of `seed`, moments zero
    int guide_init(uint32_t seed, uint32_t batch) {
        if (!guide_ready) {
            HIP_TRY(galloc(guide_allocs, &g_params, G_PARAMS));
            HIP_TRY(galloc(guide_allocs, &g_m, G_PARAMS));
            HIP_TRY(galloc(guide_allocs, &g_v, G_PARAMS));
            HIP_TRY(galloc(guide_allocs, &g_grad, G_PARAMS));
            HIP_TRY(galloc(guide_allocs, &g_partial, size_t(GUIDE_TRAIN_BLOCKS) * G_PARAMS));
            HIP_TRY(galloc(guide_allocs, &g_bias, G_BIASES));
            HIP_TRY(galloc(guide_allocs, &g_packed, G_PACK_ALL));
            HIP_TRY(galloc(guide_allocs, &g_count, 4));
            HIP_TRY(galloc(guide_allocs, &g_gather_blk, GUIDE_GATHER_BLOCKS));
            HIP_TRY(galloc(guide_allocs, &g_wsum, 4));
            HIP_TRY(galloc(guide_allocs, &g_ring, 2));
        }
        g_ring_cap = 0;
        std::vector<float> h(G_PARAMS);
        uint64_t x = 0x243F6A8885A308D3ull ^ uint64_t(seed);
        auto next = [&]() {  // splitmix64 -> uniform in [-1, 1)
            x += 0x9E3779B97F4A7C15ull;
            uint64_t z = x;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            return float(double(z >> 11) * 0x1p-53 * 2.0 - 1.0);
        };
        for (int l = 0; l < G_LAYERS; ++l) {
            const float bw = 1.0f / std::sqrt(float(g_in(l)));
            for (int i = 0; i < g_in(l) * g_out(l); ++i) h[size_t(g_w_off(l) + i)] = bw * next();
            for (int i = 0; i < g_out(l); ++i) h[size_t(g_b_off(l) + i)] = 0.1f * next();
        }
        HIP_TRY(hipMemcpyAsync(g_params, h.data(), G_PARAMS * sizeof(float), hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemsetAsync(g_m, 0, G_PARAMS * sizeof(float), stream));
        HIP_TRY(hipMemsetAsync(g_v, 0, G_PARAMS * sizeof(float), stream));
        launch_guide_pack(g_params, g_packed, g_bias, stream);
        HIP_TRY(hipStreamSynchronize(stream));
        g_steps = 0;
        guide_ready = true;
        return ensure_guide_batch(batch);
    }
    // sample storage for a ring of `batch` samples; the ring restarts empty when
    // its capacity changes
    int ensure_guide_batch(uint32_t batch) {
        if (batch > g_batch_cap) {
            HIP_TRY(galloc(guide_allocs, &g_samples, size_t(batch) * (G_IN + 4)));
            g_batch_cap = batch;
        }
        if (g_ring_cap != batch) {
            HIP_TRY(hipMemsetAsync(g_ring, 0, 2 * sizeof(uint32_t), stream));
            g_ring_cap = batch;
        }
        return AGX_OK;
    }
    int ensure_wave(size_t n, int L, int64_t wh) {
        if (n <= cap_paths && L <= cap_L && wh <= cap_wh) return AGX_OK;
        free_wave();
        const size_t nl = n * size_t(std::max(L, 1));
        HIP_TRY(walloc(&pb.coef, n));
        HIP_TRY(walloc(&pb.radiance, n));
        HIP_TRY(walloc(&pb.gb_albedo, n));
        HIP_TRY(walloc(&pb.gb_normal, n));
        HIP_TRY(walloc(&pb.rng, n));
        HIP_TRY(walloc(&pb.rng_tr, n));
        HIP_TRY(walloc(&pb.depth, n));
        HIP_TRY(walloc(&pb.flags, n));
        HIP_TRY(walloc(&pb.mis_a, n));
        HIP_TRY(walloc(&pb.mis_b, n));
        HIP_TRY(walloc(&pb.nee, nl));
        HIP_TRY(walloc(&pb.ray_o, 2 * n));
        HIP_TRY(walloc(&pb.ray_d, 2 * n));
        HIP_TRY(walloc(&pb.hit, 2 * n));
        HIP_TRY(walloc(&pb.hit_prim, 2 * n));
        HIP_TRY(walloc(&pb.sh_o, nl));
        HIP_TRY(walloc(&pb.sh_d, nl));
        HIP_TRY(walloc(&pb.occluded, nl));
        HIP_TRY(walloc(&pb.q_closest, 2 * n));
        HIP_TRY(walloc(&pb.q_shadow, nl));
        HIP_TRY(walloc(&pb.q_in, n));
        HIP_TRY(walloc(&pb.q_out, n));
        HIP_TRY(walloc(&counters, 8));
        HIP_TRY(walloc(&qc_blk, queue_compact_blocks(uint32_t(n))));
        HIP_TRY(walloc(&pb.film_sum, size_t(10 * wh)));
        HIP_TRY(walloc(&pb.film_comp, size_t(10 * wh)));
        HIP_TRY(walloc(&pb.film_count, size_t(wh)));
        cap_paths = n;
        cap_L = std::max(L, 1);
        cap_wh = wh;
        return AGX_OK;
    }
};

namespace {

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0;
    return ms;
}

// ThinLensCamera::init_from_params (thin_lens.cpp:34-55); look_at frame:
// z = normalize(dst - pos), x = normalize(up x z), y = z x x.
void camera_frame(const agx_camera_desc &c, RenderConst *rc) {
    auto sub = [](const float *a, const float *b, float *o) { for (int i = 0; i < 3; ++i) o[i] = a[i] - b[i]; };
    auto norm = [](float *v) {
        const float l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (int i = 0; i < 3; ++i) v[i] = v[i] / l;
    };
    auto cross = [](const float *a, const float *b, float *o) {
        o[0] = a[1] * b[2] - a[2] * b[1];
        o[1] = a[2] * b[0] - a[0] * b[2];
        o[2] = a[0] * b[1] - a[1] * b[0];
    };
    float z[3], x[3], y[3];
    sub(c.dst, c.pos, z);
    norm(z);
    cross(c.up, z, x);
    norm(x);
    cross(z, x, y);
    rc->cam_pos = make_float4(c.pos[0], c.pos[1], c.pos[2], 0);
    rc->cam_x = make_float4(x[0], x[1], x[2], 0);
    rc->cam_y = make_float4(y[0], y[1], y[2], 0);
    rc->cam_z = make_float4(z[0], z[1], z[2], 0);
    rc->focal = c.focal_distance;
    rc->film_h = 2 * c.focal_distance * std::tan(c.fov_rad / 2);
    rc->film_w = c.film_aspect * rc->film_h;
    rc->lens_r = c.lens_radius;
}

void fill_traversal_stats(const DevCounters hc[2], agx_stats_t *st) {
    st->closest_node_visits = hc[0].node_visits;
    st->closest_tri_tests = hc[0].tri_tests;
    st->closest_hits = hc[0].hits;
    st->shadow_node_visits = hc[1].node_visits;
    st->shadow_tri_tests = hc[1].tri_tests;
    st->closest_bvh_bytes = hc[0].node_visits * st->node_bytes + hc[0].tri_tests * st->prim_bytes +
                            hc[0].hits * st->info_bytes;
    st->shadow_bvh_bytes = hc[1].node_visits * st->node_bytes + hc[1].tri_tests * st->prim_bytes;
}

// Paths of one wavefront pass.  Larger passes keep the late bounces of a pass
// (fewer surviving paths) large enough to fill the chip: on the 1M-triangle
// 1080p scene 4M -> 33.5M paths took the closest-hit time per 16 spp from 33.2
// to 24.3 ms.  Default 2^25 paths (about 320 B + 53 B per light each, ~11 GB
// with one light), limited to a quarter of the free device memory.
size_t max_paths_in_flight(int lights) {
    if (const char *s = std::getenv("AGX_MAX_PATHS")) {
        const long long v = std::atoll(s);
        if (v >= 1024) return size_t(v);
    }
    size_t cap = size_t(1) << 25;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && free_b > 0) {
        const size_t per_path = 320 + 53 * size_t(std::max(lights, 1));
        cap = std::min(cap, std::max<size_t>(size_t(1) << 20, free_b / 4 / per_path));
    }
    return cap;
}

// Validation of a render's parameters and its RenderConst (image, depths,
// filter borders, camera); shared by agx_render and the graph-captured batch.
int fill_render_const(agx_ctx *c, const agx_render_params *p, RenderConst *out) {
    if (!c->has_scene) return fail(AGX_E_NO_SCENE, "no scene uploaded");
    if (p->width <= 0 || p->height <= 0 || p->spp <= 0) return fail(AGX_E_INVALID_ARG, "width, height and spp must be positive");
    if (p->max_depth < 1 || p->min_depth < 0 || !(p->cont_prob > 0) || !(p->eps > 0))
        return fail(AGX_E_INVALID_ARG, "invalid depth / cont_prob / eps");
    // BoxFilter radius_ = radius - EPS (box_filter.cpp:15-21); GaussianFilter keeps
    // its radius and needs alpha > 0 (gaussian_filter.cpp:45-56)
    if (p->filter != AGX_FILTER_BOX && p->filter != AGX_FILTER_GAUSSIAN)
        return fail(AGX_E_INVALID_ARG, "unknown film filter");
    const bool gaussian = p->filter == AGX_FILTER_GAUSSIAN;
    const float fr = gaussian ? p->filter_radius : p->filter_radius - p->eps;
    if (!(fr > 0)) return fail(AGX_E_INVALID_ARG, "invalid film filter radius");
    if (gaussian && !(p->filter_alpha > 0)) return fail(AGX_E_INVALID_ARG, "invalid Gaussian filter alpha");
    if (fr > 64.0f) return fail(AGX_E_LIMIT, "film filter radius above 64 pixels");
    HIP_TRY(hipSetDevice(c->device));
    RenderConst &rc = *out;
    rc.W = p->width;
    rc.H = p->height;
    rc.min_depth = p->min_depth;
    rc.max_depth = p->max_depth;
    rc.specular_depth = p->specular_depth;
    rc.cont_prob = p->cont_prob;
    rc.eps = p->eps;
    rc.filter_r = fr;
    // a filter wider than half a pixel splats; sample_bound_of (framebuffer.h:36-47)
    // adds floor-based borders around the image
    rc.splat = (gaussian || fr > 0.5f) ? 1 : 0;
    rc.nomis = p->use_mis == 0 ? 1 : 0;  // PathTracingRenderer: trace_std or trace_nomis (pt.cpp:29-31)
    rc.filter_kind = p->filter;
    if (p->task_grid_size < 0) return fail(AGX_E_INVALID_ARG, "invalid task_grid_size");
    rc.tile = p->task_grid_size > 0 ? p->task_grid_size : 32;
    rc.filter_alpha = gaussian ? p->filter_alpha : 0.0f;
    rc.filter_exp = gaussian ? std::exp(-p->filter_alpha * fr * fr) : 0.0f;
    rc.border0 = rc.splat ? -int32_t(std::floor(0.5f - fr)) : 0;
    rc.border1 = rc.splat ? int32_t(std::floor(fr + 0.49999f)) : 0;
    rc.EW = p->width + rc.border0 + rc.border1;
    rc.EH = p->height + rc.border0 + rc.border1;
    rc.seed = p->seed;
    rc.sample_begin = p->sample_begin;
    rc.L = c->n_lights;
    camera_frame(p->camera, &rc);
    return AGX_OK;
}

// device buffer freed at scope exit (debug entry points only)
struct DevBuf {
    void *p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
    template <class T> T *get() { return static_cast<T *>(p); }
};

}  // namespace

extern "C" {

int agx_abi_version(void) { return AGX_ABI_VERSION; }

const char *agx_last_error(void) { return g_err.c_str(); }

int agx_create(int device_ordinal, agx_ctx **out) {
    if (!out) return fail(AGX_E_INVALID_ARG, "out is NULL");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return fail(AGX_E_HIP, "no HIP device available");
    if (device_ordinal < 0 || device_ordinal >= count) return fail(AGX_E_INVALID_ARG, "device ordinal out of range");
    HIP_TRY(hipSetDevice(device_ordinal));
    auto *c = new agx_ctx();
    c->device = device_ordinal;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(AGX_E_HIP, "hipStreamCreate failed");
    }
    if (hipMalloc(reinterpret_cast<void **>(&c->dcnt), 2 * sizeof(DevCounters)) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void **>(&c->h_counters), agx_ctx::RING * 4 * sizeof(uint32_t),
                      hipHostMallocDefault) !=
            hipSuccess) {
        agx_destroy(c);
        return fail(AGX_E_OUT_OF_MEMORY, "allocation of counters failed");
    }
    for (auto &e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            agx_destroy(c);
            return fail(AGX_E_HIP, "hipEventCreate failed");
        }
    // the staging array of agx_trace_rays_noalloc, resolved once per device
    if (const hipError_t e = noalloc_scratch_init(device_ordinal); e != hipSuccess) {
        agx_destroy(c);
        return fail(AGX_E_HIP, std::string("staging array address: ") + hipGetErrorString(e));
    }
    *out = c;
    return AGX_OK;
}

void agx_destroy(agx_ctx *c) {
    if (!c) return;
    for (agx_ctx *w : c->batch_workers) agx_destroy(w);
    c->batch_workers.clear();
    for (agx_ctx *w : c->graph_pool) agx_destroy(w);
    c->graph_pool.clear();
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->free_scene();
    c->free_wave();
    c->free_guide();
    if (c->dcnt) (void)hipFree(c->dcnt);
    if (c->h_counters) (void)hipHostFree(c->h_counters);
    if (c->d_resolve) (void)hipFree(c->d_resolve);
    if (c->h_resolve) (void)hipHostFree(c->h_resolve);
    for (auto &e : c->ev) if (e) (void)hipEventDestroy(e);
    for (auto &e : c->pool) if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

}  // extern "C"

// The device half of agx_upload_scene: copy a built host scene into the
// context (replacing its scene) and fill its DevScene.
static int upload_host_scene(agx_ctx *c, const HostScene &hs) {
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->free_scene();
    DevScene ds{};
    DevEntity *ents = nullptr;
    DevMaterial *mats = nullptr;
    vec4 *bn = nullptr, *bp = nullptr, *bi = nullptr, *tn = nullptr, *b4n = nullptr, *b4p = nullptr;
    int32_t *tp = nullptr, *ai = nullptr;
    float *ap = nullptr, *gr = nullptr, *etx = nullptr, *epr = nullptr, *eap = nullptr;
    int32_t *eai = nullptr;
    DevMedium *md = nullptr;
    auto track = [&](void *p) { if (p) c->scene_allocs.push_back(p); };
    hipError_t e = hipSuccess;
    if (e == hipSuccess) { e = upload(&ents, hs.ents, c->stream); track(ents); }
    if (e == hipSuccess) { e = upload(&etx, hs.env_texels, c->stream); track(etx); }
    if (e == hipSuccess) { e = upload(&epr, hs.env_probs, c->stream); track(epr); }
    if (e == hipSuccess) { e = upload(&eap, hs.env_alias_p, c->stream); track(eap); }
    if (e == hipSuccess) { e = upload(&eai, hs.env_alias_i, c->stream); track(eai); }
    if (e == hipSuccess) { e = upload(&mats, hs.mats, c->stream); track(mats); }
    if (e == hipSuccess) { e = upload(&bn, hs.blas_nodes, c->stream); track(bn); }
    if (e == hipSuccess) { e = upload(&bp, hs.blas_prims, c->stream); track(bp); }
    if (e == hipSuccess) { e = upload(&bi, hs.blas_info, c->stream); track(bi); }
    if (e == hipSuccess) { e = upload(&tn, hs.tlas_nodes, c->stream); track(tn); }
    if (e == hipSuccess) { e = upload(&tp, hs.tlas_prims, c->stream); track(tp); }
    if (e == hipSuccess) { e = upload(&ap, hs.alias_prob, c->stream); track(ap); }
    if (e == hipSuccess) { e = upload(&ai, hs.alias_idx, c->stream); track(ai); }
    if (e == hipSuccess) { e = upload(&md, hs.media, c->stream); track(md); }
    if (e == hipSuccess) { e = upload(&gr, hs.grids, c->stream); track(gr); }
    uint16_t *gr16 = nullptr;
    if (e == hipSuccess) { e = upload(&gr16, hs.grids16, c->stream); track(gr16); }
    if (e == hipSuccess) { e = upload(&b4n, hs.b4_nodes, c->stream); track(b4n); }
    if (e == hipSuccess) { e = upload(&b4p, hs.b4_prims, c->stream); track(b4p); }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        c->free_scene();
        return fail(e == hipErrorOutOfMemory ? AGX_E_OUT_OF_MEMORY : AGX_E_HIP,
                    std::string("scene upload: ") + hipGetErrorString(e));
    }
    ds.ents = ents;
    ds.mats = mats;
    ds.blas_nodes = bn;
    ds.blas_prims = bp;
    ds.blas_info = bi;
    ds.tlas_nodes = tn;
    ds.tlas_prims = tp;
    ds.alias_prob = ap;
    ds.alias_idx = ai;
    ds.media = md;
    ds.grids = gr;
    ds.grids16 = gr16;
    ds.accel = hs.accel;
    // AGX_TRACE_BLOCK=128 or 256 sets the traversal block size of contexts left at the default (128)
    ds.trace_block = c->trace_block;
    if (const char *e = std::getenv("AGX_TRACE_BLOCK")) {
        if (std::string(e) == "256") ds.trace_block = 256;
        else if (std::string(e) == "128") ds.trace_block = 128;
    }
    ds.b4_nodes = b4n;
    ds.b4_prims = b4p;
    // a BVH4 traversal keeps at most 3 pending children per level
    ds.bvh4_large_stack = 3 * hs.b4_depth + 1 > BVH4_STACK_SMALL ? 1 : 0;
    if (hs.accel != AGX_ACCEL_ATRC && 3 * hs.b4_depth + 1 > BVH4_STACK_LARGE) {
        c->free_scene();
        return fail(AGX_E_LIMIT, "BVH4 depth " + std::to_string(hs.b4_depth) + " exceeds the traversal stack");
    }
    c->accel = hs.accel;
    ds.features = hs.media.empty() ? 0u : SF_MEDIA;
    if (hs.env_kind == AGX_ENV_IBL) ds.features |= SF_IBL;
    for (const DevMaterial &m : hs.mats) {
        if (m.kind == AGX_MAT_DISNEY) ds.features |= SF_DISNEY;
        if (m.kind == AGX_MAT_GLASS || m.kind == AGX_MAT_MIRROR || m.kind == AGX_MAT_INVISIBLE_SURFACE)
            ds.features |= SF_DELTA;
    }
    ds.n_ent = int32_t(hs.ents.size());
    ds.aggregate = hs.aggregate;
    ds.n_lights = int32_t(hs.lights.size());
    // stack entries needed: TLAS pending nodes (depth + 1) + BLAS (depth + 1) + 1
    // The per-thread private stack is the default: at 44 VGPRs it keeps 8 waves per
    // SIMD, while the 32 KiB LDS column per block caps occupancy at 5 (measured
    // on the 1M-triangle scene: closest-hit 83.8 ms private vs 99.5 ms LDS for 8 spp).
    // AGX_TRAVERSAL_STACK=lds selects the LDS column when the depths fit it.
    ds.lds_stack = 0;
    if (const char *e = std::getenv("AGX_TRAVERSAL_STACK"))
        if (std::string(e) == "lds")
            ds.lds_stack = (hs.max_tlas_depth + 1) + (hs.max_blas_depth + 1) + 1 <= LDS_STACK ? 1 : 0;
    ds.env_kind = hs.env_kind;
    for (int i = 0; i < MAX_LIGHTS; ++i) ds.lights[i] = i < ds.n_lights ? hs.lights[size_t(i)] : -2;
    ds.sky_top = hs.sky_top;
    ds.sky_bottom = hs.sky_bottom;
    ds.env_tex = hs.env_tex;
    ds.env_texels = etx;
    ds.env_probs = epr;
    ds.env_alias_p = eap;
    ds.env_alias_i = eai;
    ds.env_pw = hs.env_pw;
    ds.env_ph = hs.env_ph;
    ds.world = hs.world;
    c->ds = ds;
    c->n_lights = ds.n_lights;
    c->has_scene = true;
    return AGX_OK;
}

extern "C" {

int agx_upload_scene(agx_ctx *c, const agx_scene_desc *desc) {
    if (!c || !desc) return fail(AGX_E_INVALID_ARG, "null argument");
    HostScene hs;
    std::string err;
    const int rc = build_host_scene(*desc, &hs, &err);
    if (rc != AGX_OK) return fail(rc, err);
    return upload_host_scene(c, hs);
}

int agx_render(agx_ctx *c, const agx_render_params *p, const volatile int32_t *cancel_flag, float *rgb,
               float *albedo, float *normal, float *denoise) {
    if (!c || !p) return fail(AGX_E_INVALID_ARG, "null argument");
    const auto t_call = std::chrono::steady_clock::now();
    RenderConst rc{};
    if (int r0 = fill_render_const(c, p, &rc); r0 != AGX_OK) return r0;
    const bool guide = p->guide != 0;
    const bool guide_train = guide && p->guide_train != 0;
    const uint32_t guide_batch = p->guide_batch > 0 ? uint32_t(p->guide_batch) : 65536u;
    const int guide_steps = p->guide_steps > 0 ? p->guide_steps : 4;
    const float guide_lr = p->guide_lr > 0 ? p->guide_lr : 1e-2f;
    if (guide) {
        if (rc.nomis) return fail(AGX_E_UNSUPPORTED, "path guiding runs with trace_std (use_mis = 1) only");
        const float a = p->guide_alpha > 0 ? p->guide_alpha : 0.5f;
        if (!(a > 0 && a < 1)) return fail(AGX_E_INVALID_ARG, "guide_alpha must lie in (0, 1)");
        if (guide_batch > (1u << 20)) return fail(AGX_E_INVALID_ARG, "guide_batch above 1 << 20");
        if (hipError_t e = guide_prepare(c->device); e != hipSuccess)
            return fail(AGX_E_UNSUPPORTED, std::string("path guiding: ") + hipGetErrorString(e));
        rc.guide = 1;
        rc.guide_alpha = a;
        rc.guide_record = guide_train ? 1 : 0;
    }

    const int64_t wh = int64_t(p->width) * p->height;
    // paths per sample layer: the image, or the image and its border when samples splat
    const int64_t layer = rc.splat ? int64_t(rc.EW) * rc.EH : wh;
    const int64_t total = layer * p->spp;
    if (c->pass_paths_L != rc.L) {
        c->pass_paths = max_paths_in_flight(rc.L);
        if (c->pass_paths_limit) c->pass_paths = std::min(c->pass_paths, c->pass_paths_limit);
        c->pass_paths_L = rc.L;
    }
    const size_t cap = c->pass_paths;
    int64_t per_pass;
    if (layer <= int64_t(cap)) per_pass = layer * std::min<int64_t>(p->spp, int64_t(cap) / layer);
    else per_pass = rc.splat ? layer : int64_t(cap);  // splatting needs whole layers in a pass
    // a training render takes one sample per pixel per pass, so that the network
    // learns after every sample layer
    if (guide_train && (layer <= int64_t(cap) || rc.splat)) per_pass = layer;
    rc.guide_rec_shift = guide_rec_shift(size_t(per_pass));
    int r = c->ensure_wave(size_t(per_pass), rc.L, wh);
    if (r != AGX_OK) return r;
    if (guide) {
        if ((r = c->ensure_guide_wave(size_t(per_pass))) != AGX_OK) return r;
        // media: with guiding at medium points, scattering points get their own
        // network evaluation (k_medium_pass)
        rc.guide_medium_depth = p->guide_medium_depth > 0 ? p->guide_medium_depth : 0;
        rc.guide_media = ((c->ds.features & SF_MEDIA) && rc.guide_medium_depth > 0) ? 1 : 0;
        if (rc.guide_media && (r = c->ensure_guide_media(size_t(per_pass))) != AGX_OK) return r;
        if (!c->guide_ready || p->guide_reset) {
            if ((r = c->guide_init(p->guide_seed, guide_batch)) != AGX_OK) return r;
        } else if ((r = c->ensure_guide_batch(guide_batch)) != AGX_OK) {
            return r;
        }
    }

    const bool stats = p->collect_stats != 0;
    agx_stats_t st{};
    st.node_bytes = c->accel != AGX_ACCEL_ATRC ? 16 * BVH4_NODE_VEC4 : 32;
    st.prim_bytes = 48;
    st.info_bytes = 96;
    if (stats) HIP_TRY(hipMemsetAsync(c->dcnt, 0, 2 * sizeof(DevCounters), c->stream));
    launch_film_clear(c->pb, wh, c->stream);
    HIP_TRY(hipEventRecord(c->ev[0], c->stream));
    // Timing events: (start, end, kind) per launch group; read after the render.
    struct Timed { size_t a, b; int kind; };
    std::vector<Timed> timed;
    size_t next_ev = 0;
    auto mark = [&](size_t *idx) -> int {
        hipEvent_t e = c->event(next_ev);
        if (!e) return fail(AGX_E_HIP, "hipEventCreate failed");
        *idx = next_ev++;
        HIP_TRY(hipEventRecord(e, c->stream));
        return AGX_OK;
    };
    // Queue sizes are read back asynchronously into a pinned ring; the host
    // checks termination two bounces behind, so the GPU never idles waiting.
    struct Readback { size_t ev; int slot; };
    std::vector<Readback> reads;
    int ring_next = 0;
    auto readback = [&](const uint32_t *dev) -> int {
        const int slot = ring_next++ % agx_ctx::RING;
        HIP_TRY(hipMemcpyAsync(c->h_counters + 4 * slot, dev, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
        size_t e;
        if (int rr = mark(&e)) return rr;
        reads.push_back({e, slot});
        return AGX_OK;
    };
    const uint32_t n_closest_max = uint32_t(2 * per_pass), n_shadow_max = uint32_t(per_pass * std::max(rc.L, 1));
    const int max_bounces = p->max_depth + std::max(p->specular_depth, 0) + 2;
    if (max_bounces + 1 > agx_ctx::RING)
        return fail(AGX_E_UNSUPPORTED, "max_depth + specular_depth too large (at most " +
                                           std::to_string(agx_ctx::RING - 3) + ")");
    bool cancelled = false;
    uint32_t guide_samples_h = 0, guide_valid_h = 0;
    for (int64_t base = 0; base < total && !cancelled; base += per_pass) {
        rc.work_base = base;
        rc.n_paths = int32_t(std::min<int64_t>(per_pass, total - base));
        rc.guide_rec_rot = uint32_t(uint64_t(base / per_pass) * 2654435761ull) & ((1u << rc.guide_rec_shift) - 1u);
        uint32_t *cin = c->counters, *cout = c->counters + 4;
        size_t a, b;
        if ((r = mark(&a))) return r;
        launch_reset_counters(cout, c->stream);
        c->pb.cnt_out = cout;
        launch_raygen(c->ds, rc, c->pb, c->stream);
        if ((r = mark(&b))) return r;
        timed.push_back({a, b, 2});
        if ((r = readback(cout))) return r;
        st.paths += uint64_t(rc.n_paths);
        const size_t first_read = reads.size() - 1;
        std::swap(cin, cout);
        std::swap(c->pb.q_in, c->pb.q_out);
        for (int bounce = 0; bounce < max_bounces; ++bounce) {
            c->pb.cnt_in = cin;
            c->pb.cnt_out = cout;
            size_t e0, e1, e2, e3;
            if ((r = mark(&e0))) return r;
            launch_trace_closest(c->ds, c->pb, n_closest_max, stats, c->dcnt, c->stream);
            if ((r = mark(&e1))) return r;
            if (c->capture_count) {
                launch_capture_rays(c->pb, 0, c->capture_o[0], c->capture_d[0], c->capture_count, c->capture_cap,
                                    n_closest_max, c->stream);
                launch_capture_rays(c->pb, 1, c->capture_o[1], c->capture_d[1], c->capture_count + 1,
                                    c->capture_cap, n_shadow_max, c->stream);
            }
            if (guide) {
                // the network's mixture for every vertex about to be shaded (in media,
                // at the scattering point the vertex's medium sample reaches)
                if (rc.guide_media) launch_medium_pass(c->ds, rc, c->pb, uint32_t(rc.n_paths), c->stream);
                launch_guide_features(c->ds, rc, c->pb, uint32_t(rc.n_paths), c->stream);
                size_t ef, eg;
                if ((r = mark(&ef))) return r;
                launch_guide_infer(c->g_packed, c->g_bias, c->pb.g_feat, c->pb.g_out, cin + 2, uint32_t(rc.n_paths),
                                   c->stream);
                if ((r = mark(&eg))) return r;
                timed.push_back({e1, ef, 2});
                timed.push_back({ef, eg, 3});
                e1 = eg;
            }
            launch_trace_shadow(c->ds, c->pb, n_shadow_max, stats, c->dcnt + 1, c->stream);
            if ((r = mark(&e2))) return r;
            launch_reset_counters(cout, c->stream);
            launch_shade(c->ds, rc, c->pb, uint32_t(rc.n_paths), c->stream);
            launch_queue_compact(c->pb, uint32_t(rc.n_paths), rc.L, c->qc_blk, c->stream);
            if ((r = mark(&e3))) return r;
            timed.push_back({e0, e1, 0});
            timed.push_back({e1, e2, 1});
            timed.push_back({e2, e3, 2});
            if ((r = readback(cout))) return r;
            std::swap(cin, cout);
            std::swap(c->pb.q_in, c->pb.q_out);
            // the queue produced two bounces ago: if it was empty, everything since is a no-op
            if (bounce >= 1) {
                const Readback &rb = reads[first_read + size_t(bounce)];
                HIP_TRY(hipEventSynchronize(c->pool[rb.ev]));
                if (c->h_counters[4 * rb.slot + 2] == 0) break;
            }
            if (cancel_flag && *cancel_flag) { cancelled = true; break; }
        }
        if (cancelled) break;
        if ((r = mark(&a))) return r;
        launch_film(rc, c->pb, c->stream);
        if ((r = mark(&b))) return r;
        timed.push_back({a, b, 2});
        if (guide_train) {
            // samples from this pass's recording paths, then Adam steps, then the
            // BF16 weights of the next pass
            size_t ta, tb;
            if ((r = mark(&ta))) return r;
            HIP_TRY(hipMemsetAsync(c->g_count, 0, sizeof(uint32_t), c->stream));
            launch_guide_gather(rc, c->pb, c->g_samples, guide_batch, c->g_count, c->g_ring, c->g_gather_blk,
                                c->stream);
            launch_guide_ring_commit(c->g_ring, c->g_count, guide_batch, c->g_samples, c->g_wsum, c->stream);
            for (int k = 0; k < guide_steps; ++k) {
                // matrix-core step: forward with the current BF16 weights, backward, reduce, Adam
                // (which also repacks the BF16 weights)
                launch_guide_train_mfma(c->g_packed, c->g_bias, c->g_samples, c->g_ring + 1, guide_batch, c->g_wsum,
                                        c->g_partial, agx_ctx::GUIDE_TRAIN_BLOCKS, c->stream);
                c->g_steps += 1;
                launch_guide_reduce_adam(c->g_partial, agx_ctx::GUIDE_TRAIN_BLOCKS, c->g_params, c->g_m, c->g_v,
                                         guide_lr, 0.9f, 0.99f, 1e-8f,
                                         int(std::min<uint64_t>(c->g_steps, 1u << 30)), c->stream, c->g_packed,
                                         c->g_bias);
            }
            HIP_TRY(hipMemcpyAsync(&guide_samples_h, c->g_count, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
            HIP_TRY(hipMemcpyAsync(&guide_valid_h, c->g_ring + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
            if ((r = mark(&tb))) return r;
            timed.push_back({ta, tb, 4});
            b = tb;
        }
        HIP_TRY(hipEventSynchronize(c->pool[b]));
        if (guide_train) {
            st.guide_samples += std::min<uint32_t>(guide_samples_h, guide_batch);
            st.guide_trained += uint64_t(guide_valid_h) * uint64_t(guide_steps);
        }
        // sum the ray queue sizes of this pass (every read has completed)
        for (size_t k = first_read; k < reads.size(); ++k) {
            const uint32_t *q = c->h_counters + 4 * reads[k].slot;
            st.closest_rays += q[0];
            st.shadow_rays += q[1];
            if (guide && k + 1 < reads.size()) st.guide_rows += q[2];  // shaded by the next bounce
        }
        // every bounce launched both traversal kernels (an empty queue exits at once)
        st.trace_closest_launches += uint64_t(reads.size() - first_read - 1);
        st.trace_shadow_launches += uint64_t(reads.size() - first_read - 1);
        reads.resize(first_read);
        ring_next = 0;
        if (cancel_flag && *cancel_flag) cancelled = true;
    }
    HIP_TRY(hipEventRecord(c->ev[1], c->stream));
    HIP_TRY(hipEventSynchronize(c->ev[1]));
    st.total_ms = elapsed(c->ev[0], c->ev[1]);
    for (const Timed &t : timed) {
        const double ms = elapsed(c->pool[t.a], c->pool[t.b]);
        if (t.kind == 0) st.trace_closest_ms += ms;
        else if (t.kind == 1) st.trace_shadow_ms += ms;
        else if (t.kind == 3) st.guide_infer_ms += ms;
        else if (t.kind == 4) st.guide_train_ms += ms;
        else st.shade_ms += ms;
    }
    if (stats) {
        DevCounters hc[2]{};
        HIP_TRY(hipMemcpy(hc, c->dcnt, sizeof(hc), hipMemcpyDeviceToHost));
        fill_traversal_stats(hc, &st);
    }
    st.guide_steps = c->g_steps;
    c->stats = st;
    const auto t_device_end = std::chrono::steady_clock::now();
    // normalise into host outputs
    if (rgb || albedo || normal || denoise) {
        // device planes and pinned staging persist across renders (no allocation per call)
        if (size_t(wh) > c->resolve_cap) {
            if (c->d_resolve) (void)hipFree(c->d_resolve);
            if (c->h_resolve) (void)hipHostFree(c->h_resolve);
            c->d_resolve = nullptr;
            c->h_resolve = nullptr;
            c->resolve_cap = 0;
            HIP_TRY(hipMalloc(reinterpret_cast<void **>(&c->d_resolve), size_t(wh) * 10 * sizeof(float)));
            HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&c->h_resolve), size_t(wh) * 10 * sizeof(float),
                                  hipHostMallocDefault));
            c->resolve_cap = size_t(wh);
        }
        float *dout = c->d_resolve, *hout = c->h_resolve;
        launch_film_resolve(c->pb, wh, dout, dout + 3 * wh, dout + 6 * wh, dout + 9 * wh, c->stream);
        hipError_t e = hipSuccess;
        struct Plane { float *user; size_t off, n; };
        const Plane planes[4] = {{rgb, 0, 3}, {albedo, 3, 3}, {normal, 6, 3}, {denoise, 9, 1}};
        for (const Plane &pl : planes)
            if (pl.user && e == hipSuccess)
                e = hipMemcpyAsync(hout + pl.off * size_t(wh), dout + pl.off * size_t(wh),
                                   pl.n * size_t(wh) * sizeof(float), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return fail(AGX_E_HIP, std::string("film readback: ") + hipGetErrorString(e));
        for (const Plane &pl : planes)
            if (pl.user) std::memcpy(pl.user, hout + pl.off * size_t(wh), pl.n * size_t(wh) * sizeof(float));
    }
    const auto t_end = std::chrono::steady_clock::now();
    c->stats.host_ms = std::chrono::duration<double, std::milli>(t_end - t_call).count();
    c->stats.readback_ms = std::chrono::duration<double, std::milli>(t_end - t_device_end).count();
    if (cancelled) return fail(AGX_E_CANCELLED, "render cancelled");
    return AGX_OK;
}

int agx_trace_rays(agx_ctx *c, const float *o, const float *d, const float *tmin, const float *tmax, uint32_t n,
                   int32_t any_hit, int32_t *entity_id, int32_t *prim_id, float *t, float *u, float *v) {
    if (!c || (n && (!o || !d || !tmin || !tmax || !entity_id))) return fail(AGX_E_INVALID_ARG, "null argument");
    if (!c->has_scene) return fail(AGX_E_NO_SCENE, "no scene uploaded");
    if (!n) return AGX_OK;
    HIP_TRY(hipSetDevice(c->device));
    std::vector<float4> ho(n), hd(n);
    for (uint32_t i = 0; i < n; ++i) {
        ho[i] = make_float4(o[3 * i], o[3 * i + 1], o[3 * i + 2], tmin[i]);
        hd[i] = make_float4(d[3 * i], d[3 * i + 1], d[3 * i + 2], tmax[i]);
    }
    char *buf = nullptr;
    const size_t bytes = size_t(n) * (16 + 16 + 4 * 5);
    HIP_TRY(hipMalloc(reinterpret_cast<void **>(&buf), bytes));
    float4 *dO = reinterpret_cast<float4 *>(buf);
    float4 *dD = dO + n;
    int32_t *dE = reinterpret_cast<int32_t *>(dD + n);
    int32_t *dP = dE + n;
    float *dT = reinterpret_cast<float *>(dP + n);
    float *dU = dT + n;
    float *dV = dU + n;
    hipError_t e = hipMemcpyAsync(dO, ho.data(), n * sizeof(float4), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dD, hd.data(), n * sizeof(float4), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(c->dcnt, 0, 2 * sizeof(DevCounters), c->stream);
    if (e == hipSuccess) e = hipEventRecord(c->ev[0], c->stream);
    if (e == hipSuccess)
        launch_trace_rays(c->ds, dO, dD, n, any_hit, dE, dP, dT, dU, dV, c->dcnt + (any_hit ? 1 : 0), c->stream);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(c->ev[1], c->stream);
    std::vector<int32_t> he(n), hp(n);
    std::vector<float> ht(n), hu(n), hv(n);
    if (e == hipSuccess) e = hipMemcpyAsync(he.data(), dE, n * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && !any_hit) {
        e = hipMemcpyAsync(hp.data(), dP, n * 4, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(ht.data(), dT, n * 4, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(hu.data(), dU, n * 4, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(hv.data(), dV, n * 4, hipMemcpyDeviceToHost, c->stream);
    }
    DevCounters hc[2]{};
    if (e == hipSuccess) e = hipMemcpyAsync(hc, c->dcnt, sizeof(hc), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(buf);
    if (e != hipSuccess) return fail(AGX_E_HIP, std::string("trace_rays: ") + hipGetErrorString(e));
    std::memcpy(entity_id, he.data(), n * 4);
    if (!any_hit) {
        if (prim_id) std::memcpy(prim_id, hp.data(), n * 4);
        if (t) std::memcpy(t, ht.data(), n * 4);
        if (u) std::memcpy(u, hu.data(), n * 4);
        if (v) std::memcpy(v, hv.data(), n * 4);
    }
    agx_stats_t st{};
    st.node_bytes = c->accel != AGX_ACCEL_ATRC ? 16 * BVH4_NODE_VEC4 : 32;
    st.prim_bytes = 48;
    st.info_bytes = 96;
    if (any_hit) st.shadow_rays = n; else st.closest_rays = n;
    fill_traversal_stats(hc, &st);
    st.total_ms = elapsed(c->ev[0], c->ev[1]);
    if (any_hit) { st.trace_shadow_ms = st.total_ms; st.trace_shadow_launches = 1; }
    else { st.trace_closest_ms = st.total_ms; st.trace_closest_launches = 1; }
    c->stats = st;
    return AGX_OK;
}

static_assert(AGX_NOALLOC_MAX_RAYS == NOALLOC_MAX_RAYS, "staging capacity");

int agx_trace_rays_noalloc(agx_ctx *c, const float *o, const float *d, const float *tmin, const float *tmax,
                           uint32_t n, int32_t any_hit, int32_t *entity_id, int32_t *prim_id, float *t, float *u,
                           float *v) {
    if (!c || (n && (!o || !d || !tmin || !tmax || !entity_id))) return fail(AGX_E_INVALID_ARG, "null argument");
    if (!c->has_scene) return fail(AGX_E_NO_SCENE, "no scene uploaded");
    if (n > NOALLOC_MAX_RAYS)
        return fail(AGX_E_LIMIT, "agx_trace_rays_noalloc takes at most " + std::to_string(NOALLOC_MAX_RAYS) +
                                     " rays per call (" + std::to_string(n) + " given)");
    if (!n) return AGX_OK;
    HIP_TRY(hipSetDevice(c->device));
    char *base = static_cast<char *>(noalloc_scratch(c->device));
    if (!base) return fail(AGX_E_HIP, "staging array not resolved on this device");
    const size_t M = NOALLOC_MAX_RAYS;
    auto word = [&](size_t k) { return base + 4 * M * k; };  // layout: launch_trace_rays_staged
    hipError_t e = hipMemcpyAsync(word(8), o, size_t(n) * 12, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(word(11), d, size_t(n) * 12, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(word(14), tmin, size_t(n) * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(word(15), tmax, size_t(n) * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipEventRecord(c->ev[0], c->stream);
    if (e == hipSuccess) launch_trace_rays_staged(c->ds, base, n, any_hit, c->dcnt, c->stream);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(c->ev[1], c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(entity_id, word(16), size_t(n) * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && !any_hit) {
        if (prim_id) e = hipMemcpyAsync(prim_id, word(17), size_t(n) * 4, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess && t) e = hipMemcpyAsync(t, word(18), size_t(n) * 4, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess && u) e = hipMemcpyAsync(u, word(19), size_t(n) * 4, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess && v) e = hipMemcpyAsync(v, word(20), size_t(n) * 4, hipMemcpyDeviceToHost, c->stream);
    }
    DevCounters hc[2]{};
    if (e == hipSuccess) e = hipMemcpyAsync(hc, c->dcnt, sizeof(hc), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(AGX_E_HIP, std::string("trace_rays_noalloc: ") + hipGetErrorString(e));
    agx_stats_t st{};
    st.node_bytes = c->accel != AGX_ACCEL_ATRC ? 16 * BVH4_NODE_VEC4 : 32;
    st.prim_bytes = 48;
    st.info_bytes = 96;
    if (any_hit) st.shadow_rays = n; else st.closest_rays = n;
    fill_traversal_stats(hc, &st);
    st.total_ms = elapsed(c->ev[0], c->ev[1]);
    if (any_hit) { st.trace_shadow_ms = st.total_ms; st.trace_shadow_launches = 1; }
    else { st.trace_closest_ms = st.total_ms; st.trace_closest_launches = 1; }
    c->stats = st;
    return AGX_OK;
}

int agx_get_stats(agx_ctx *c, agx_stats_t *out) {
    if (!c || !out) return fail(AGX_E_INVALID_ARG, "null argument");
    *out = c->stats;
    return AGX_OK;
}

// ---------------------------------------------------------------- batched rendering
}  // extern "C"

namespace {

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}
// AGX_BATCH_TRACE=1: the graph batch prints the time of each phase to stderr
bool batch_trace() {
    static const bool on = [] {
        const char *e = std::getenv("AGX_BATCH_TRACE");
        return e && e[0] == '1';
    }();
    return on;
}

// Renders items[order[0..m)] with k worker contexts (each worker thread takes
// the next item, builds + uploads it, renders it).  spp_override > 0 replaces
// the items' spp (autotuner probes; images are then not written).  Returns the
// first failing status; *first_err receives its message.
int run_batch(agx_ctx *c, agx_batch_item *items, const std::vector<uint32_t> &order, int k, int spp_override,
              uint64_t *rays_out, std::string *first_err) {
    std::atomic<size_t> next{0};
    std::atomic<uint64_t> rays{0};
    std::mutex mu;
    int status = AGX_OK;
    auto work = [&](int wi) {
        agx_ctx *w = c->batch_workers[size_t(wi)];
        for (;;) {
            const size_t j = next.fetch_add(1);
            if (j >= order.size()) break;
            agx_batch_item &it = items[order[j]];
            agx_render_params p = *it.params;
            if (spp_override > 0) p.spp = std::min(p.spp, spp_override);
            const auto t0 = Clock::now();
            int r = agx_upload_scene(w, it.scene);
            const double b_ms = ms_since(t0);
            const auto t1 = Clock::now();
            if (r == AGX_OK) r = agx_render(w, &p, nullptr, spp_override > 0 ? nullptr : it.rgb, nullptr, nullptr, nullptr);
            const double r_ms = ms_since(t1);
            const uint64_t nr = r == AGX_OK ? w->stats.closest_rays + w->stats.shadow_rays : 0;
            rays += nr;
            if (spp_override <= 0) {
                it.status = r;
                it.worker = wi;
                it.rays = nr;
                it.build_ms = b_ms;
                it.render_ms = r_ms;
            }
            if (r != AGX_OK) {
                std::lock_guard<std::mutex> lk(mu);
                if (status == AGX_OK) {
                    status = r;
                    *first_err = "batch item " + std::to_string(order[j]) + ": " + agx_last_error();
                }
            }
        }
    };
    std::vector<std::thread> threads;
    for (int wi = 1; wi < k; ++wi) threads.emplace_back(work, wi);
    work(0);
    for (auto &t : threads) t.join();
    *rays_out = rays.load();
    return status;
}


// ---------------------------------------------------------------- graph-captured batch
// Paths per wavefront pass of a render without guiding (agx_render's sizing).
int64_t plain_pass_size(agx_ctx *c, const RenderConst &rc, int64_t layer, int spp) {
    if (c->pass_paths_L != rc.L) {
        c->pass_paths = max_paths_in_flight(rc.L);
        if (c->pass_paths_limit) c->pass_paths = std::min(c->pass_paths, c->pass_paths_limit);
        c->pass_paths_L = rc.L;
    }
    const int64_t cap = int64_t(c->pass_paths);
    if (layer <= cap) return layer * std::min<int64_t>(spp, cap / layer);
    return rc.splat ? layer : cap;
}

// Allocations a captured render needs (path buffers, resolve planes), made
// before the capture: nothing may allocate inside it.
int graph_prepare(agx_ctx *c, const agx_render_params *p) {
    RenderConst rc{};
    if (int r = fill_render_const(c, p, &rc); r != AGX_OK) return r;
    const int64_t wh = int64_t(p->width) * p->height;
    const int64_t layer = rc.splat ? int64_t(rc.EW) * rc.EH : wh;
    const int64_t per_pass = plain_pass_size(c, rc, layer, p->spp);
    if (int r = c->ensure_wave(size_t(per_pass), rc.L, wh); r != AGX_OK) return r;
    if (size_t(wh) > c->resolve_cap) {
        if (c->d_resolve) (void)hipFree(c->d_resolve);
        if (c->h_resolve) (void)hipHostFree(c->h_resolve);
        c->d_resolve = nullptr;
        c->h_resolve = nullptr;
        c->resolve_cap = 0;
        HIP_TRY(hipMalloc(reinterpret_cast<void **>(&c->d_resolve), size_t(wh) * 10 * sizeof(float)));
        HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&c->h_resolve), size_t(wh) * 10 * sizeof(float),
                              hipHostMallocDefault));
        c->resolve_cap = size_t(wh);
    }
    return AGX_OK;
}

// One scene's whole render enqueued on a stream with no host wait, for capture:
// every bounce up to the depth bound is launched (a kernel whose queue is empty
// exits at once), the rays of each bounce are summed on the device into *rays,
// and the image is resolved into the context's device planes.  No memset, no
// allocation, no host synchronisation.
int enqueue_render(agx_ctx *c, const agx_render_params *p, hipStream_t st, unsigned long long *rays) {
    RenderConst rc{};
    if (int r = fill_render_const(c, p, &rc); r != AGX_OK) return r;
    const int64_t wh = int64_t(p->width) * p->height;
    const int64_t layer = rc.splat ? int64_t(rc.EW) * rc.EH : wh;
    const int64_t total = layer * p->spp;
    const int64_t per_pass = plain_pass_size(c, rc, layer, p->spp);
    if (size_t(per_pass) > c->cap_paths || std::max(rc.L, 1) > c->cap_L || wh > c->cap_wh ||
        size_t(wh) > c->resolve_cap)
        return fail(AGX_E_INVALID_ARG, "graph batch: render buffers not prepared");
    rc.guide_rec_shift = guide_rec_shift(size_t(per_pass));
    const uint32_t n_closest_max = uint32_t(2 * per_pass), n_shadow_max = uint32_t(per_pass * std::max(rc.L, 1));
    // without delta surfaces no bounce extends the depth (path_tracing.cpp:155-162),
    // so the specular allowance adds only empty launches
    const int spec = (c->ds.features & SF_DELTA) ? std::max(p->specular_depth, 0) : 0;
    const int max_bounces = p->max_depth + spec + 2;
    PathBuffers pb = c->pb;
    launch_film_clear(pb, wh, st);
    for (int64_t base = 0; base < total; base += per_pass) {
        rc.work_base = base;
        rc.n_paths = int32_t(std::min<int64_t>(per_pass, total - base));
        uint32_t *cin = c->counters, *cout = c->counters + 4;
        launch_reset_counters(cout, st);
        pb.cnt_out = cout;
        launch_raygen(c->ds, rc, pb, st);
        std::swap(cin, cout);
        std::swap(pb.q_in, pb.q_out);
        for (int bounce = 0; bounce < max_bounces; ++bounce) {
            pb.cnt_in = cin;
            pb.cnt_out = cout;
            launch_count_rays(cin, rays, st);
            launch_trace_closest(c->ds, pb, n_closest_max, false, c->dcnt, st);
            launch_trace_shadow(c->ds, pb, n_shadow_max, false, c->dcnt + 1, st);
            launch_reset_counters(cout, st);
            launch_shade(c->ds, rc, pb, uint32_t(rc.n_paths), st);
            launch_queue_compact(pb, uint32_t(rc.n_paths), rc.L, c->qc_blk, st);
            std::swap(cin, cout);
            std::swap(pb.q_in, pb.q_out);
        }
        launch_film(rc, pb, st);
    }
    float *dout = c->d_resolve;
    launch_film_resolve(pb, wh, dout, dout + 3 * wh, dout + 6 * wh, dout + 9 * wh, st);
    return AGX_OK;
}

}  // namespace

extern "C" {

int agx_render_batch(agx_ctx *c, agx_batch_item *items, uint32_t n, int32_t workers, agx_batch_report *report) {
    if (!c || (n && !items)) return fail(AGX_E_INVALID_ARG, "null argument");
    if (workers < 0 || workers > 8) return fail(AGX_E_INVALID_ARG, "workers must lie in [0, 8]");
    for (uint32_t i = 0; i < n; ++i) {
        if (!items[i].scene || !items[i].params) return fail(AGX_E_INVALID_ARG, "batch item without scene or params");
        if (items[i].params->guide) return fail(AGX_E_UNSUPPORTED, "guided rendering is not available in batches");
        items[i].status = AGX_E_CANCELLED;
        items[i].worker = -1;
        items[i].rays = 0;
        items[i].build_ms = items[i].render_ms = 0;
    }
    agx_batch_report rep{};
    const int need = workers > 0 ? workers : 4;
    while (int(c->batch_workers.size()) < need) {
        agx_ctx *w = nullptr;
        const int r = agx_create(c->device, &w);
        if (r != AGX_OK) return r;
        // several scenes in flight share the chip: passes of 4 M paths keep each
        // worker's buffers small and its allocations few as scene sizes vary
        w->pass_paths_limit = size_t(1) << 22;
        c->batch_workers.push_back(w);
    }
    std::string err;
    int k = workers;
    if (k == 0) {
        // probe: the (up to) 4 cheapest items at 2 spp with 1, 2 and 4 workers;
        // cost estimate = triangles + pixels * spp
        std::vector<uint32_t> idx(n);
        for (uint32_t i = 0; i < n; ++i) idx[i] = i;
        auto cost = [&](uint32_t i) {
            uint64_t tris = 0;
            const agx_scene_desc &d = *items[i].scene;
            for (uint32_t e = 0; e < d.entity_count; ++e) tris += d.entities[e].tri_count;
            const agx_render_params &p = *items[i].params;
            return double(tris) + double(p.width) * double(p.height) * double(p.spp);
        };
        std::sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return cost(a) < cost(b); });
        idx.resize(std::min<size_t>(idx.size(), 4));
        double best = -1;
        k = 1;
        for (int kk : {1, 2, 4}) {
            if (kk > 1 && size_t(kk) > idx.size()) break;
            uint64_t rays = 0;
            const auto t0 = Clock::now();
            const int r = run_batch(c, items, idx, kk, 2, &rays, &err);
            const double ms = ms_since(t0);
            if (r != AGX_OK) return fail(r, err);
            const double rate = rays / std::max(ms * 1e-3, 1e-9);
            rep.tune_rays_per_s[kk - 1] = rate;
            if (rate > best) { best = rate; k = kk; }
        }
        rep.tuned = 1;
    }
    std::vector<uint32_t> all(n);
    for (uint32_t i = 0; i < n; ++i) all[i] = i;
    uint64_t rays = 0;
    const auto t0 = Clock::now();
    const int r = run_batch(c, items, all, k, 0, &rays, &err);
    rep.wall_ms = ms_since(t0);
    rep.workers = k;
    rep.rays = rays;
    if (report) *report = rep;
    if (r != AGX_OK) return fail(r, err);
    return AGX_OK;
}

int agx_render_batch_graph(agx_ctx *c, agx_batch_item *items, uint32_t n, int32_t branches,
                           agx_batch_report *report) {
    if (!c || (n && !items)) return fail(AGX_E_INVALID_ARG, "null argument");
    if (branches < 0 || branches > 8) return fail(AGX_E_INVALID_ARG, "branches must lie in [0, 8]");
    for (uint32_t i = 0; i < n; ++i) {
        if (!items[i].scene || !items[i].params) return fail(AGX_E_INVALID_ARG, "batch item without scene or params");
        if (items[i].params->guide) return fail(AGX_E_UNSUPPORTED, "guided rendering is not available in batches");
        items[i].status = AGX_E_CANCELLED;
        items[i].worker = -1;
        items[i].rays = 0;
        items[i].build_ms = items[i].render_ms = 0;
    }
    HIP_TRY(hipSetDevice(c->device));
    const auto t0 = Clock::now();
    agx_batch_report rep{};
    // cost estimate of an item: paths it traces plus a BVH term
    auto cost = [&](uint32_t i) {
        const agx_render_params *p = items[i].params;
        uint64_t tris = 0;
        for (uint32_t e = 0; e < items[i].scene->entity_count; ++e) tris += items[i].scene->entities[e].tri_count;
        return double(p->width) * p->height * p->spp + 4.0 * double(items[i].scene->entity_count) + double(tris);
    };
    // smallest builds first: the first group builds quickly and the GPU starts
    // while the larger scenes are still being built.  The order ignores the
    // render cost, so the group left to render after the last build costs
    // what an average group does, not the most.
    auto build_cost = [&](uint32_t i) {
        uint64_t tris = 0;
        for (uint32_t e = 0; e < items[i].scene->entity_count; ++e) tris += items[i].scene->entities[e].tri_count;
        return double(tris) + 4.0 * double(items[i].scene->entity_count);
    };
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return build_cost(a) < build_cost(b); });
    const uint32_t group_size = 8;
    const uint32_t n_groups = (n + group_size - 1) / group_size;

    // everything the graphs touch, released on every exit
    struct Held {
        std::vector<agx_ctx *> ctxs;
        std::vector<hipStream_t> streams;
        std::vector<hipEvent_t> fork, join, start, done, copied;
        float *stage[2] = {nullptr, nullptr};  // pinned host images of two groups
        unsigned long long *rays_pin[2] = {nullptr, nullptr};  // pinned ray counters of two groups
        unsigned long long *rays = nullptr;
        std::vector<hipGraphExec_t> execs;
        std::vector<std::thread> builders;
        std::atomic<bool> stop{false};
        agx_ctx *owner = nullptr;  // the calling context
        bool keep = false;         // the call succeeded: ctxs go back to owner->graph_pool
        ~Held() {
            stop = true;
            for (auto &t : builders) if (t.joinable()) t.join();
            for (hipGraphExec_t e : execs) if (e) (void)hipGraphExecDestroy(e);
            for (float *st : stage) if (st) (void)hipHostFree(st);
            for (unsigned long long *rp : rays_pin) if (rp) (void)hipHostFree(rp);
            for (auto *v : {&fork, &join, &start, &done, &copied})
                for (hipEvent_t e : *v) if (e) (void)hipEventDestroy(e);
            for (hipStream_t st : streams) (void)hipStreamDestroy(st);
            if (rays) (void)hipFree(rays);
            for (agx_ctx *w : ctxs) {
                if (!w) continue;
                if (keep) owner->graph_pool.push_back(w);
                else agx_destroy(w);
            }
        }
    } held;
    held.owner = c;
    // SETS sets of group_size contexts, used by the groups in turn: a group's
    // contexts take the scenes of group g + SETS once it has finished, so path
    // buffers are allocated once and grow with the scenes instead of being
    // allocated and freed per scene
    std::vector<uint32_t> pos(n);
    for (uint32_t j = 0; j < n; ++j) pos[order[j]] = j;
    constexpr uint32_t SETS = 3;
    const uint32_t pool_size = std::min(n, SETS * group_size);
    held.ctxs.assign(pool_size, nullptr);
    for (auto &w : held.ctxs) {
        // the contexts the last call left, then new ones (their pass limits,
        // block sizes and scenes are set per launch, as for reuse within a call)
        if (!c->graph_pool.empty()) {
            w = c->graph_pool.back();
            c->graph_pool.pop_back();
        } else if (int r = agx_create(c->device, &w); r != AGX_OK) {
            return r;
        }
    }
    auto ctx_of = [&](uint32_t i) {
        return held.ctxs[((pos[i] / group_size) % SETS) * group_size + pos[i] % group_size];
    };
    HIP_TRY(hipMalloc(reinterpret_cast<void **>(&held.rays), std::max<size_t>(n, 1) * sizeof(unsigned long long)));
    HIP_TRY(hipMemset(held.rays, 0, std::max<size_t>(n, 1) * sizeof(unsigned long long)));
    HIP_TRY(hipDeviceSynchronize());

    // host builds on every host core, in group order; a built scene is uploaded
    // into its pooled context once the group that last used that context has
    // finished (groups finish in order)
    std::vector<std::atomic<int>> built(n);  // 0 pending, 1 built, 2 failed
    for (auto &b : built) b = 0;
    std::mutex mu;
    std::condition_variable cv;
    std::atomic<uint32_t> next{0};
    std::atomic<int64_t> build_end_ns{0};
    int64_t finished_upto = -1;  // highest finished group (guarded by mu)
    // one builder per core (the main thread mostly waits on events)
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int n_threads = int(std::max(1u, std::min({hw, 16u, std::max(n, 1u)})));
    std::string build_err;
    auto build = [&]() {
        for (;;) {
            if (held.stop) break;
            const uint32_t j = next.fetch_add(1);
            if (j >= n) break;
            const uint32_t i = order[j];
            const auto tb = Clock::now();
            HostScene hs;
            std::string err;
            int r = build_host_scene(*items[i].scene, &hs, &err);
            if (r == AGX_OK) {
                const int64_t g = int64_t(j / group_size);
                std::unique_lock<std::mutex> lk(mu);
                while (!held.stop && g >= int64_t(SETS) && finished_upto < g - int64_t(SETS))
                    cv.wait_for(lk, std::chrono::milliseconds(5));
                lk.unlock();
                if (held.stop) break;
                r = upload_host_scene(ctx_of(i), hs);
                if (r != AGX_OK) err = agx_last_error();
            }
            items[i].build_ms = ms_since(tb);
            items[i].status = r;
            {
                std::lock_guard<std::mutex> lk(mu);
                if (r != AGX_OK && build_err.empty()) build_err = "batch item " + std::to_string(i) + ": " + err;
                built[i] = r == AGX_OK ? 1 : 2;
                build_end_ns = std::chrono::duration_cast<std::chrono::nanoseconds>(Clock::now() - t0).count();
            }
            cv.notify_all();
        }
    };
    for (int t = 0; t < n_threads; ++t) held.builders.emplace_back(build);
    auto wait_group = [&](uint32_t g) -> int {
        std::unique_lock<std::mutex> lk(mu);
        const uint32_t a = g * group_size, b = std::min(n, a + group_size);
        cv.wait(lk, [&] {
            for (uint32_t j = a; j < b; ++j) if (built[order[j]] == 0) return false;
            return true;
        });
        for (uint32_t j = a; j < b; ++j)
            if (built[order[j]] == 2) return fail(items[order[j]].status, build_err);
        return AGX_OK;
    };

    // per-branch streams and events (branches at most 8)
    held.streams.assign(8, nullptr);
    for (auto &st : held.streams) HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    held.fork.assign(n_groups + 64, nullptr);
    held.join.assign(size_t(n_groups + 64) * 8, nullptr);
    held.start.assign(n_groups + 64, nullptr);
    held.done.assign(n_groups + 64, nullptr);
    held.copied.assign(n_groups + 64, nullptr);
    for (auto *v : {&held.fork, &held.join})
        for (auto &e : *v) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (auto *v : {&held.start, &held.done})
        for (auto &e : *v) HIP_TRY(hipEventCreate(&e));
    for (auto &e : held.copied) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    // pinned staging for the images of two groups: a group's images are copied
    // out on the stream right after its graph, and to the callers' buffers on
    // the host once the group is done
    size_t stage_floats = 0;
    for (uint32_t g = 0; g < n_groups; ++g) {
        size_t f = 0;
        for (uint32_t j = g * group_size; j < std::min(n, (g + 1) * group_size); ++j)
            f += 3 * size_t(items[order[j]].params->width) * size_t(items[order[j]].params->height);
        stage_floats = std::max(stage_floats, f);
    }
    for (float *&st : held.stage)
        HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&st), std::max<size_t>(stage_floats, 1) * sizeof(float),
                              hipHostMallocDefault));
    for (unsigned long long *&rp : held.rays_pin)
        HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&rp), std::max<size_t>(n, 1) * sizeof(unsigned long long),
                              hipHostMallocDefault));
    size_t ev = 0;  // next free event slot

    struct Config {
        int branches;
        int pass_log2;
        int trace_block;
    };
    // Capture the renders of items sel with configuration cfg into one graph
    // (costliest first, each onto the least loaded branch; a branch renders its
    // scenes one after another), launch it on the context's stream between
    // timing events.  spp_override > 0 renders the items at that spp (probes).
    auto launch = [&](const std::vector<uint32_t> &sel, const Config &cfg, int spp_override, size_t *slot) -> int {
        std::vector<agx_render_params> prm(sel.size());
        for (size_t q = 0; q < sel.size(); ++q) {
            agx_ctx *w = ctx_of(sel[q]);
            prm[q] = *items[sel[q]].params;
            if (spp_override > 0) prm[q].spp = std::min(prm[q].spp, spp_override);
            if (w->pass_paths_limit != (size_t(1) << cfg.pass_log2)) {
                w->pass_paths_limit = size_t(1) << cfg.pass_log2;
                w->pass_paths_L = -1;
            }
            w->ds.trace_block = cfg.trace_block;
            if (int r = graph_prepare(w, &prm[q]); r != AGX_OK) return r;
        }
        std::vector<size_t> by_cost(sel.size());
        for (size_t q = 0; q < sel.size(); ++q) by_cost[q] = q;
        std::stable_sort(by_cost.begin(), by_cost.end(), [&](size_t a, size_t b) { return cost(sel[a]) > cost(sel[b]); });
        const int k = cfg.branches;
        std::vector<std::vector<size_t>> lanes(static_cast<size_t>(k));
        std::vector<double> load(static_cast<size_t>(k), 0.0);
        for (size_t q : by_cost) {
            const size_t b = size_t(std::min_element(load.begin(), load.end()) - load.begin());
            lanes[b].push_back(q);
            load[b] += cost(sel[q]);
            if (spp_override <= 0) items[sel[q]].worker = int32_t(b);
        }
        if (ev >= held.start.size()) return fail(AGX_E_LIMIT, "graph batch: out of events");
        const size_t e = ev++;
        *slot = e;
        hipGraph_t graph = nullptr;
        HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
        int cap_status = AGX_OK;
        hipError_t he = hipEventRecord(held.fork[e], c->stream);
        for (int b = 0; b < k && he == hipSuccess; ++b) he = hipStreamWaitEvent(held.streams[size_t(b)], held.fork[e], 0);
        for (int b = 0; b < k && he == hipSuccess && cap_status == AGX_OK; ++b)
            for (size_t q : lanes[size_t(b)]) {
                cap_status = enqueue_render(ctx_of(sel[q]), &prm[q], held.streams[size_t(b)], held.rays + sel[q]);
                if (cap_status != AGX_OK) break;
            }
        for (int b = 0; b < k && he == hipSuccess; ++b) {
            he = hipEventRecord(held.join[e * 8 + size_t(b)], held.streams[size_t(b)]);
            if (he == hipSuccess) he = hipStreamWaitEvent(c->stream, held.join[e * 8 + size_t(b)], 0);
        }
        const hipError_t he_end = hipStreamEndCapture(c->stream, &graph);
        if (cap_status != AGX_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return cap_status;  // the message is set
        }
        if (he != hipSuccess || he_end != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return fail(AGX_E_HIP, std::string("graph capture: ") + hipGetErrorString(he != hipSuccess ? he : he_end));
        }
        hipGraphExec_t exec = nullptr;
        const hipError_t hi = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (hi != hipSuccess) return fail(AGX_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(hi));
        held.execs.push_back(exec);
        HIP_TRY(hipEventRecord(held.start[e], c->stream));
        HIP_TRY(hipGraphLaunch(exec, c->stream));
        HIP_TRY(hipEventRecord(held.done[e], c->stream));
        return AGX_OK;
    };
    auto group_items = [&](uint32_t g) {
        std::vector<uint32_t> v;
        for (uint32_t j = g * group_size; j < std::min(n, (g + 1) * group_size); ++j) v.push_back(order[j]);
        return v;
    };

    auto trace = [&](const char *what, uint32_t g) {
        if (batch_trace()) std::fprintf(stderr, "[batch graph] %8.1f ms  %s %u\n", ms_since(t0), what, g);
    };
    Config cfg{branches > 0 ? branches : 4, 22, 128};
    if (branches == 0 && n > 0) {
        // autotuner: coordinate descent over branches, pass size and traversal block
        // size, each configuration timed on the first group's scenes at 2 spp
        // (device time between events); later groups build meanwhile
        if (int r = wait_group(0); r != AGX_OK) return r;
        trace("first group built", 0);
        const std::vector<uint32_t> probe = group_items(0);
        auto time_cfg = [&](const Config &k, double *rate) -> int {
            HIP_TRY(hipMemset(held.rays, 0, std::max<size_t>(n, 1) * sizeof(unsigned long long)));
            size_t slot = 0;
            if (int r = launch(probe, k, 2, &slot); r != AGX_OK) return r;
            HIP_TRY(hipEventSynchronize(held.done[slot]));
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, held.start[slot], held.done[slot]));
            std::vector<unsigned long long> hr(n, 0ull);
            HIP_TRY(hipMemcpy(hr.data(), held.rays, n * sizeof(unsigned long long), hipMemcpyDeviceToHost));
            uint64_t rays = 0;
            for (uint32_t i : probe) rays += hr[i];
            *rate = double(rays) / std::max(double(ms) * 1e-3, 1e-9);
            ++rep.tune_configs;
            trace("probe timed", uint32_t(rep.tune_configs));
            return AGX_OK;
        };
        double best = -1;
        // warm-up launch (first-use costs), not compared
        if (int r = time_cfg(cfg, &best); r != AGX_OK) return r;
        best = -1;
        for (int b : {2, 4, 8}) {
            double rate = 0;
            const Config k{b, cfg.pass_log2, cfg.trace_block};
            if (int r = time_cfg(k, &rate); r != AGX_OK) return r;
            rep.tune_rays_per_s[b == 2 ? 0 : (b == 4 ? 1 : 2)] = rate;
            if (rate > best) { best = rate; cfg.branches = b; }
        }
        for (int pl : {21, 23}) {
            double rate = 0;
            const Config k{cfg.branches, pl, cfg.trace_block};
            if (int r = time_cfg(k, &rate); r != AGX_OK) return r;
            if (rate > best * 1.02) { best = rate; cfg.pass_log2 = pl; }
        }
        {
            double rate = 0;
            const Config k{cfg.branches, cfg.pass_log2, 256};
            if (int r = time_cfg(k, &rate); r != AGX_OK) return r;
            rep.tune_rays_per_s[3] = rate;
            if (rate > best * 1.02) { best = rate; cfg.trace_block = 256; }
        }
        rep.tuned = 1;
        HIP_TRY(hipMemset(held.rays, 0, std::max<size_t>(n, 1) * sizeof(unsigned long long)));
        HIP_TRY(hipDeviceSynchronize());
    }

    // the batch: group after group, each captured and launched as soon as its
    // scenes are built; a finished group's images are copied out and its
    // contexts released while later groups render
    const auto tr = Clock::now();
    std::vector<size_t> slots(n_groups, 0);
    std::vector<uint64_t> hr(std::max<size_t>(n, 1), 0);
    // a launched group's images and ray counts to pinned host memory, on the stream
    auto copy_out = [&](uint32_t g) -> int {
        size_t off = 0;
        for (uint32_t i : group_items(g)) {
            const size_t f = 3 * size_t(items[i].params->width) * size_t(items[i].params->height);
            HIP_TRY(hipMemcpyAsync(held.stage[g % 2] + off, ctx_of(i)->d_resolve, f * sizeof(float),
                                   hipMemcpyDeviceToHost, c->stream));
            off += f;
        }
        HIP_TRY(hipMemcpyAsync(held.rays_pin[g % 2], held.rays, size_t(n) * sizeof(unsigned long long),
                               hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipEventRecord(held.copied[slots[g]], c->stream));
        return AGX_OK;
    };
    auto finish_group = [&](uint32_t g) -> int {
        HIP_TRY(hipEventSynchronize(held.done[slots[g]]));
        trace("group done on the device", g);
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, held.start[slots[g]], held.done[slots[g]]));
        HIP_TRY(hipEventSynchronize(held.copied[slots[g]]));
        {   // the images to the callers' buffers, one thread per image (their first
            // touch of fresh pages dominates the copy)
            std::vector<std::thread> cp;
            size_t off = 0;
            for (uint32_t i : group_items(g)) {
                const size_t f = 3 * size_t(items[i].params->width) * size_t(items[i].params->height);
                if (items[i].rgb)
                    cp.emplace_back([dst = items[i].rgb, src = held.stage[g % 2] + off, f]() {
                        std::memcpy(dst, src, f * sizeof(float));
                    });
                off += f;
            }
            for (auto &t : cp) t.join();
        }
        for (uint32_t i : group_items(g)) {
            const unsigned long long r = held.rays_pin[g % 2][i];
            items[i].rays = r;
            items[i].render_ms = ms;
            items[i].status = AGX_OK;
            rep.rays += r;
        }
        {   // its contexts may now take the scenes of group g + SETS
            std::lock_guard<std::mutex> lk(mu);
            finished_upto = int64_t(g);
        }
        cv.notify_all();
        return AGX_OK;
    };
    trace("tuned; rendering starts", 0);
    for (uint32_t g = 0; g < n_groups; ++g) {
        if (int r = wait_group(g); r != AGX_OK) return r;
        trace("group built", g);
        if (int r = launch(group_items(g), cfg, 0, &slots[g]); r != AGX_OK) return r;
        if (int r = copy_out(g); r != AGX_OK) return r;
        trace("group launched", g);
        if (g >= 1) {
            if (int r = finish_group(g - 1); r != AGX_OK) return r;
            trace("group finished", g - 1);
        }
    }
    if (n_groups) {
        if (int r = finish_group(n_groups - 1); r != AGX_OK) return r;
        trace("group finished", n_groups - 1);
    }
    rep.render_ms = ms_since(tr);
    for (auto &t : held.builders) t.join();
    held.builders.clear();
    rep.workers = cfg.branches;
    rep.pass_paths_log2 = cfg.pass_log2;
    rep.trace_block = cfg.trace_block;
    rep.groups = int32_t(n_groups);
    rep.build_threads = n_threads;
    rep.build_wall_ms = double(build_end_ns.load()) * 1e-6;
    rep.wall_ms = ms_since(t0);
    if (report) *report = rep;
    held.keep = true;
    return AGX_OK;
}

// ---------------------------------------------------------------- host-only test hooks
int agx_debug_triangle_bvh(const agx_vertex *tris, uint32_t n, uint32_t *node_count, uint32_t *max_depth,
                           float *nodes_out, float *prims_out) {
    if (!tris || !n || !node_count) return fail(AGX_E_INVALID_ARG, "null argument");
    std::vector<vec4> nodes, prims, info;
    double sa = 0;
    const int depth = build_triangle_bvh(tris, n, &nodes, &prims, &info, &sa);
    *node_count = uint32_t(nodes.size() / 2);
    if (max_depth) *max_depth = uint32_t(depth);
    if (nodes_out) std::memcpy(nodes_out, nodes.data(), nodes.size() * sizeof(vec4));
    if (prims_out) std::memcpy(prims_out, prims.data(), prims.size() * sizeof(vec4));
    return AGX_OK;
}

int agx_debug_copy_kernel(agx_ctx *c, uint64_t bytes, double *ms) {
    if (!c || bytes < 16) return fail(AGX_E_INVALID_ARG, "invalid argument");
    HIP_TRY(hipSetDevice(c->device));
    const uint64_t n = bytes / 16;
    float4 *buf = nullptr;
    HIP_TRY(hipMalloc(reinterpret_cast<void **>(&buf), 2 * n * sizeof(float4)));
    // one untimed launch (first-touch of the pages, code object load), then the
    // average of REPS timed launches
    constexpr int REPS = 5;
    hipError_t e = hipMemsetAsync(buf, 0, n * sizeof(float4), c->stream);
    if (e == hipSuccess) launch_copy_calibration(buf, buf + n, n, c->stream);
    if (e == hipSuccess) e = hipEventRecord(c->ev[0], c->stream);
    for (int i = 0; i < REPS && e == hipSuccess; ++i) launch_copy_calibration(buf, buf + n, n, c->stream);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(c->ev[1], c->stream);
    if (e == hipSuccess) e = hipEventSynchronize(c->ev[1]);
    if (e == hipSuccess && ms) *ms = elapsed(c->ev[0], c->ev[1]) / REPS;
    (void)hipFree(buf);
    if (e != hipSuccess) return fail(AGX_E_HIP, std::string("copy kernel: ") + hipGetErrorString(e));
    return AGX_OK;
}

int agx_debug_replay_rays(agx_ctx *c, const agx_render_params *p, uint32_t max_rays, int reps, uint64_t out[2],
                          double out_ms[4]) {
    if (!c || !p || !out || !out_ms || max_rays == 0 || reps < 1) return fail(AGX_E_INVALID_ARG, "invalid argument");
    if (!c->has_scene) return fail(AGX_E_NO_SCENE, "no scene uploaded");
    HIP_TRY(hipSetDevice(c->device));
    DevBuf rays, cnt, res;
    HIP_TRY(hipMalloc(&rays.p, size_t(max_rays) * 4 * sizeof(float4)));
    HIP_TRY(hipMalloc(&cnt.p, 2 * sizeof(uint32_t)));
    HIP_TRY(hipMalloc(&res.p, size_t(max_rays) * 5 * sizeof(float)));
    HIP_TRY(hipMemsetAsync(cnt.p, 0, 2 * sizeof(uint32_t), c->stream));
    float4 *base = rays.get<float4>();
    c->capture_o[0] = base;
    c->capture_d[0] = base + max_rays;
    c->capture_o[1] = base + 2 * size_t(max_rays);
    c->capture_d[1] = base + 3 * size_t(max_rays);
    c->capture_count = cnt.get<uint32_t>();
    c->capture_cap = max_rays;
    const int r = agx_render(c, p, nullptr, nullptr, nullptr, nullptr, nullptr);
    c->capture_count = nullptr;
    c->capture_o[0] = c->capture_o[1] = c->capture_d[0] = c->capture_d[1] = nullptr;
    c->capture_cap = 0;
    if (r != AGX_OK) return r;
    const agx_stats_t st = c->stats;
    uint32_t n[2];
    HIP_TRY(hipMemcpy(n, cnt.p, sizeof(n), hipMemcpyDeviceToHost));
    if (n[0] > max_rays || n[1] > max_rays)
        return fail(AGX_E_LIMIT, "more rays than max_rays (" + std::to_string(std::max(n[0], n[1])) + ")");
    int32_t *ent = res.get<int32_t>(), *prim = ent + max_rays;
    float *t = reinterpret_cast<float *>(prim + max_rays), *u = t + max_rays, *v = u + max_rays;
    for (int which = 0; which < 2; ++which) {
        double ms = 0;
        for (int k = 0; k <= reps; ++k) {
            HIP_TRY(hipEventRecord(c->ev[0], c->stream));
            launch_trace_rays(c->ds, base + 2 * size_t(which) * max_rays, base + (2 * size_t(which) + 1) * max_rays,
                              n[which], which, ent, prim, t, u, v, nullptr, c->stream);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipEventRecord(c->ev[1], c->stream));
            HIP_TRY(hipEventSynchronize(c->ev[1]));
            if (k > 0) ms += elapsed(c->ev[0], c->ev[1]);
        }
        out[which] = n[which];
        out_ms[which] = ms / reps;
    }
    out_ms[2] = st.trace_closest_ms;
    out_ms[3] = st.trace_shadow_ms;
    c->stats = st;
    return AGX_OK;
}

int agx_debug_scene_summary(const agx_scene_desc *desc, agx_debug_summary *out) {
    if (!desc || !out) return fail(AGX_E_INVALID_ARG, "null argument");
    HostScene hs;
    std::string err;
    const int rc = build_host_scene(*desc, &hs, &err);
    if (rc != AGX_OK) return fail(rc, err);
    out->entities = uint32_t(hs.ents.size());
    out->lights = uint32_t(hs.lights.size());
    out->blas_nodes = uint32_t(hs.blas_nodes.size() / 2);
    out->blas_prims = uint32_t(hs.blas_prims.size() / 3);
    out->tlas_nodes = uint32_t(hs.tlas_nodes.size() / 2);
    out->max_blas_depth = uint32_t(hs.max_blas_depth);
    out->max_tlas_depth = uint32_t(hs.max_tlas_depth);
    out->device_bytes = uint64_t(hs.ents.size() * sizeof(DevEntity) + hs.mats.size() * sizeof(DevMaterial) +
                                 (hs.blas_nodes.size() + hs.blas_prims.size() + hs.blas_info.size() +
                                  hs.tlas_nodes.size()) * sizeof(vec4) +
                                 (hs.tlas_prims.size() + hs.alias_idx.size() + hs.alias_prob.size()) * 4);
    out->world_centre_x = hs.world.x;
    out->world_centre_y = hs.world.y;
    out->world_centre_z = hs.world.z;
    out->world_radius = hs.world.w;
    out->bvh4_nodes = hs.b4_nodes.size() / BVH4_NODE_VEC4;
    out->bvh4_leaves = hs.b4_leaves;
    out->bvh4_depth = uint32_t(hs.b4_depth);
    out->bvh4_sah = hs.b4_sah;
    out->device_bytes += (hs.b4_nodes.size() + hs.b4_prims.size()) * sizeof(vec4);
    out->grid_texels_f32 = hs.grids.size();
    out->grid_texels_f16 = hs.grids16.size();
    out->device_bytes += hs.grids.size() * 4 + hs.grids16.size() * 2;
    return AGX_OK;
}

int agx_debug_bvh4(const agx_scene_desc *desc, uint64_t *n_nodes, uint64_t *n_prims, float *nodes_out,
                   float *prims_out) {
    if (!desc || !n_nodes || !n_prims) return fail(AGX_E_INVALID_ARG, "null argument");
    if (desc->accel != AGX_ACCEL_BVH4 && desc->accel != AGX_ACCEL_BVH4_WATERTIGHT)
        return fail(AGX_E_INVALID_ARG, "descriptor does not ask for a BVH4");
    HostScene hs;
    std::string err;
    const int rc = build_host_scene(*desc, &hs, &err);
    if (rc != AGX_OK) return fail(rc, err);
    *n_nodes = hs.b4_nodes.size() / BVH4_NODE_VEC4;
    *n_prims = hs.b4_prims.size() / 3;
    if (nodes_out) std::memcpy(nodes_out, hs.b4_nodes.data(), hs.b4_nodes.size() * sizeof(vec4));
    if (prims_out) std::memcpy(prims_out, hs.b4_prims.data(), hs.b4_prims.size() * sizeof(vec4));
    return AGX_OK;
}

int agx_debug_guide_infer(agx_ctx *c, const float *params, const float *inputs, uint32_t n, float *out) {
    static_assert(AGX_GUIDE_PARAMS == G_PARAMS, "guide parameter count");
    if (!c || !params || (n && (!inputs || !out))) return fail(AGX_E_INVALID_ARG, "null argument");
    if (!n) return AGX_OK;
    HIP_TRY(hipSetDevice(c->device));
    if (hipError_t e = guide_prepare(c->device); e != hipSuccess)
        return fail(AGX_E_UNSUPPORTED, std::string("guiding kernels: ") + hipGetErrorString(e) + " (LDS per block " +
                                           std::to_string(guide_smem_bytes()) + " B needed)");
    DevBuf dp, dpack, dbias, din, dout;
    HIP_TRY(hipMalloc(&dp.p, G_PARAMS * sizeof(float)));
    HIP_TRY(hipMalloc(&dpack.p, G_PACK_ALL * 2));
    HIP_TRY(hipMalloc(&dbias.p, G_BIASES * sizeof(float)));
    HIP_TRY(hipMalloc(&din.p, size_t(n) * G_IN * sizeof(float)));
    HIP_TRY(hipMalloc(&dout.p, size_t(n) * G_ROW_OUT * sizeof(float)));
    HIP_TRY(hipMemcpyAsync(dp.p, params, G_PARAMS * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(din.p, inputs, size_t(n) * G_IN * sizeof(float), hipMemcpyHostToDevice, c->stream));
    launch_guide_pack(dp.get<float>(), dpack.get<__bf16>(), dbias.get<float>(), c->stream);
    launch_guide_infer(dpack.get<__bf16>(), dbias.get<float>(), din.get<float>(), dout.get<float>(), nullptr, n,
                       c->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, dout.p, size_t(n) * G_ROW_OUT * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return AGX_OK;
}

int agx_debug_guide_train(agx_ctx *c, float *params, float *m, float *v, const float *samples, uint32_t n,
                          float wscale, float lr, int t, float *grad_out) {
    if (!c || !params || !m || !v || (n && !samples)) return fail(AGX_E_INVALID_ARG, "null argument");
    if (t < 1) return fail(AGX_E_INVALID_ARG, "Adam step t must be >= 1");
    HIP_TRY(hipSetDevice(c->device));
    if (hipError_t e = guide_prepare(c->device); e != hipSuccess)
        return fail(AGX_E_UNSUPPORTED, std::string("guiding kernels: ") + hipGetErrorString(e));
    constexpr int BLOCKS = 64;
    DevBuf dp, dm, dv, ds, dpart, dg, dscal;
    HIP_TRY(hipMalloc(&dp.p, G_PARAMS * sizeof(float)));
    HIP_TRY(hipMalloc(&dm.p, G_PARAMS * sizeof(float)));
    HIP_TRY(hipMalloc(&dv.p, G_PARAMS * sizeof(float)));
    HIP_TRY(hipMalloc(&ds.p, (size_t(n) + 1) * (G_IN + 4) * sizeof(float)));
    HIP_TRY(hipMalloc(&dpart.p, size_t(BLOCKS) * G_PARAMS * sizeof(float)));
    HIP_TRY(hipMalloc(&dg.p, G_PARAMS * sizeof(float)));
    HIP_TRY(hipMalloc(&dscal.p, 2 * sizeof(float)));
    {
        // device scalars of the training kernel: sample count and weight sum (= 1 / wscale)
        const uint32_t nn = n;
        const float ws = wscale > 0.0f ? 1.0f / wscale : 0.0f;
        HIP_TRY(hipMemcpyAsync(dscal.p, &nn, 4, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(static_cast<char *>(dscal.p) + 4, &ws, 4, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    HIP_TRY(hipMemcpyAsync(dp.p, params, G_PARAMS * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(dm.p, m, G_PARAMS * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(dv.p, v, G_PARAMS * sizeof(float), hipMemcpyHostToDevice, c->stream));
    if (n) HIP_TRY(hipMemcpyAsync(ds.p, samples, size_t(n) * (G_IN + 4) * sizeof(float), hipMemcpyHostToDevice, c->stream));
    launch_guide_train(dp.get<float>(), ds.get<float>(), dscal.get<uint32_t>(), n, dscal.get<float>() + 1,
                       dpart.get<float>(), BLOCKS, c->stream);
    // the renders' one-launch reduction + Adam runs on a copy of the partials
    // and of the state, and must give the same parameters bit for bit
    DevBuf dpart2, dp2, dm2, dv2;
    if (lr != 0.0f) {
        HIP_TRY(hipMalloc(&dpart2.p, size_t(BLOCKS) * G_PARAMS * sizeof(float)));
        HIP_TRY(hipMalloc(&dp2.p, G_PARAMS * sizeof(float)));
        HIP_TRY(hipMalloc(&dm2.p, G_PARAMS * sizeof(float)));
        HIP_TRY(hipMalloc(&dv2.p, G_PARAMS * sizeof(float)));
        HIP_TRY(hipMemcpyAsync(dpart2.p, dpart.p, size_t(BLOCKS) * G_PARAMS * sizeof(float), hipMemcpyDeviceToDevice,
                               c->stream));
        HIP_TRY(hipMemcpyAsync(dp2.p, dp.p, G_PARAMS * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(dm2.p, dm.p, G_PARAMS * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(dv2.p, dv.p, G_PARAMS * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    }
    launch_guide_reduce(dpart.get<float>(), BLOCKS, dg.get<float>(), c->stream);
    DevBuf dpk, dbi, dpk2, dbi2;
    if (lr != 0.0f) {
        // the renders' Adam also repacks the BF16 weights: check it against the packing kernel
        HIP_TRY(hipMalloc(&dpk.p, G_PACK_ALL * sizeof(__bf16)));
        HIP_TRY(hipMalloc(&dbi.p, G_BIASES * sizeof(float)));
        HIP_TRY(hipMalloc(&dpk2.p, G_PACK_ALL * sizeof(__bf16)));
        HIP_TRY(hipMalloc(&dbi2.p, G_BIASES * sizeof(float)));
        launch_guide_adam(dp.get<float>(), dm.get<float>(), dv.get<float>(), dg.get<float>(), lr, 0.9f, 0.99f, 1e-8f,
                          t, c->stream, dpk.get<__bf16>(), dbi.get<float>());
        launch_guide_pack(dp.get<float>(), dpk2.get<__bf16>(), dbi2.get<float>(), c->stream);
        std::vector<uint16_t> a(G_PACK_ALL), b(G_PACK_ALL);
        std::vector<float> ba(G_BIASES), bb(G_BIASES);
        HIP_TRY(hipMemcpyAsync(a.data(), dpk.p, G_PACK_ALL * 2, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipMemcpyAsync(b.data(), dpk2.p, G_PACK_ALL * 2, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipMemcpyAsync(ba.data(), dbi.p, G_BIASES * 4, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipMemcpyAsync(bb.data(), dbi2.p, G_BIASES * 4, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        if (a != b || std::memcmp(ba.data(), bb.data(), G_BIASES * 4) != 0)
            return fail(AGX_E_HIP, "Adam's repacked weights differ from the packing kernel's");
        launch_guide_reduce_adam(dpart2.get<float>(), BLOCKS, dp2.get<float>(), dm2.get<float>(), dv2.get<float>(), lr,
                                 0.9f, 0.99f, 1e-8f, t, c->stream, nullptr, nullptr);
        std::vector<float> pa(G_PARAMS), pb(G_PARAMS), va(G_PARAMS), vb(G_PARAMS);
        HIP_TRY(hipMemcpyAsync(pa.data(), dp.p, G_PARAMS * 4, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipMemcpyAsync(pb.data(), dp2.p, G_PARAMS * 4, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipMemcpyAsync(va.data(), dv.p, G_PARAMS * 4, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipMemcpyAsync(vb.data(), dv2.p, G_PARAMS * 4, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        if (std::memcmp(pa.data(), pb.data(), G_PARAMS * 4) != 0 || std::memcmp(va.data(), vb.data(), G_PARAMS * 4) != 0)
            return fail(AGX_E_HIP, "the one-launch reduction + Adam differs from the two-pass reduction + Adam");
    }
    HIP_TRY(hipGetLastError());
    if (grad_out) HIP_TRY(hipMemcpyAsync(grad_out, dg.p, G_PARAMS * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(params, dp.p, G_PARAMS * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(m, dm.p, G_PARAMS * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(v, dv.p, G_PARAMS * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return AGX_OK;
}

int agx_debug_guide_grad_mfma(agx_ctx *c, const float *params, const float *samples, uint32_t n, float wscale,
                              float *grad_out) {
    if (!c || !params || !grad_out || (n && !samples)) return fail(AGX_E_INVALID_ARG, "null argument");
    HIP_TRY(hipSetDevice(c->device));
    if (hipError_t e = guide_prepare(c->device); e != hipSuccess)
        return fail(AGX_E_UNSUPPORTED, std::string("guiding kernels: ") + hipGetErrorString(e));
    constexpr int BLOCKS = 64;
    const uint32_t cap = std::max<uint32_t>(n, 1);
    DevBuf dp, dpack, dbias, ds, dpart, dg, dscal;
    HIP_TRY(hipMalloc(&dp.p, G_PARAMS * sizeof(float)));
    HIP_TRY(hipMalloc(&dpack.p, G_PACK_ALL * sizeof(__bf16)));
    HIP_TRY(hipMalloc(&dbias.p, G_BIASES * sizeof(float)));
    HIP_TRY(hipMalloc(&ds.p, size_t(cap) * (G_IN + 4) * sizeof(float)));
    HIP_TRY(hipMalloc(&dpart.p, size_t(BLOCKS) * G_PARAMS * sizeof(float)));
    HIP_TRY(hipMalloc(&dg.p, G_PARAMS * sizeof(float)));
    HIP_TRY(hipMalloc(&dscal.p, 2 * sizeof(float)));
    const uint32_t nn = n;
    const float ws = wscale > 0.0f ? 1.0f / wscale : 0.0f;
    HIP_TRY(hipMemcpyAsync(dscal.p, &nn, 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(static_cast<char *>(dscal.p) + 4, &ws, 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(dp.p, params, G_PARAMS * sizeof(float), hipMemcpyHostToDevice, c->stream));
    if (n) HIP_TRY(hipMemcpyAsync(ds.p, samples, size_t(n) * (G_IN + 4) * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    launch_guide_pack(dp.get<float>(), dpack.get<__bf16>(), dbias.get<float>(), c->stream);
    launch_guide_train_mfma(dpack.get<__bf16>(), dbias.get<float>(), ds.get<float>(), dscal.get<uint32_t>(), cap,
                            dscal.get<float>() + 1, dpart.get<float>(), BLOCKS, c->stream);
    launch_guide_reduce(dpart.get<float>(), BLOCKS, dg.get<float>(), c->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(grad_out, dg.p, G_PARAMS * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return AGX_OK;
}

}  // extern "C"
