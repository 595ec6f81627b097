// Atrc's renderer-facing API (src/tracer/include/agz/tracer/core/renderer.h:16-64,
// core/render_target.h:64-74, core/renderer_interactor.h:17-65) and the GPU
// path tracer registered as "gpu_pt" beside Atrc's CPU "pt"
// (src/factory/src/creator/renderer_creators.cpp:115-151).
#pragma once

#include <atomic>
#include <future>
#include <memory>
#include <string>
#include <vector>

#include "factory.h"

namespace atrc {

template <class T> struct Image2D {
    int w = 0, h = 0;
    std::vector<T> data;  // row-major, row 0 = film y in [0, 1/h)
    Image2D() = default;
    Image2D(int width, int height) : w(width), h(height), data(size_t(width) * size_t(height)) {}
    T &at(int y, int x) { return data[size_t(y) * size_t(w) + size_t(x)]; }
    const T &at(int y, int x) const { return data[size_t(y) * size_t(w) + size_t(x)]; }
};

struct RenderTarget {
    Image2D<Vec3> image, albedo, normal;
    Image2D<float> denoise;
};

// FilmFilterApplier: film size + filter (render_target.h)
struct FilmFilterApplier {
    int width = 0, height = 0;
    std::shared_ptr<FilmFilter> filter;
};

class RendererInteractor {
public:
    virtual ~RendererInteractor() = default;
    virtual void begin() {}
    virtual void end() {}
    virtual void new_stage() {}
    virtual void end_stage() {}
    virtual void progress(double percent) {}
};

class Renderer {
public:
    virtual ~Renderer() = default;
    virtual RenderTarget render(FilmFilterApplier filter, Scene &scene, RendererInteractor &reporter) = 0;
    // Renderer::render_async / stop_async (core/core/renderer.cpp:5-35)
    std::future<RenderTarget> render_async(FilmFilterApplier filter, Scene &scene, RendererInteractor &reporter);
    void stop_async();

protected:
    std::atomic<int32_t> stop_rendering_{0};
};

// PTRendererParams (src/tracer/include/agz/tracer/create/renderer.h:10-24)
struct PTRendererParams {
    int worker_count = 0;
    int task_grid_size = 32;
    int spp = 1;
    int min_depth = 5;
    int max_depth = 10;
    float cont_prob = 0.9f;
    bool use_mis = true;
    int specular_depth = 20;
    int device = 0;          // HIP device ordinal (gpu_pt only)
    uint64_t seed = 42;      // PCG32 stream seed (gpu_pt only)
    // gpu_pt only (no counterpart in Atrc's "pt"):
    int accel = 0;           // AGX_ACCEL_*: "atrc" (default), "bvh4", "bvh4_watertight"
    bool guide = false;      // path guiding (agx_render_params.guide*)
    float guide_alpha = 0.5f;
    int guide_batch = 65536;
    int guide_steps = 4;
    float guide_lr = 1e-2f;
    uint32_t guide_seed = 0;
    int guide_medium_depth = 0;  // medium scattering points guided up to this depth (0: none)
};

std::shared_ptr<Renderer> create_gpu_pt_renderer(const PTRendererParams &params);

class PostProcessor {
public:
    virtual ~PostProcessor() = default;
    virtual void process(RenderTarget &rt) = 0;
};

// RenderSession (src/factory/src/utility/render_session.cpp:14-132)
struct RenderSetting {
    int width = 0, height = 0;
    std::shared_ptr<FilmFilter> film_filter;
    std::shared_ptr<Camera> camera;
    std::shared_ptr<Renderer> renderer;
    std::shared_ptr<RendererInteractor> reporter;
    std::vector<std::shared_ptr<PostProcessor>> post_processors;
    float eps = 3e-4f;
};

struct RenderSession {
    std::shared_ptr<Scene> scene;
    std::shared_ptr<RenderSetting> settings;
    RenderTarget execute();
};

// the current global EPS (common.cpp:7-16); set by RenderSession::execute
float EPS();
void set_eps(float eps);

RenderSession create_render_session(std::shared_ptr<Scene> scene, const ConfigGroup &rendering,
                                    CreatingContext &ctx);
std::vector<RenderSession> parse_render_sessions(const ConfigGroup &root, CreatingContext &ctx);

// image writers used by save_to_img
void write_png(const std::string &file, int w, int h, const std::vector<uint8_t> &rgb);
void write_hdr(const std::string &file, int w, int h, const std::vector<float> &rgb);

}  // namespace atrc
