// gpu_pt renderer, reporters, post-processors and render sessions.
#include "renderer.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <mutex>

namespace atrc {

namespace {
float g_eps = 3e-4f;
}
float EPS() { return g_eps; }
void set_eps(float e) { g_eps = e; }

std::future<RenderTarget> Renderer::render_async(FilmFilterApplier filter, Scene &scene,
                                                 RendererInteractor &reporter) {
    stop_rendering_ = 0;
    return std::async(std::launch::async, [this, filter, &scene, &reporter] { return render(filter, scene, reporter); });
}

void Renderer::stop_async() { stop_rendering_ = 1; }

// ---------------------------------------------------------------- gpu_pt
namespace {

std::string agx_error(int code) {
    return "gpu_pt: " + std::string(agx_last_error()) + " (status " + std::to_string(code) + ")";
}

class GpuPathTracingRenderer : public Renderer {
    PTRendererParams params_;
    agx_ctx *ctx_ = nullptr;
    std::mutex mu_;

public:
    explicit GpuPathTracingRenderer(const PTRendererParams &p) : params_(p) {
        if (p.spp <= 0) throw ObjectConstructionException("invalid spp value");
    }
    ~GpuPathTracingRenderer() override {
        if (ctx_) agx_destroy(ctx_);
    }

    RenderTarget render(FilmFilterApplier filter, Scene &scene, RendererInteractor &reporter) override {
        std::lock_guard<std::mutex> lk(mu_);
        if (!filter.filter || (filter.filter->type != "box" && filter.filter->type != "gaussian"))
            throw ObjectConstructionException("gpu_pt supports the box and gaussian film filters");
        if (!scene.camera) throw ObjectConstructionException("scene has no camera");
        int rc;
        if (!ctx_ && (rc = agx_create(params_.device, &ctx_)) != AGX_OK) throw std::runtime_error(agx_error(rc));
        agx_scene_desc sd = scene.desc();
        sd.accel = params_.accel;
        if ((rc = agx_upload_scene(ctx_, &sd)) != AGX_OK) throw std::runtime_error(agx_error(rc));
        agx_render_params rp{};
        rp.width = filter.width;
        rp.height = filter.height;
        rp.spp = params_.spp;
        rp.min_depth = params_.min_depth;
        rp.max_depth = params_.max_depth;
        rp.specular_depth = params_.specular_depth;
        rp.cont_prob = params_.cont_prob;
        rp.use_mis = params_.use_mis ? 1 : 0;
        rp.eps = EPS();
        rp.filter_radius = filter.filter->radius;
        rp.filter = filter.filter->type == "gaussian" ? AGX_FILTER_GAUSSIAN : AGX_FILTER_BOX;
        rp.filter_alpha = filter.filter->alpha;
        rp.task_grid_size = params_.task_grid_size;
        rp.seed = params_.seed;
        rp.camera = scene.camera->desc;
        if (params_.guide) {
            rp.guide = 1;
            rp.guide_alpha = params_.guide_alpha;
            rp.guide_train = 1;
            rp.guide_batch = params_.guide_batch;
            rp.guide_steps = params_.guide_steps;
            rp.guide_lr = params_.guide_lr;
            rp.guide_reset = 1;
            rp.guide_seed = params_.guide_seed;
            rp.guide_medium_depth = params_.guide_medium_depth;
        }
        const size_t wh = size_t(filter.width) * size_t(filter.height);
        std::vector<float> rgb(wh * 3), alb(wh * 3), nor(wh * 3), den(wh);
        reporter.begin();
        reporter.new_stage();
        rc = agx_render(ctx_, &rp, reinterpret_cast<const volatile int32_t *>(&stop_rendering_), rgb.data(),
                        alb.data(), nor.data(), den.data());
        if (rc != AGX_OK && rc != AGX_E_CANCELLED) throw std::runtime_error(agx_error(rc));
        reporter.progress(100);
        reporter.end_stage();
        reporter.end();
        RenderTarget rt;
        rt.image = Image2D<Vec3>(filter.width, filter.height);
        rt.albedo = Image2D<Vec3>(filter.width, filter.height);
        rt.normal = Image2D<Vec3>(filter.width, filter.height);
        rt.denoise = Image2D<float>(filter.width, filter.height);
        for (size_t i = 0; i < wh; ++i) {
            rt.image.data[i] = {rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]};
            rt.albedo.data[i] = {alb[3 * i], alb[3 * i + 1], alb[3 * i + 2]};
            rt.normal.data[i] = {nor[3 * i], nor[3 * i + 1], nor[3 * i + 2]};
            rt.denoise.data[i] = den[i];
        }
        return rt;
    }
};

class StdOutput : public RendererInteractor {
public:
    void begin() override { std::cout << "[gpu_pt] rendering" << std::endl; }
    void progress(double p) override { std::cout << "[gpu_pt] progress " << p << "%" << std::endl; }
    void end() override { std::cout << "[gpu_pt] done" << std::endl; }
};
class NoOutput : public RendererInteractor {};

class GammaCorrector : public PostProcessor {
    float gamma_;
public:
    explicit GammaCorrector(float g) : gamma_(g) {
        if (!(g > 0)) throw ObjectConstructionException("invalid gamma value");
    }
    void process(RenderTarget &rt) override {
        for (auto &p : rt.image.data) p = {std::pow(p.x, gamma_), std::pow(p.y, gamma_), std::pow(p.z, gamma_)};
    }
};

// save_to_img (src/tracer/src/core/post_processor/save_to_img.cpp:41-67):
// flip vertically, c^gamma clamped to [0, 1], 8-bit PNG or float HDR
class SaveToImage : public PostProcessor {
    std::string file_, ext_;
    float gamma_;
public:
    SaveToImage(std::string f, std::string ext, float g) : file_(std::move(f)), ext_(std::move(ext)), gamma_(g) {
        if (ext_ != "png" && ext_ != "hdr") throw ObjectConstructionException("unsupported image extension: " + ext_);
        if (!(gamma_ > 0)) throw ObjectConstructionException("invalid gamma value");
    }
    void process(RenderTarget &rt) override {
        const int w = rt.image.w, h = rt.image.h;
        const auto parent = std::filesystem::path(file_).parent_path();
        if (!parent.empty()) std::filesystem::create_directories(parent);
        if (ext_ == "hdr") {
            std::vector<float> rgb(size_t(w) * h * 3);
            for (int y = 0; y < h; ++y)
                for (int x = 0; x < w; ++x) {
                    const Vec3 &c = rt.image.at(h - 1 - y, x);
                    float *o = &rgb[(size_t(y) * w + x) * 3];
                    o[0] = c.x; o[1] = c.y; o[2] = c.z;
                }
            write_hdr(file_, w, h, rgb);
            return;
        }
        std::vector<uint8_t> rgb(size_t(w) * h * 3);
        auto q = [this](float c) {
            const float v = std::pow(c, gamma_);
            return uint8_t((v < 0 ? 0.0f : (v > 1 ? 1.0f : v)) * 255);
        };
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) {
                const Vec3 &c = rt.image.at(h - 1 - y, x);
                uint8_t *o = &rgb[(size_t(y) * w + x) * 3];
                o[0] = q(c.x); o[1] = q(c.y); o[2] = q(c.z);
            }
        write_png(file_, w, h, rgb);
    }
};

uint32_t crc32(const uint8_t *p, size_t n, uint32_t c = 0xffffffffu) {
    static uint32_t table[256];
    static bool init = false;
    if (!init) {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t v = i;
            for (int k = 0; k < 8; ++k) v = (v & 1) ? 0xedb88320u ^ (v >> 1) : v >> 1;
            table[i] = v;
        }
        init = true;
    }
    for (size_t i = 0; i < n; ++i) c = table[(c ^ p[i]) & 0xff] ^ (c >> 8);
    return c;
}

}  // namespace

// PNG with stored (uncompressed) deflate blocks
void write_png(const std::string &file, int w, int h, const std::vector<uint8_t> &rgb) {
    std::vector<uint8_t> raw;
    raw.reserve(size_t(h) * (size_t(w) * 3 + 1));
    for (int y = 0; y < h; ++y) {
        raw.push_back(0);
        raw.insert(raw.end(), rgb.begin() + ptrdiff_t(size_t(y) * w * 3), rgb.begin() + ptrdiff_t(size_t(y + 1) * w * 3));
    }
    std::vector<uint8_t> z = {0x78, 0x01};
    for (size_t off = 0; off < raw.size() || off == 0;) {
        const size_t len = std::min<size_t>(65535, raw.size() - off);
        const bool last = off + len >= raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back(uint8_t(len & 0xff)); z.push_back(uint8_t(len >> 8));
        z.push_back(uint8_t(~len & 0xff)); z.push_back(uint8_t((~len >> 8) & 0xff));
        z.insert(z.end(), raw.begin() + ptrdiff_t(off), raw.begin() + ptrdiff_t(off + len));
        off += len;
        if (last) break;
    }
    uint32_t a = 1, b = 0;
    for (uint8_t v : raw) { a = (a + v) % 65521; b = (b + a) % 65521; }
    const uint32_t adler = (b << 16) | a;
    for (int s = 24; s >= 0; s -= 8) z.push_back(uint8_t(adler >> s));
    std::ofstream o(file, std::ios::binary);
    if (!o) throw std::runtime_error("cannot write " + file);
    const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    o.write(reinterpret_cast<const char *>(sig), 8);
    auto chunk = [&](const char *type, const std::vector<uint8_t> &data) {
        const uint32_t n = uint32_t(data.size());
        const uint8_t len[4] = {uint8_t(n >> 24), uint8_t(n >> 16), uint8_t(n >> 8), uint8_t(n)};
        o.write(reinterpret_cast<const char *>(len), 4);
        std::vector<uint8_t> td(type, type + 4);
        td.insert(td.end(), data.begin(), data.end());
        o.write(reinterpret_cast<const char *>(td.data()), std::streamsize(td.size()));
        const uint32_t c = crc32(td.data(), td.size()) ^ 0xffffffffu;
        const uint8_t cb[4] = {uint8_t(c >> 24), uint8_t(c >> 16), uint8_t(c >> 8), uint8_t(c)};
        o.write(reinterpret_cast<const char *>(cb), 4);
    };
    std::vector<uint8_t> ihdr = {uint8_t(w >> 24), uint8_t(w >> 16), uint8_t(w >> 8), uint8_t(w),
                                 uint8_t(h >> 24), uint8_t(h >> 16), uint8_t(h >> 8), uint8_t(h), 8, 2, 0, 0, 0};
    chunk("IHDR", ihdr);
    chunk("IDAT", z);
    chunk("IEND", {});
}

// Radiance RGBE (.hdr), flat scanlines
void write_hdr(const std::string &file, int w, int h, const std::vector<float> &rgb) {
    std::ofstream o(file, std::ios::binary);
    if (!o) throw std::runtime_error("cannot write " + file);
    o << "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y " << h << " +X " << w << "\n";
    for (size_t i = 0; i < size_t(w) * h; ++i) {
        const float r = rgb[3 * i], g = rgb[3 * i + 1], b = rgb[3 * i + 2];
        const float m = std::max(r, std::max(g, b));
        uint8_t e[4] = {0, 0, 0, 0};
        if (m > 1e-32f) {
            int ex;
            const float f = std::frexp(m, &ex) * 256.0f / m;
            e[0] = uint8_t(std::max(0.0f, r) * f);
            e[1] = uint8_t(std::max(0.0f, g) * f);
            e[2] = uint8_t(std::max(0.0f, b) * f);
            e[3] = uint8_t(ex + 128);
        }
        o.write(reinterpret_cast<const char *>(e), 4);
    }
}

std::shared_ptr<Renderer> create_gpu_pt_renderer(const PTRendererParams &p) {
    return std::make_shared<GpuPathTracingRenderer>(p);
}

void register_renderers(CreatingContext &ctx) {
    // same parameter names and defaults as Atrc's "pt" (renderer_creators.cpp:124-137)
    ctx.renderer.add_creator("gpu_pt", [](const ConfigGroup &p, CreatingContext &) {
        PTRendererParams r;
        r.worker_count = p.child_int_or("worker_count", 0);
        r.task_grid_size = p.child_int_or("task_grid_size", 32);
        r.spp = p.child_int("spp");
        r.min_depth = p.child_int_or("min_depth", 5);
        r.max_depth = p.child_int_or("max_depth", 10);
        r.cont_prob = p.child_real_or("cont_prob", 0.9f);
        r.use_mis = p.child_int_or("use_mis", 1) != 0;
        r.specular_depth = p.child_int_or("specular_depth", 20);
        r.device = p.child_int_or("device", 0);
        r.seed = uint64_t(p.child_int_or("seed", 42));
        const std::string accel = p.child_str_or("accel", "atrc");
        if (accel == "atrc") r.accel = AGX_ACCEL_ATRC;
        else if (accel == "bvh4") r.accel = AGX_ACCEL_BVH4;
        else if (accel == "bvh4_watertight") r.accel = AGX_ACCEL_BVH4_WATERTIGHT;
        else throw ObjectConstructionException("unknown accel: " + accel);
        r.guide = p.child_int_or("guide", 0) != 0;
        r.guide_alpha = p.child_real_or("guide_alpha", 0.5f);
        r.guide_batch = p.child_int_or("guide_batch", 65536);
        r.guide_steps = p.child_int_or("guide_steps", 4);
        r.guide_lr = p.child_real_or("guide_lr", 1e-2f);
        r.guide_seed = uint32_t(p.child_int_or("guide_seed", 0));
        r.guide_medium_depth = p.child_int_or("guide_medium_depth", 0);
        return create_gpu_pt_renderer(r);
    });
    ctx.reporter.add_creator("stdout", [](const ConfigGroup &, CreatingContext &) {
        return std::shared_ptr<RendererInteractor>(std::make_shared<StdOutput>());
    });
    ctx.reporter.add_creator("noout", [](const ConfigGroup &, CreatingContext &) {
        return std::shared_ptr<RendererInteractor>(std::make_shared<NoOutput>());
    });
    ctx.post_processor.add_creator("gamma", [](const ConfigGroup &p, CreatingContext &) {
        float g;
        if (auto n = p.find_child("gamma")) g = n->as_value().as_real();
        else g = 1 / p.child_real("inv_gamma");
        return std::shared_ptr<PostProcessor>(std::make_shared<GammaCorrector>(g));
    });
    ctx.post_processor.add_creator("save_to_img", [](const ConfigGroup &p, CreatingContext &c) {
        const std::string file = c.path_mapper ? c.path_mapper->map(p.child_str("filename")) : p.child_str("filename");
        float g = 1;
        if (auto n = p.find_child_value("gamma")) g = n->as_real();
        else if (auto m = p.find_child_value("inv_gamma")) g = 1 / m->as_real();
        std::string ext;
        if (auto n = p.find_child_value("ext")) ext = n->as_str();
        else if (ends_with(file, ".png")) ext = "png";
        else if (ends_with(file, ".jpg")) ext = "jpg";
        else if (ends_with(file, ".hdr")) ext = "hdr";
        else throw ObjectConstructionException("unknown image file format: " + file);
        return std::shared_ptr<PostProcessor>(std::make_shared<SaveToImage>(file, ext, g));
    });
}

RenderTarget RenderSession::execute() {  // render_session.cpp:76-96
    set_eps(settings->eps);
    scene->camera = settings->camera;
    FilmFilterApplier f;
    f.width = settings->width;
    f.height = settings->height;
    f.filter = settings->film_filter;
    RenderTarget rt = settings->renderer->render(f, *scene, *settings->reporter);
    for (auto &pp : settings->post_processors) pp->process(rt);
    return rt;
}

RenderSession create_render_session(std::shared_ptr<Scene> scene, const ConfigGroup &rc, CreatingContext &ctx) {
    auto s = std::make_shared<RenderSetting>();  // parse_rendering_settings (render_session.cpp:14-66)
    s->width = rc.child_int("width");
    s->height = rc.child_int("height");
    if (s->width <= 0 || s->height <= 0) throw ObjectConstructionException("invalid film size");
    if (auto g = rc.find_child_group("film_filter")) s->film_filter = ctx.create<FilmFilter>(*g);
    else {
        s->film_filter = std::make_shared<FilmFilter>();
        s->film_filter->radius = 0.5f;
    }
    ctx.film_aspect = float(s->width) / float(s->height);
    s->camera = ctx.create<Camera>(rc.child_group("camera"));
    s->renderer = ctx.create<Renderer>(rc.child_group("renderer"));
    s->reporter = ctx.create<RendererInteractor>(rc.child_group("reporter"));
    if (auto n = rc.find_child("post_processors")) {
        const auto &arr = n->as_array();
        for (size_t i = 0; i < arr.size(); ++i) {
            const auto &g = arr.at(i).as_group();
            if (ends_with(g.child_str("type"), "//")) continue;
            s->post_processors.push_back(ctx.create<PostProcessor>(g));
        }
    }
    if (auto v = rc.find_child_value("eps")) s->eps = v->as_real();
    return RenderSession{std::move(scene), std::move(s)};
}

std::vector<RenderSession> parse_render_sessions(const ConfigGroup &root, CreatingContext &ctx) {
    const ConfigGroup &scene_cfg = root.child_group("scene");
    const ConfigNode &rendering = root.child("rendering");
    ctx.reference_root = &scene_cfg;
    auto scene = ctx.create<Scene>(scene_cfg);
    std::vector<RenderSession> out;
    if (rendering.is_array()) {
        const auto &arr = rendering.as_array();
        for (size_t i = 0; i < arr.size(); ++i) out.push_back(create_render_session(scene, arr.at_group(i), ctx));
    } else {
        out.push_back(create_render_session(scene, rendering.as_group(), ctx));
    }
    return out;
}

}  // namespace atrc
