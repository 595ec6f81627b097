// extern "C" surface of the C++ host (libatrc_host.so): run Atrc JSON render
// sessions through gpu_pt, or build the scene descriptor a JSON file describes.
#include <cstring>
#include <filesystem>
#include <memory>
#include <string>

#include "../../include/atrc_host.h"
#include "factory.h"
#include "renderer.h"

namespace {

thread_local std::string g_err;

struct HostScene {
    atrc::CreatingContext ctx;
    atrc::PathMapper mapper;
    std::shared_ptr<atrc::ConfigGroup> root;
    std::shared_ptr<atrc::Scene> scene;
    std::vector<atrc::RenderSession> sessions;
};

void setup_paths(HostScene *h, const char *scene_dir) {
    h->mapper.add_replacer("${working-directory}",
                           std::filesystem::absolute(std::filesystem::current_path()).lexically_normal().string());
    h->mapper.add_replacer("${scene-directory}", scene_dir && *scene_dir ? std::string(scene_dir) : std::string("."));
    h->ctx.path_mapper = &h->mapper;
}

}  // namespace

extern "C" {

const char *atrc_host_last_error(void) { return g_err.c_str(); }

void *atrc_host_load(const char *json_text, const char *scene_dir) {
    try {
        auto h = std::make_unique<HostScene>();
        setup_paths(h.get(), scene_dir);
        h->root = atrc::json_to_config(json_text);
        h->sessions = atrc::parse_render_sessions(*h->root, h->ctx);
        h->scene = h->sessions.at(0).scene;
        h->scene->camera = h->sessions.at(0).settings->camera;
        return h.release();
    } catch (const std::exception &e) {
        g_err = e.what();
        return nullptr;
    }
}

const agx_scene_desc *atrc_host_scene_desc(void *handle) {
    auto *h = static_cast<HostScene *>(handle);
    return h ? &h->scene->desc() : nullptr;
}

int atrc_host_session_info(void *handle, int index, atrc_host_session *out) {
    auto *h = static_cast<HostScene *>(handle);
    if (!h || !out || index < 0 || size_t(index) >= h->sessions.size()) {
        g_err = "invalid session index";
        return 1;
    }
    const auto &s = *h->sessions[size_t(index)].settings;
    out->width = s.width;
    out->height = s.height;
    out->eps = s.eps;
    out->filter_radius = s.film_filter ? s.film_filter->radius : 0.5f;
    out->camera = s.camera->desc;
    out->session_count = int32_t(h->sessions.size());
    out->filter = s.film_filter && s.film_filter->type == "gaussian" ? AGX_FILTER_GAUSSIAN : AGX_FILTER_BOX;
    out->filter_alpha = s.film_filter ? s.film_filter->alpha : 0.0f;
    return 0;
}

int atrc_host_execute(void *handle, int index, float *rgb, float *albedo, float *normal, float *denoise) {
    auto *h = static_cast<HostScene *>(handle);
    if (!h || index < 0 || size_t(index) >= h->sessions.size()) {
        g_err = "invalid session index";
        return 1;
    }
    try {
        atrc::RenderTarget rt = h->sessions[size_t(index)].execute();
        const size_t wh = rt.image.data.size();
        for (size_t i = 0; i < wh; ++i) {
            if (rgb) { rgb[3 * i] = rt.image.data[i].x; rgb[3 * i + 1] = rt.image.data[i].y; rgb[3 * i + 2] = rt.image.data[i].z; }
            if (albedo) { albedo[3 * i] = rt.albedo.data[i].x; albedo[3 * i + 1] = rt.albedo.data[i].y; albedo[3 * i + 2] = rt.albedo.data[i].z; }
            if (normal) { normal[3 * i] = rt.normal.data[i].x; normal[3 * i + 1] = rt.normal.data[i].y; normal[3 * i + 2] = rt.normal.data[i].z; }
            if (denoise) denoise[i] = rt.denoise.data[i];
        }
        return 0;
    } catch (const std::exception &e) {
        g_err = e.what();
        return 1;
    }
}

void atrc_host_free(void *handle) { delete static_cast<HostScene *>(handle); }

int atrc_host_write_bin_mesh(const char *filename, const agx_vertex *tris, uint32_t tri_count) {
    try {
        atrc::save_bin_mesh(filename, std::vector<agx_vertex>(tris, tris + size_t(tri_count) * 3));
        return 0;
    } catch (const std::exception &e) {
        g_err = e.what();
        return 1;
    }
}

}  // extern "C"
