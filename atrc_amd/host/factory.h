// Atrc's object factory (src/factory/include/agz/factory/context.h:27-300,
// src/factory/src/creator/*) for the object families the `gpu_pt` path needs.
// Scene objects here are plain descriptors: the factory turns Atrc JSON into
// the agx_scene_desc that crosses the C ABI of include/atrc_gpu.h.
#pragma once

#include <functional>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/atrc_gpu.h"
#include "config.h"

namespace atrc {

class CreatingObjectException : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

// ---------------------------------------------------------------- scene objects
struct Texture2D {
    std::string type = "constant";  // "constant", "hdr" or "checker_board"
    Vec3 texel;                      // "constant" texel
    agx_texture2d_desc desc{};       // every kind; data points into image (hdr)
    std::shared_ptr<std::vector<float>> image;
};
struct Geometry {
    agx_entity_desc proto{};         // geometry part of the entity descriptor
    std::shared_ptr<std::vector<agx_vertex>> mesh;
};
struct Material {
    agx_material_desc desc{};
};
struct Texture3D {
    agx_texture3d_desc desc{};       // data points into grid (image3d) or is NULL (constant)
    std::shared_ptr<std::vector<float>> grid;
};
struct Medium {
    std::string type;                // "void", "homogeneous" or "heterogeneous"
    agx_medium_desc desc{};
    std::shared_ptr<Texture3D> density, albedo, g;
};
struct Entity {
    agx_entity_desc desc{};
    std::shared_ptr<std::vector<agx_vertex>> mesh;
    std::shared_ptr<Material> material;
    std::shared_ptr<Medium> med_in, med_out;  // NULL or "void" = no medium
};
struct EnvirLight {
    agx_env_desc desc{};
    std::shared_ptr<Texture2D> tex;  // ibl: owns the texels desc.tex points to
};
struct Aggregate {
    int32_t kind = AGX_AGGREGATE_NATIVE;
    int32_t max_leaf_size = 5;
};
struct Camera {
    agx_camera_desc desc{};          // film_aspect filled when the camera is created
};
struct FilmFilter {
    std::string type = "box";
    float radius = 0.5f;
    float alpha = 0;
};

// DefaultScene (src/tracer/src/core/scene/scene_default.cpp): owns every
// descriptor and buffer that agx_scene_desc points to.
class Scene {
public:
    std::vector<std::shared_ptr<Entity>> entities;
    std::shared_ptr<EnvirLight> env;
    Aggregate aggregate;
    std::shared_ptr<Camera> camera;

    // Descriptor view, valid while this Scene lives and is not modified.
    const agx_scene_desc &desc();

private:
    std::vector<agx_entity_desc> ents_;
    std::vector<agx_material_desc> mats_;
    std::vector<agx_medium_desc> meds_;
    agx_scene_desc desc_{};
};

// ---------------------------------------------------------------- path mapping
class PathMapper {
public:
    void add_replacer(const std::string &key, const std::string &value) { rep_[key] = value; }
    std::string map(const std::string &s) const;
private:
    std::map<std::string, std::string> rep_;
};

class Renderer;
class RendererInteractor;
class PostProcessor;

// ---------------------------------------------------------------- factories
class CreatingContext;

template <class T> using CreatorFn = std::function<std::shared_ptr<T>(const ConfigGroup &, CreatingContext &)>;

template <class T> class Factory {
public:
    explicit Factory(std::string name) : name_(std::move(name)) {}
    void add_creator(const std::string &type, CreatorFn<T> fn) { creators_[type] = std::move(fn); }
    const CreatorFn<T> *get(const std::string &type) const {
        auto it = creators_.find(type);
        return it == creators_.end() ? nullptr : &it->second;
    }
    const std::string &name() const { return name_; }
    std::map<std::vector<std::string>, std::shared_ptr<T>> references;  // ReferenceCreator cache

private:
    std::string name_;
    std::map<std::string, CreatorFn<T>> creators_;
};

class CreatingContext {
public:
    CreatingContext();
    const PathMapper *path_mapper = nullptr;
    const ConfigGroup *reference_root = nullptr;

    // type-dispatched creation with Atrc's "reference" creator
    // (context.h:179-208) and hierarchical error messages (context.h:280-300)
    template <class T> std::shared_ptr<T> create(const ConfigGroup &params);

    float film_aspect = 1;  // argument Atrc passes to camera creators

    Factory<Texture2D> texture2d{"texture2d"};
    Factory<Texture3D> texture3d{"texture3d"};
    Factory<Geometry> geometry{"geometry"};
    Factory<Material> material{"material"};
    Factory<Medium> medium{"medium"};
    Factory<Entity> entity{"entity"};
    Factory<EnvirLight> envir_light{"envir_light"};
    Factory<Aggregate> aggregate{"aggregate"};
    Factory<Camera> camera{"camera"};
    Factory<FilmFilter> film_filter{"film_filter"};
    Factory<Renderer> renderer{"renderer"};
    Factory<RendererInteractor> reporter{"renderer_interactor"};
    Factory<PostProcessor> post_processor{"post_processor"};
    Factory<Scene> scene{"scene"};

    template <class T> Factory<T> &factory();
};

// Mesh files: ".bm" (size_t n, then n triangles of 3 x {pos[3], nor[3], uv[2]}
// floats; src/factory/src/utility/bin_mesh.cpp:7-27) or Wavefront OBJ.
std::vector<agx_vertex> load_triangle_mesh(const std::string &filename);
void save_bin_mesh(const std::string &filename, const std::vector<agx_vertex> &tris);

// Registers the renderer creators ("gpu_pt"); defined with the renderer.
void register_renderers(CreatingContext &ctx);

}  // namespace atrc
