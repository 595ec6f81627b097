// Creators of the Atrc object families on the gpu_pt path.
// Parameter names and defaults follow src/factory/src/creator/*_creators.cpp.
#include "factory.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <sstream>

#include "renderer.h"

namespace atrc {

// ---------------------------------------------------------------- factory plumbing
template <> Factory<Texture2D> &CreatingContext::factory<Texture2D>() { return texture2d; }
template <> Factory<Geometry> &CreatingContext::factory<Geometry>() { return geometry; }
template <> Factory<Material> &CreatingContext::factory<Material>() { return material; }
template <> Factory<Medium> &CreatingContext::factory<Medium>() { return medium; }
template <> Factory<Entity> &CreatingContext::factory<Entity>() { return entity; }
template <> Factory<EnvirLight> &CreatingContext::factory<EnvirLight>() { return envir_light; }
template <> Factory<Aggregate> &CreatingContext::factory<Aggregate>() { return aggregate; }
template <> Factory<Camera> &CreatingContext::factory<Camera>() { return camera; }
template <> Factory<FilmFilter> &CreatingContext::factory<FilmFilter>() { return film_filter; }
template <> Factory<Renderer> &CreatingContext::factory<Renderer>() { return renderer; }
template <> Factory<RendererInteractor> &CreatingContext::factory<RendererInteractor>() { return reporter; }
template <> Factory<PostProcessor> &CreatingContext::factory<PostProcessor>() { return post_processor; }
template <> Factory<Scene> &CreatingContext::factory<Scene>() { return scene; }

template <class T> std::shared_ptr<T> CreatingContext::create(const ConfigGroup &params) {
    Factory<T> &f = factory<T>();
    try {
        const std::string &type = params.child_str("type");
        if (type == "reference") {  // ReferenceCreator (context.h:179-208)
            const ConfigArray &arr = params.child_array("name");
            if (arr.size() < 1) throw CreatingObjectException("empty reference name sequence");
            std::vector<std::string> names;
            for (size_t i = 0; i < arr.size(); ++i) names.push_back(arr.at_str(i));
            auto it = f.references.find(names);
            if (it != f.references.end()) return it->second;
            if (!reference_root) throw CreatingObjectException("no reference root");
            const ConfigGroup *g = reference_root;
            for (size_t i = 0; i + 1 < names.size(); ++i) g = &g->child_group(names[i]);
            auto obj = create<T>(g->child_group(names.back()));
            f.references[names] = obj;
            return obj;
        }
        const CreatorFn<T> *fn = f.get(type);
        if (!fn) throw CreatingObjectException("unknown creator type name: " + type);
        try {
            return (*fn)(params, *this);
        } catch (const std::exception &e) {
            throw CreatingObjectException(std::string(e.what()) + "\n  in creating object with creator: " + type);
        }
    } catch (const std::exception &e) {
        throw CreatingObjectException(std::string(e.what()) + "\n  in creating object with factory: " + f.name());
    }
}

template std::shared_ptr<Scene> CreatingContext::create<Scene>(const ConfigGroup &);
template std::shared_ptr<Camera> CreatingContext::create<Camera>(const ConfigGroup &);
template std::shared_ptr<FilmFilter> CreatingContext::create<FilmFilter>(const ConfigGroup &);
template std::shared_ptr<Renderer> CreatingContext::create<Renderer>(const ConfigGroup &);
template std::shared_ptr<RendererInteractor> CreatingContext::create<RendererInteractor>(const ConfigGroup &);
template std::shared_ptr<PostProcessor> CreatingContext::create<PostProcessor>(const ConfigGroup &);

std::string PathMapper::map(const std::string &s) const {  // BasicPathMapper::map (context.h:163-177)
    std::string r = s;
    for (bool done = false; !done;) {
        done = true;
        for (auto &kv : rep_) {
            size_t pos;
            while ((pos = r.find(kv.first)) != std::string::npos) {
                r.replace(pos, kv.first.size(), kv.second);
                done = false;
            }
        }
    }
    return std::filesystem::absolute(std::filesystem::path(r)).lexically_normal().string();
}

// ---------------------------------------------------------------- meshes
std::vector<agx_vertex> load_triangle_mesh(const std::string &filename) {
    std::vector<agx_vertex> out;
    if (ends_with(filename, ".bm")) {  // bin_mesh.cpp:7-27
        std::ifstream in(filename, std::ios::binary);
        if (!in) throw ObjectConstructionException("failed to open mesh file: " + filename);
        uint64_t n = 0;
        in.read(reinterpret_cast<char *>(&n), sizeof(n));
        if (!in || n > (uint64_t(1) << 32)) throw ObjectConstructionException("invalid binary mesh: " + filename);
        out.resize(size_t(n) * 3);
        in.read(reinterpret_cast<char *>(out.data()), std::streamsize(out.size() * sizeof(agx_vertex)));
        if (!in) throw ObjectConstructionException("truncated binary mesh: " + filename);
        return out;
    }
    if (!ends_with(filename, ".obj")) throw ObjectConstructionException("unsupported mesh file format: " + filename);
    std::ifstream in(filename);
    if (!in) throw ObjectConstructionException("failed to open mesh file: " + filename);
    std::vector<Vec3> P, N;
    std::vector<Vec2> T;
    std::string line;
    auto resolve = [](long idx, size_t n) -> long { return idx < 0 ? long(n) + idx : idx - 1; };
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string tag;
        ls >> tag;
        if (tag == "v") { Vec3 v; ls >> v.x >> v.y >> v.z; P.push_back(v); }
        else if (tag == "vn") { Vec3 v; ls >> v.x >> v.y >> v.z; N.push_back(v); }
        else if (tag == "vt") { Vec2 v; ls >> v.x >> v.y; T.push_back(v); }
        else if (tag == "f") {
            struct Idx { long p = -1, t = -1, n = -1; };
            std::vector<Idx> face;
            std::string tok;
            while (ls >> tok) {
                Idx id;
                long vals[3] = {0, 0, 0};
                bool has[3] = {false, false, false};
                size_t field = 0, start = 0;
                for (size_t i = 0; i <= tok.size(); ++i) {
                    if (i == tok.size() || tok[i] == '/') {
                        if (i > start) { vals[field] = std::stol(tok.substr(start, i - start)); has[field] = true; }
                        ++field;
                        start = i + 1;
                        if (field > 2) break;
                    }
                }
                if (!has[0]) throw ObjectConstructionException("invalid OBJ face in " + filename);
                id.p = resolve(vals[0], P.size());
                if (has[1]) id.t = resolve(vals[1], T.size());
                if (has[2]) id.n = resolve(vals[2], N.size());
                face.push_back(id);
            }
            for (size_t k = 1; k + 1 < face.size(); ++k) {
                const Idx tri[3] = {face[0], face[k], face[k + 1]};
                Vec3 fn{0, 0, 0};
                {
                    const Vec3 &a = P.at(size_t(tri[0].p)), &b = P.at(size_t(tri[1].p)), &c = P.at(size_t(tri[2].p));
                    const float ux = b.x - a.x, uy = b.y - a.y, uz = b.z - a.z;
                    const float vx = c.x - a.x, vy = c.y - a.y, vz = c.z - a.z;
                    fn = {uy * vz - uz * vy, uz * vx - ux * vz, ux * vy - uy * vx};
                }
                for (const Idx &id : tri) {
                    agx_vertex v{};
                    const Vec3 &p = P.at(size_t(id.p));
                    v.pos[0] = p.x; v.pos[1] = p.y; v.pos[2] = p.z;
                    const Vec3 n = id.n >= 0 ? N.at(size_t(id.n)) : fn;
                    v.nor[0] = n.x; v.nor[1] = n.y; v.nor[2] = n.z;
                    if (id.t >= 0) { v.uv[0] = T.at(size_t(id.t)).x; v.uv[1] = T.at(size_t(id.t)).y; }
                    out.push_back(v);
                }
            }
        }
    }
    if (out.empty()) throw ObjectConstructionException("empty mesh: " + filename);
    return out;
}

void save_bin_mesh(const std::string &filename, const std::vector<agx_vertex> &tris) {
    std::ofstream o(filename, std::ios::binary);
    if (!o) throw ObjectConstructionException("failed to write " + filename);
    const uint64_t n = tris.size() / 3;
    o.write(reinterpret_cast<const char *>(&n), sizeof(n));
    o.write(reinterpret_cast<const char *>(tris.data()), std::streamsize(tris.size() * sizeof(agx_vertex)));
}

// ---------------------------------------------------------------- scene descriptor
const agx_scene_desc &Scene::desc() {
    ents_.clear();
    mats_.clear();
    std::map<const Material *, int32_t> mat_index;
    for (auto &e : entities) {
        agx_entity_desc d = e->desc;
        auto it = mat_index.find(e->material.get());
        if (it == mat_index.end()) {
            it = mat_index.emplace(e->material.get(), int32_t(mats_.size())).first;
            mats_.push_back(e->material->desc);
        }
        d.material = it->second;
        if (e->mesh) {
            d.tris = e->mesh->data();
            d.tri_count = uint32_t(e->mesh->size() / 3);
        }
        ents_.push_back(d);
    }
    desc_ = agx_scene_desc{};
    desc_.abi_version = AGX_ABI_VERSION;
    desc_.entity_count = uint32_t(ents_.size());
    desc_.entities = ents_.data();
    desc_.material_count = uint32_t(mats_.size());
    desc_.materials = mats_.data();
    if (env) desc_.env = env->desc;
    desc_.aggregate = aggregate.kind;
    desc_.max_leaf_size = aggregate.max_leaf_size;
    return desc_;
}

// ---------------------------------------------------------------- creators
namespace {

void copy3(float *dst, const Vec3 &v) { dst[0] = v.x; dst[1] = v.y; dst[2] = v.z; }

Vec3 texture_value(CreatingContext &ctx, const ConfigGroup &params, const std::string &name) {
    return ctx.create<Texture2D>(params.child_group(name))->texel;
}

void set_transforms(agx_entity_desc &d, const Transform3 &l2w) {
    l2w.to_rows12(d.local_to_world);
    l2w.inverse().to_rows12(d.world_to_local);
}

void reject_normal_map(const ConfigGroup &params) {
    if (params.find_child("normal_map"))
        throw ObjectConstructionException("normal maps are not supported by gpu_pt");
}

}  // namespace

CreatingContext::CreatingContext() {
    // texture2d: only constant textures reach the GPU path (texture2d_creators.cpp:50-65)
    texture2d.add_creator("constant", [](const ConfigGroup &p, CreatingContext &) {
        auto t = std::make_shared<Texture2D>();
        t->texel = p.child_spectrum("texel");
        return t;
    });
    for (const char *unsupported : {"image", "hdr", "checker_board", "solid_image"})
        texture2d.add_creator(unsupported, [n = std::string(unsupported)](const ConfigGroup &, CreatingContext &)
                                               -> std::shared_ptr<Texture2D> {
            throw ObjectConstructionException("texture type '" + n + "' is not supported by gpu_pt");
        });

    // geometry (geometry_creators.cpp)
    geometry.add_creator("quad", [](const ConfigGroup &p, CreatingContext &) {
        auto g = std::make_shared<Geometry>();
        g->proto.geometry = AGX_GEOM_QUAD;
        const Transform3 l2w = p.child_transform3("transform");
        const char *names[4] = {"A", "B", "C", "D"};
        const char *tnames[4] = {"tA", "tB", "tC", "tD"};
        for (int k = 0; k < 4; ++k) {
            copy3(g->proto.p[k], l2w.apply_to_point(p.child_vec3(names[k])));  // Quad applies the transform
            const Vec2 t = p.child_vec2_or(tnames[k], Vec2{0, 0});
            g->proto.t[k][0] = t.x;
            g->proto.t[k][1] = t.y;
        }
        set_transforms(g->proto, Transform3());
        return g;
    });
    geometry.add_creator("triangle", [](const ConfigGroup &p, CreatingContext &) {
        auto g = std::make_shared<Geometry>();
        g->proto.geometry = AGX_GEOM_TRIANGLE;
        const char *names[3] = {"A", "B", "C"};
        const char *tnames[3] = {"tA", "tB", "tC"};
        for (int k = 0; k < 3; ++k) {
            copy3(g->proto.p[k], p.child_vec3(names[k]));
            const Vec2 t = p.child_vec2_or(tnames[k], Vec2{0, 0});
            g->proto.t[k][0] = t.x;
            g->proto.t[k][1] = t.y;
        }
        set_transforms(g->proto, p.child_transform3("transform"));
        return g;
    });
    for (const char *kind : {"sphere", "disk"})
        geometry.add_creator(kind, [sphere = std::string(kind) == "sphere"](const ConfigGroup &p, CreatingContext &) {
            auto g = std::make_shared<Geometry>();
            g->proto.geometry = sphere ? AGX_GEOM_SPHERE : AGX_GEOM_DISK;
            g->proto.radius = p.child_real("radius");
            if (!(g->proto.radius > 0)) throw ObjectConstructionException("invalid radius");
            set_transforms(g->proto, p.child_transform3("transform"));
            return g;
        });
    for (const char *kind : {"triangle_bvh", "triangle_bvh_noembree", "triangle_bvh_embree"})
        geometry.add_creator(kind, [](const ConfigGroup &p, CreatingContext &ctx) {
            auto g = std::make_shared<Geometry>();
            g->proto.geometry = AGX_GEOM_TRIANGLE_BVH;
            const Transform3 l2w = p.child_transform3("transform");
            const std::string file = ctx.path_mapper ? ctx.path_mapper->map(p.child_str("filename"))
                                                     : p.child_str("filename");
            auto tris = std::make_shared<std::vector<agx_vertex>>(load_triangle_mesh(file));
            // TriangleBVH::load (triangle_bvh.cpp:533-558): world-space positions and normals
            for (agx_vertex &v : *tris) {
                const Vec3 wp = l2w.apply_to_point(Vec3{v.pos[0], v.pos[1], v.pos[2]});
                const Vec3 wn = l2w.apply_to_vector(Vec3{v.nor[0], v.nor[1], v.nor[2]});
                copy3(v.pos, wp);
                copy3(v.nor, wn);
            }
            g->mesh = tris;
            set_transforms(g->proto, Transform3());
            return g;
        });
    for (const char *unsupported : {"double_sided", "transform_wrapper"})
        geometry.add_creator(unsupported, [n = std::string(unsupported)](const ConfigGroup &, CreatingContext &)
                                              -> std::shared_ptr<Geometry> {
            throw ObjectConstructionException("geometry type '" + n + "' is not supported by gpu_pt yet");
        });

    // material (material_creators.cpp)
    material.add_creator("ideal_diffuse", [](const ConfigGroup &p, CreatingContext &ctx) {
        reject_normal_map(p);
        auto m = std::make_shared<Material>();
        m->desc.kind = AGX_MAT_IDEAL_DIFFUSE;
        copy3(m->desc.color, texture_value(ctx, p, "albedo"));
        return m;
    });
    material.add_creator("metal", [](const ConfigGroup &p, CreatingContext &ctx) {
        reject_normal_map(p);
        auto m = std::make_shared<Material>();
        m->desc.kind = AGX_MAT_METAL;
        copy3(m->desc.color, texture_value(ctx, p, "color"));
        copy3(m->desc.eta, texture_value(ctx, p, "eta"));  // kept as written; the swap happens in the library
        copy3(m->desc.k, texture_value(ctx, p, "k"));
        m->desc.roughness = texture_value(ctx, p, "roughness").x;
        m->desc.anisotropic = texture_value(ctx, p, "anisotropic").x;
        return m;
    });
    material.add_creator("glass", [](const ConfigGroup &p, CreatingContext &ctx) {
        auto m = std::make_shared<Material>();
        m->desc.kind = AGX_MAT_GLASS;
        m->desc.ior = texture_value(ctx, p, "ior").x;
        bool refl = false, refr = false;
        if (auto n = p.find_child_group("color_map")) {
            const Vec3 c = ctx.create<Texture2D>(*n)->texel;
            copy3(m->desc.color, c);
            copy3(m->desc.color2, c);
            refl = refr = true;
        }
        if (auto n = p.find_child_group("color_reflection_map")) {
            copy3(m->desc.color, ctx.create<Texture2D>(*n)->texel);
            refl = true;
            m->desc.glass_separate_refraction = 1;
        }
        if (auto n = p.find_child_group("color_refraction_map")) {
            copy3(m->desc.color2, ctx.create<Texture2D>(*n)->texel);
            refr = true;
            m->desc.glass_separate_refraction = 1;
        }
        if (!refl || !refr) throw CreatingObjectException("empty color reflection/refraction map");
        if (p.find_child("bssrdf_dmfp")) throw ObjectConstructionException("BSSRDF glass is not supported by gpu_pt");
        return m;
    });
    material.add_creator("mirror", [](const ConfigGroup &p, CreatingContext &ctx) {
        auto m = std::make_shared<Material>();
        m->desc.kind = AGX_MAT_MIRROR;
        copy3(m->desc.color, texture_value(ctx, p, "color_map"));
        copy3(m->desc.eta, texture_value(ctx, p, "eta"));
        copy3(m->desc.k, texture_value(ctx, p, "k"));
        return m;
    });
    material.add_creator("ideal_black", [](const ConfigGroup &, CreatingContext &) {
        auto m = std::make_shared<Material>();
        m->desc.kind = AGX_MAT_IDEAL_BLACK;
        return m;
    });
    for (const char *unsupported : {"disney", "dream_works_fabric", "invisible_surface", "paper", "phong"})
        material.add_creator(unsupported, [n = std::string(unsupported)](const ConfigGroup &, CreatingContext &)
                                              -> std::shared_ptr<Material> {
            throw ObjectConstructionException("material type '" + n + "' is not supported by gpu_pt yet");
        });

    // medium (medium_creators.cpp): only "void" on this path
    medium.add_creator("void", [](const ConfigGroup &, CreatingContext &) {
        auto m = std::make_shared<Medium>();
        m->type = "void";
        return m;
    });
    for (const char *unsupported : {"homogeneous", "heterogeneous"})
        medium.add_creator(unsupported, [n = std::string(unsupported)](const ConfigGroup &, CreatingContext &)
                                            -> std::shared_ptr<Medium> {
            throw ObjectConstructionException("medium type '" + n + "' is not supported by gpu_pt yet");
        });

    // entity (entity_creators.cpp:29-60)
    entity.add_creator("geometric", [](const ConfigGroup &p, CreatingContext &ctx) {
        auto e = std::make_shared<Entity>();
        auto g = ctx.create<Geometry>(p.child_group("geometry"));
        e->material = ctx.create<Material>(p.child_group("material"));
        if (auto n = p.find_child_group("med_in")) ctx.create<Medium>(*n);
        if (auto n = p.find_child_group("med_out")) ctx.create<Medium>(*n);
        e->desc = g->proto;
        e->mesh = g->mesh;
        e->desc.medium_in = -1;
        e->desc.medium_out = -1;
        copy3(e->desc.emit_radiance, p.child_spectrum_or("emit_radiance", Vec3{0, 0, 0}));
        e->desc.no_denoise = p.child_int_or("no_denoise", 0);
        e->desc.power = p.child_real_or("power", -1);
        return e;
    });

    // environment light (envir_light_creators.cpp)
    envir_light.add_creator("native_sky", [](const ConfigGroup &p, CreatingContext &) {
        auto l = std::make_shared<EnvirLight>();
        l->desc.kind = AGX_ENV_NATIVE_SKY;
        copy3(l->desc.top, p.child_spectrum("top"));
        copy3(l->desc.bottom, p.child_spectrum("bottom"));
        l->desc.power = p.child_real_or("power", -1);
        return l;
    });
    envir_light.add_creator("ibl", [](const ConfigGroup &, CreatingContext &) -> std::shared_ptr<EnvirLight> {
        throw ObjectConstructionException("environment light 'ibl' is not supported by gpu_pt yet");
    });

    // aggregate (aggregate_creators.cpp)
    aggregate.add_creator("native", [](const ConfigGroup &, CreatingContext &) {
        auto a = std::make_shared<Aggregate>();
        a->kind = AGX_AGGREGATE_NATIVE;
        return a;
    });
    aggregate.add_creator("bvh", [](const ConfigGroup &p, CreatingContext &) {
        auto a = std::make_shared<Aggregate>();
        a->kind = AGX_AGGREGATE_BVH;
        a->max_leaf_size = p.child_int_or("max_leaf_size", 5);
        if (a->max_leaf_size < 1) throw ObjectConstructionException("invalid max_leaf_size value");
        return a;
    });

    // camera (camera_creators.cpp)
    camera.add_creator("thin_lens", [](const ConfigGroup &p, CreatingContext &ctx) {
        auto c = std::make_shared<Camera>();
        copy3(c->desc.pos, p.child_vec3("pos"));
        copy3(c->desc.dst, p.child_vec3("dst"));
        copy3(c->desc.up, p.child_vec3("up"));
        c->desc.fov_rad = p.child_real("fov") * (3.14159265358979323846f / 180);
        c->desc.lens_radius = p.child_real_or("lens_radius", 0);
        c->desc.focal_distance = p.child_real_or("focal_distance", 1);
        c->desc.film_aspect = ctx.film_aspect;
        return c;
    });
    camera.add_creator("pinhole", [](const ConfigGroup &p, CreatingContext &ctx) {
        auto c = std::make_shared<Camera>();
        copy3(c->desc.pos, p.child_vec3("pos"));
        copy3(c->desc.dst, p.child_vec3("dst"));
        copy3(c->desc.up, p.child_vec3("up"));
        const float width = p.child_real("width");
        const float height = width / ctx.film_aspect;
        const float dist = p.child_real("dist");
        c->desc.fov_rad = 2 * std::atan(height / (2 * dist));
        c->desc.lens_radius = 0;
        c->desc.focal_distance = 1;
        c->desc.film_aspect = ctx.film_aspect;
        return c;
    });

    // film filter (film_filter_creators.cpp)
    film_filter.add_creator("box", [](const ConfigGroup &p, CreatingContext &) {
        auto f = std::make_shared<FilmFilter>();
        f->type = "box";
        f->radius = p.child_real("radius");
        return f;
    });
    film_filter.add_creator("gaussian", [](const ConfigGroup &p, CreatingContext &) {
        auto f = std::make_shared<FilmFilter>();
        f->type = "gaussian";
        f->radius = p.child_real("radius");
        f->alpha = p.child_real("alpha");
        return f;
    });

    // scene (scene_creators.cpp:15-66)
    scene.add_creator("default", [](const ConfigGroup &p, CreatingContext &ctx) {
        auto s = std::make_shared<Scene>();
        if (auto arr = p.find_child_array("entities")) {
            for (size_t i = 0; i < arr->size(); ++i) {
                const ConfigGroup &g = arr->at_group(i);
                if (ends_with(g.child_str("type"), "//")) continue;
                s->entities.push_back(ctx.create<Entity>(g));
            }
        }
        if (auto g = p.find_child_group("env")) s->env = ctx.create<EnvirLight>(*g);
        if (auto g = p.find_child_group("aggregate")) s->aggregate = *ctx.create<Aggregate>(*g);
        return s;
    });

    register_renderers(*this);
}

template std::shared_ptr<Texture2D> CreatingContext::create<Texture2D>(const ConfigGroup &);
template std::shared_ptr<Geometry> CreatingContext::create<Geometry>(const ConfigGroup &);
template std::shared_ptr<Material> CreatingContext::create<Material>(const ConfigGroup &);
template std::shared_ptr<Medium> CreatingContext::create<Medium>(const ConfigGroup &);
template std::shared_ptr<Entity> CreatingContext::create<Entity>(const ConfigGroup &);
template std::shared_ptr<EnvirLight> CreatingContext::create<EnvirLight>(const ConfigGroup &);
template std::shared_ptr<Aggregate> CreatingContext::create<Aggregate>(const ConfigGroup &);

}  // namespace atrc
