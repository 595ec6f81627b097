// Creators of the Atrc object families on the gpu_pt path.
// Parameter names and defaults follow src/factory/src/creator/*_creators.cpp.
#include "factory.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <limits>
#include <fstream>
#include <sstream>

#include "renderer.h"

namespace atrc {

// ---------------------------------------------------------------- factory plumbing
template <> Factory<Texture2D> &CreatingContext::factory<Texture2D>() { return texture2d; }
template <> Factory<Texture3D> &CreatingContext::factory<Texture3D>() { return texture3d; }
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
// STL triangle soup (binary: 80-byte header, uint32 count, then per triangle
// a facet normal, three vertices and a uint16 attribute; ASCII: "solid" ...
// "facet normal" / "vertex" lines).  Each corner takes the facet normal, or
// (B - A) x (C - A) when the file's normal is zero; uv = 0.  (Atrc reads STL
// through agz-utils' mesh loader, which is not in the source tree: its choice
// of corner normals is not pinned.)
std::vector<agx_vertex> load_stl(const std::string &filename) {
    std::ifstream in(filename, std::ios::binary);
    if (!in) throw ObjectConstructionException("failed to open mesh file: " + filename);
    std::string bytes((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    std::vector<agx_vertex> out;
    auto emit = [&out](const float n[3], const float v[9]) {
        float nn[3] = {n[0], n[1], n[2]};
        if (nn[0] == 0 && nn[1] == 0 && nn[2] == 0) {
            const float ux = v[3] - v[0], uy = v[4] - v[1], uz = v[5] - v[2];
            const float vx = v[6] - v[0], vy = v[7] - v[1], vz = v[8] - v[2];
            nn[0] = uy * vz - uz * vy; nn[1] = uz * vx - ux * vz; nn[2] = ux * vy - uy * vx;
        }
        for (int k = 0; k < 3; ++k) {
            agx_vertex a{};
            for (int j = 0; j < 3; ++j) { a.pos[j] = v[3 * k + j]; a.nor[j] = nn[j]; }
            out.push_back(a);
        }
    };
    // binary when the size matches the triangle count (ASCII files may also start with "solid")
    if (bytes.size() >= 84) {
        uint32_t n = 0;
        std::memcpy(&n, bytes.data() + 80, 4);
        if (bytes.size() == 84 + size_t(n) * 50) {
            for (uint32_t i = 0; i < n; ++i) {
                float f[12];
                std::memcpy(f, bytes.data() + 84 + size_t(i) * 50, sizeof(f));
                emit(f, f + 3);
            }
            if (out.empty()) throw ObjectConstructionException("empty mesh: " + filename);
            return out;
        }
    }
    std::istringstream ss(bytes);
    std::string tok;
    float n[3] = {0, 0, 0}, v[9];
    int nv = 0;
    while (ss >> tok) {
        if (tok == "facet") {
            std::string normal;
            ss >> normal >> n[0] >> n[1] >> n[2];
            nv = 0;
        } else if (tok == "vertex") {
            if (nv >= 3) throw ObjectConstructionException("invalid STL facet in " + filename);
            ss >> v[3 * nv] >> v[3 * nv + 1] >> v[3 * nv + 2];
            ++nv;
        } else if (tok == "endfacet") {
            if (nv != 3) throw ObjectConstructionException("invalid STL facet in " + filename);
            emit(n, v);
        }
    }
    if (!ss.eof()) throw ObjectConstructionException("invalid STL file: " + filename);
    if (out.empty()) throw ObjectConstructionException("empty mesh: " + filename);
    return out;
}

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
    if (ends_with(filename, ".stl")) return load_stl(filename);
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
    meds_.clear();
    std::map<const Material *, int32_t> mat_index;
    std::map<const Medium *, int32_t> med_index;
    // one descriptor per distinct medium object; "void" is index -1
    auto medium_index = [&](const std::shared_ptr<Medium> &m) -> int32_t {
        if (!m || m->desc.kind == AGX_MEDIUM_VOID) return -1;
        auto it = med_index.find(m.get());
        if (it == med_index.end()) {
            it = med_index.emplace(m.get(), int32_t(meds_.size())).first;
            meds_.push_back(m->desc);
        }
        return it->second;
    };
    for (auto &e : entities) {
        agx_entity_desc d = e->desc;
        auto it = mat_index.find(e->material.get());
        if (it == mat_index.end()) {
            it = mat_index.emplace(e->material.get(), int32_t(mats_.size())).first;
            mats_.push_back(e->material->desc);
        }
        d.material = it->second;
        d.medium_in = medium_index(e->med_in);
        d.medium_out = medium_index(e->med_out);
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
    desc_.medium_count = uint32_t(meds_.size());
    desc_.media = meds_.empty() ? nullptr : meds_.data();
    return desc_;
}

// ---------------------------------------------------------------- creators
namespace {

void copy3(float *dst, const Vec3 &v) { dst[0] = v.x; dst[1] = v.y; dst[2] = v.z; }

// material parameters on this path are constant textures
Vec3 constant_texel(const Texture2D &t) {
    if (t.type != "constant")
        throw ObjectConstructionException("texture type '" + t.type + "' is not supported by gpu_pt materials");
    return t.texel;
}

Vec3 texture_value(CreatingContext &ctx, const ConfigGroup &params, const std::string &name) {
    return constant_texel(*ctx.create<Texture2D>(params.child_group(name)));
}

// to_transform2 (utility/config.cpp:136-190): translate / rotate / scale in the uv
// plane, kept as a Transform3 that leaves the third coordinate alone; an array
// is the product of its elements
Transform3 to_transform2(const ConfigNode &node) {
    if (node.is_array()) {
        const auto &arr = node.as_array();
        Transform3 r;
        for (size_t i = 0; i < arr.size(); ++i) r = r * to_transform2(arr.at(i));
        return r;
    }
    const auto &g = node.as_group();
    const std::string &type = g.child_str("type");
    if (ends_with(type, "//")) return Transform3();
    auto vec2 = [&](const std::string &name) {
        const auto &a = g.child_array(name);
        if (a.size() == 1) return std::pair<double, double>(a.at_real(0), a.at_real(0));
        if (a.size() == 2) return std::pair<double, double>(a.at_real(0), a.at_real(1));
        throw ConfigException("invalid array size for vec2 (" + std::to_string(a.size()) + ")");
    };
    if (type == "translate") {
        const auto o = vec2("offset");
        return Transform3::translate(o.first, o.second, 0);
    }
    if (type == "rotate") {
        double rad;
        if (auto p = g.find_child("rad")) rad = p->as_value().as_real();
        else rad = g.child_real("deg") / 180 * 3.14159265358979323846f;
        return Transform3::rotate(0, 0, 1, rad);
    }
    if (type == "scale") {
        if (g.child("ratio").is_array()) {
            const auto s = vec2("ratio");
            return Transform3::scale(s.first, s.second, 1);
        }
        const double s = g.child_real("ratio");
        return Transform3::scale(s, s, 1);
    }
    throw ConfigException("unknown transform2 type (type = " + type + ")");
}

// Texture2DCommonParams (core/texture2d.h:12-50): inv_v, inv_u, swap_uv, then the
// user transform (applied first); wrap modes and inv_gamma
void tex2d_common(const ConfigGroup &p, agx_texture2d_desc *d) {
    Transform3 ret;
    if (p.child_int_or("inv_v", 0)) ret = ret * (Transform3::translate(0, 1, 0) * Transform3::scale(1, -1, 1));
    if (p.child_int_or("inv_u", 0)) ret = ret * (Transform3::translate(1, 0, 0) * Transform3::scale(-1, 1, 1));
    if (p.child_int_or("swap_uv", 0)) {
        Transform3 s;
        s.m[0][0] = 0; s.m[0][1] = 1; s.m[1][0] = 1; s.m[1][1] = 0;
        ret = ret * s;
    }
    if (auto n = p.find_child("transform")) ret = ret * to_transform2(*n);
    const float row[6] = {float(ret.m[0][0]), float(ret.m[0][1]), float(ret.m[0][3]),
                          float(ret.m[1][0]), float(ret.m[1][1]), float(ret.m[1][3])};
    for (int i = 0; i < 6; ++i) d->transform[i] = row[i];
    const auto wrap = [](const std::string &w, const char *axis) {
        if (w == "clamp") return int32_t(AGX_WRAP_CLAMP);
        if (w == "repeat") return int32_t(AGX_WRAP_REPEAT);
        if (w == "mirror") return int32_t(AGX_WRAP_MIRROR);
        throw ObjectConstructionException(std::string("invalid wrap_") + axis + " value: " + w +
                                          " (expect clamp/repeat/mirror)");
    };
    d->wrap[0] = wrap(p.child_str_or("wrap_u", "clamp"), "u");
    d->wrap[1] = wrap(p.child_str_or("wrap_v", "clamp"), "v");
    d->inv_gamma = p.child_real_or("inv_gamma", 1);
}

// Radiance RGBE (.hdr) reader: "#?RADIANCE"/"#?RGBE" header, "-Y H +X W"
// resolution, flat or new-style run-length scanlines; texel = mantissa *
// 2^(e - 136), rows top to bottom as stored
std::shared_ptr<std::vector<float>> load_hdr(const std::string &filename, int *w, int *h) {
    std::ifstream f(filename, std::ios::binary);
    if (!f) throw ObjectConstructionException("failed to load texture from " + filename);
    std::string line;
    std::getline(f, line);
    if (line.rfind("#?", 0) != 0) throw ObjectConstructionException("not a radiance hdr file: " + filename);
    for (;;) {
        if (!std::getline(f, line)) throw ObjectConstructionException("truncated hdr header: " + filename);
        if (line.empty() || line == "\r") break;
        if (line.rfind("FORMAT=", 0) == 0 && line.find("32-bit_rle_rgbe") == std::string::npos)
            throw ObjectConstructionException("unsupported hdr format: " + line);
    }
    std::getline(f, line);
    char ya[3] = {0}, xa[3] = {0};
    int H = 0, W = 0;
    if (std::sscanf(line.c_str(), "%2s %d %2s %d", ya, &H, xa, &W) != 4 || std::string(ya) != "-Y" ||
        std::string(xa) != "+X" || W <= 0 || H <= 0 || int64_t(W) * H > (int64_t(1) << 28))
        throw ObjectConstructionException("unsupported hdr resolution line: " + line);
    auto out = std::make_shared<std::vector<float>>(size_t(W) * size_t(H) * 3);
    std::vector<unsigned char> scan(size_t(W) * 4);
    auto byte = [&]() {
        const int c = f.get();
        if (c == EOF) throw ObjectConstructionException("truncated hdr data: " + filename);
        return static_cast<unsigned char>(c);
    };
    for (int y = 0; y < H; ++y) {
        unsigned char head[4] = {byte(), byte(), byte(), byte()};
        if (W >= 8 && W < 32768 && head[0] == 2 && head[1] == 2 && ((head[2] << 8) | head[3]) == W) {
            for (int c = 0; c < 4; ++c)
                for (int x = 0; x < W;) {
                    int n = byte();
                    if (n > 128) {
                        n -= 128;
                        const unsigned char v = byte();
                        if (x + n > W) throw ObjectConstructionException("bad hdr run length: " + filename);
                        for (int k = 0; k < n; ++k) scan[size_t(x++) * 4 + size_t(c)] = v;
                    } else {
                        if (n == 0 || x + n > W) throw ObjectConstructionException("bad hdr run length: " + filename);
                        for (int k = 0; k < n; ++k) scan[size_t(x++) * 4 + size_t(c)] = byte();
                    }
                }
        } else {
            for (int c = 0; c < 4; ++c) scan[size_t(c)] = head[c];
            for (int x = 1; x < W; ++x)
                for (int c = 0; c < 4; ++c) scan[size_t(x) * 4 + size_t(c)] = byte();
        }
        for (int x = 0; x < W; ++x) {
            const unsigned char *p = &scan[size_t(x) * 4];
            const float s = p[3] ? std::ldexp(1.0f, int(p[3]) - 136) : 0.0f;
            float *o = out->data() + (size_t(y) * size_t(W) + size_t(x)) * 3;
            o[0] = float(p[0]) * s;
            o[1] = float(p[1]) * s;
            o[2] = float(p[2]) * s;
        }
    }
    *w = W;
    *h = H;
    return out;
}

void set_transforms(agx_entity_desc &d, const Transform3 &l2w) {
    l2w.to_rows12(d.local_to_world);
    l2w.inverse().to_rows12(d.world_to_local);
}

// Texture3DCommonParams::full_transform (core/texture3d.h:89-125): flips, the
// uvw permutation, then the user transform; the product applies the user
// transform first (agz-utils Transform3 product order, parity unpinned).
Transform3 tex3d_full_transform(const ConfigGroup &p) {
    Transform3 ret;
    if (p.child_int_or("inv_u", 0)) ret = ret * (Transform3::translate(1, 0, 0) * Transform3::scale(-1, 1, 1));
    if (p.child_int_or("inv_v", 0)) ret = ret * (Transform3::translate(0, 1, 0) * Transform3::scale(1, -1, 1));
    if (p.child_int_or("inv_w", 0)) ret = ret * (Transform3::translate(0, 0, 1) * Transform3::scale(1, 1, -1));
    int perm[3] = {0, 1, 2};
    if (p.find_child("uvw_perm")) {
        const ConfigArray &a = p.child_array("uvw_perm");
        if (a.size() != 3) throw ObjectConstructionException("uvw_perm needs 3 integers");
        for (int i = 0; i < 3; ++i) perm[i] = a.at_int(size_t(i));
    }
    Transform3 P;
    for (int r = 0; r < 3; ++r) {
        if (perm[r] < 0 || perm[r] >= 3) throw ObjectConstructionException("invalid uvw perm value");
        for (int c = 0; c < 4; ++c) P.m[r][c] = 0;
        P.m[r][perm[r]] = 1;
    }
    ret = ret * P;
    if (p.find_child("transform")) ret = ret * p.child_transform3("transform");
    return ret;
}

int32_t wrap_mode(const std::string &s) {  // Texture3D::str_to_wrap_func (texture3d.h:165-175)
    if (s == "clamp") return AGX_WRAP_CLAMP;
    if (s == "repeat") return AGX_WRAP_REPEAT;
    if (s == "mirror") return AGX_WRAP_MIRROR;
    throw ObjectConstructionException("invalid wrap function name: " + s);
}

// image3d texel files (factory/src/utility/texture3d_loader.cpp): a header of
// width, height, depth, then texels in (z, y, x) order; "real" = 1 float,
// "spec" = 3 floats, "gray8" = 1 byte / 255, "rgb8" = 3 bytes / 255.  The ASCII
// gray8/rgb8 readers extract single characters, as Atrc's operator>> on uint8_t does.
std::shared_ptr<std::vector<float>> load_texels(const std::string &filename, bool binary, const std::string &format,
                                                int *w, int *h, int *d, int *ch) {
    std::ifstream f(filename, binary ? std::ios::in | std::ios::binary : std::ios::in);
    if (!f) throw ObjectConstructionException("failed to open file: " + filename);
    int32_t W = 0, H = 0, D = 0;
    if (binary) {
        f.read(reinterpret_cast<char *>(&W), 4);
        f.read(reinterpret_cast<char *>(&H), 4);
        f.read(reinterpret_cast<char *>(&D), 4);
    } else {
        f >> W >> H >> D;
    }
    if (!f || W <= 0 || H <= 0 || D <= 0) throw ObjectConstructionException("failed to read width/height/depth from file");
    int C;
    if (format == "real" || format == "gray8") C = 1;
    else if (format == "spec" || format == "rgb8") C = 3;
    else throw ObjectConstructionException("unknown image3d format: " + format);
    const size_t n = size_t(W) * size_t(H) * size_t(D) * size_t(C);
    auto out = std::make_shared<std::vector<float>>(n);
    const bool bytes = format == "gray8" || format == "rgb8";
    for (size_t i = 0; i < n; ++i) {
        if (bytes) {
            unsigned char c = 0;
            if (binary) f.read(reinterpret_cast<char *>(&c), 1);
            else f >> c;
            (*out)[i] = float(c) / 255.0f;
        } else {
            float v = 0;
            if (binary) f.read(reinterpret_cast<char *>(&v), 4);
            else f >> v;
            (*out)[i] = v;
        }
        if (!f) throw ObjectConstructionException("failed to parse texel data from file");
    }
    *w = W; *h = H; *d = D; *ch = C;
    return out;
}

// helper::child_texture_or_constant (factory/include/agz/factory/help.h:11-19): the
// constant texel of child `name`, or `def` when the child is absent
Vec3 texture_value_or(CreatingContext &ctx, const ConfigGroup &params, const std::string &name, float def) {
    if (auto n = params.find_child_group(name)) return constant_texel(*ctx.create<Texture2D>(*n));
    return Vec3{def, def, def};
}

void reject_normal_map(const ConfigGroup &params) {
    if (params.find_child("normal_map"))
        throw ObjectConstructionException("normal maps are not supported by gpu_pt");
}

}  // namespace

CreatingContext::CreatingContext() {
    // texture2d (texture2d_creators.cpp:10-166): constant, hdr and checker_board
    // (materials take constant textures only; hdr / checker feed the ibl light)
    texture2d.add_creator("constant", [](const ConfigGroup &p, CreatingContext &) {
        auto t = std::make_shared<Texture2D>();
        t->texel = p.child_spectrum("texel");
        tex2d_common(p, &t->desc);
        t->desc.kind = AGX_TEX2D_CONSTANT;
        copy3(t->desc.value, t->texel);
        return t;
    });
    texture2d.add_creator("checker_board", [](const ConfigGroup &p, CreatingContext &) {
        auto t = std::make_shared<Texture2D>();
        t->type = "checker_board";
        tex2d_common(p, &t->desc);
        t->desc.kind = AGX_TEX2D_CHECKER;
        t->desc.grid_count = p.child_int("grid_count");
        if (t->desc.grid_count <= 0)
            throw ObjectConstructionException("invalid grid_count value: " + std::to_string(t->desc.grid_count));
        copy3(t->desc.value, p.child_spectrum("color1"));
        copy3(t->desc.value2, p.child_spectrum("color2"));
        return t;
    });
    texture2d.add_creator("hdr", [](const ConfigGroup &p, CreatingContext &ctx) {
        auto t = std::make_shared<Texture2D>();
        t->type = "hdr";
        tex2d_common(p, &t->desc);
        const std::string sample = p.child_str_or("sample", "linear");
        if (sample != "linear" && sample != "nearest") throw ObjectConstructionException("invalid sample method");
        int w = 0, h = 0;
        const std::string file = ctx.path_mapper ? ctx.path_mapper->map(p.child_str("filename")) : p.child_str("filename");
        t->image = load_hdr(file, &w, &h);
        t->desc.kind = AGX_TEX2D_IMAGE;
        t->desc.linear = sample == "linear" ? 1 : 0;
        t->desc.width = w;
        t->desc.height = h;
        t->desc.data = t->image->data();
        return t;
    });
    for (const char *unsupported : {"image", "solid_image"})
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
    // DoubleSidedGeometry (double_sided.cpp:22-100): the wrapped geometry with the
    // double_sided flag; the library turns back-side frames around and halves
    // the light pdf
    geometry.add_creator("double_sided", [](const ConfigGroup &p, CreatingContext &ctx) {
        auto in = ctx.create<Geometry>(p.child_group("internal"));
        if (in->proto.double_sided)
            throw ObjectConstructionException("double_sided of a double_sided geometry is not supported by gpu_pt");
        auto g = std::make_shared<Geometry>(*in);
        g->proto.double_sided = 1;
        return g;
    });
    // TransformWrapper (transform_wrapper.cpp:5-140): the wrapper's transform is
    // composed into the wrapped geometry at load (quad corners and mesh
    // vertices move, the local-to-world of triangle / sphere / disk is
    // premultiplied).  Atrc intersects in the wrapper's local space instead, so
    // hits agree up to rounding; its light pdf scales by the z axis' length
    // squared, which equals the exact area ratio for uniform scales.
    geometry.add_creator("transform_wrapper", [](const ConfigGroup &p, CreatingContext &ctx) {
        auto in = ctx.create<Geometry>(p.child_group("internal"));
        const Transform3 T = p.child_transform3("transform");
        auto g = std::make_shared<Geometry>(*in);
        switch (g->proto.geometry) {
        case AGX_GEOM_QUAD:
            for (int k = 0; k < 4; ++k)
                copy3(g->proto.p[k], T.apply_to_point(Vec3{g->proto.p[k][0], g->proto.p[k][1], g->proto.p[k][2]}));
            break;
        case AGX_GEOM_TRIANGLE_BVH: {
            auto tris = std::make_shared<std::vector<agx_vertex>>(*in->mesh);
            for (agx_vertex &v : *tris) {
                copy3(v.pos, T.apply_to_point(Vec3{v.pos[0], v.pos[1], v.pos[2]}));
                copy3(v.nor, T.apply_to_vector(Vec3{v.nor[0], v.nor[1], v.nor[2]}));
            }
            g->mesh = tris;
            break;
        }
        default: {
            Transform3 l2w;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 4; ++j) l2w.m[i][j] = g->proto.local_to_world[4 * i + j];
            set_transforms(g->proto, T * l2w);
            break;
        }
        }
        return g;
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
    material.add_creator("invisible_surface", [](const ConfigGroup &p, CreatingContext &) {
        if (p.find_child("bssrdf_dmfp")) throw ObjectConstructionException("BSSRDF surfaces are not supported by gpu_pt");
        auto m = std::make_shared<Material>();
        m->desc.kind = AGX_MAT_INVISIBLE_SURFACE;
        return m;
    });
    material.add_creator("disney", [](const ConfigGroup &p, CreatingContext &ctx) {  // material_creators.cpp:46-112
        reject_normal_map(p);
        if (p.find_child("bssrdf_dmfp")) throw ObjectConstructionException("BSSRDF surfaces are not supported by gpu_pt");
        auto m = std::make_shared<Material>();
        m->desc.kind = AGX_MAT_DISNEY;
        copy3(m->desc.color, texture_value(ctx, p, "base_color"));
        m->desc.metallic = texture_value(ctx, p, "metallic").x;
        m->desc.roughness = texture_value(ctx, p, "roughness").x;
        m->desc.transmission = texture_value_or(ctx, p, "transmission", 0).x;
        m->desc.transmission_roughness = p.find_child_group("transmission_roughness")
                                             ? texture_value(ctx, p, "transmission_roughness").x
                                             : m->desc.roughness;
        m->desc.ior = texture_value_or(ctx, p, "ior", 1.5f).x;
        copy3(m->desc.color2, texture_value_or(ctx, p, "specular_scale", 1));
        m->desc.specular_tint = texture_value_or(ctx, p, "specular_tint", 0).x;
        m->desc.anisotropic = texture_value_or(ctx, p, "anisotropic", 0).x;
        m->desc.sheen = texture_value_or(ctx, p, "sheen", 0).x;
        m->desc.sheen_tint = texture_value_or(ctx, p, "sheen_tint", 0).x;
        m->desc.clearcoat = texture_value_or(ctx, p, "clearcoat", 0).x;
        m->desc.clearcoat_gloss = texture_value_or(ctx, p, "clearcoat_gloss", 1).x;
        return m;
    });
    for (const char *unsupported : {"dream_works_fabric", "paper", "phong"})
        material.add_creator(unsupported, [n = std::string(unsupported)](const ConfigGroup &, CreatingContext &)
                                              -> std::shared_ptr<Material> {
            throw ObjectConstructionException("material type '" + n + "' is not supported by gpu_pt yet");
        });

    // texture3d (texture3d_creators.cpp:18-207)
    texture3d.add_creator("constant", [](const ConfigGroup &p, CreatingContext &) {
        auto t = std::make_shared<Texture3D>();
        tex3d_full_transform(p);  // validated as Atrc does; constant3d ignores it
        t->desc.kind = AGX_TEX3D_CONSTANT;
        t->desc.width = t->desc.height = t->desc.depth = t->desc.channels = 1;
        copy3(t->desc.value, p.child_spectrum("texel"));
        Transform3().to_rows12(t->desc.transform);
        t->desc.inv_gamma = p.child_real_or("inv_gamma", 1);
        t->desc.linear = 1;
        return t;
    });
    texture3d.add_creator("image3d", [](const ConfigGroup &p, CreatingContext &ctx) {
        auto t = std::make_shared<Texture3D>();
        tex3d_full_transform(p).to_rows12(t->desc.transform);
        t->desc.wrap[0] = wrap_mode(p.child_str_or("wrap_u", "clamp"));
        t->desc.wrap[1] = wrap_mode(p.child_str_or("wrap_v", "clamp"));
        t->desc.wrap[2] = wrap_mode(p.child_str_or("wrap_w", "clamp"));
        t->desc.inv_gamma = p.child_real_or("inv_gamma", 1);
        t->desc.linear = p.child_str_or("sampler", "linear") == "linear" ? 1 : 0;
        const std::string format = p.child_str("format");
        int w, h, d, c;
        if (p.find_child("ascii_filename"))
            t->grid = load_texels(ctx.path_mapper->map(p.child_str("ascii_filename")), false, format, &w, &h, &d, &c);
        else if (p.find_child("binary_filename"))
            t->grid = load_texels(ctx.path_mapper->map(p.child_str("binary_filename")), true, format, &w, &h, &d, &c);
        else if (p.find_child("image_filenames"))
            throw ObjectConstructionException("image3d from 2d image slices is not supported by gpu_pt");
        else
            throw ObjectConstructionException("input filename is unspecified");
        t->desc.kind = AGX_TEX3D_GRID;
        t->desc.width = w; t->desc.height = h; t->desc.depth = d; t->desc.channels = c;
        t->desc.data = t->grid->data();
        return t;
    });

    // medium (medium_creators.cpp:9-88)
    medium.add_creator("void", [](const ConfigGroup &, CreatingContext &) {
        auto m = std::make_shared<Medium>();
        m->type = "void";
        m->desc.kind = AGX_MEDIUM_VOID;
        return m;
    });
    medium.add_creator("homogeneous", [](const ConfigGroup &p, CreatingContext &) {
        auto m = std::make_shared<Medium>();
        m->type = "homogeneous";
        m->desc.kind = AGX_MEDIUM_HOMOGENEOUS;
        copy3(m->desc.sigma_a, p.child_spectrum("sigma_a"));
        copy3(m->desc.sigma_s, p.child_spectrum("sigma_s"));
        m->desc.g = p.child_real("g");
        if (m->desc.g <= -1 || m->desc.g >= 1)  // homogeneous.cpp:33-35
            throw ObjectConstructionException("invalid g value: " + std::to_string(m->desc.g));
        m->desc.max_scattering_count = p.child_int_or("max_scattering_count", std::numeric_limits<int>::max());
        return m;
    });
    medium.add_creator("heterogeneous", [](const ConfigGroup &p, CreatingContext &ctx) {
        auto m = std::make_shared<Medium>();
        m->type = "heterogeneous";
        m->desc.kind = AGX_MEDIUM_HETEROGENEOUS;
        const Transform3 l2w = p.child_transform3("transform");
        l2w.to_rows12(m->desc.local_to_world);
        l2w.inverse().to_rows12(m->desc.world_to_local);
        m->density = ctx.create<Texture3D>(p.child_group("density"));
        m->albedo = ctx.create<Texture3D>(p.child_group("albedo"));
        m->g = ctx.create<Texture3D>(p.child_group("g"));
        m->desc.density = m->density->desc;
        m->desc.albedo = m->albedo->desc;
        m->desc.g3 = m->g->desc;
        m->desc.max_scattering_count = p.child_int_or("max_scattering_count", std::numeric_limits<int>::max());
        m->desc.white_for_indirect = p.child_int_or("white_for_indirect", 0) != 0;
        return m;
    });

    // entity (entity_creators.cpp:29-60)
    entity.add_creator("geometric", [](const ConfigGroup &p, CreatingContext &ctx) {
        auto e = std::make_shared<Entity>();
        auto g = ctx.create<Geometry>(p.child_group("geometry"));
        e->material = ctx.create<Material>(p.child_group("material"));
        // create_medium_interface (entity_creators.cpp:11-27): missing = void
        if (auto n = p.find_child_group("med_in")) e->med_in = ctx.create<Medium>(*n);
        if (auto n = p.find_child_group("med_out")) e->med_out = ctx.create<Medium>(*n);
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
    // IBLCreator (envir_light_creators.cpp:11-31)
    envir_light.add_creator("ibl", [](const ConfigGroup &p, CreatingContext &ctx) {
        auto l = std::make_shared<EnvirLight>();
        l->desc.kind = AGX_ENV_IBL;
        l->tex = ctx.create<Texture2D>(p.child_group("tex"));
        l->desc.tex = l->tex->desc;
        l->desc.no_importance_sampling = p.child_int_or("no_importance_sampling", 0) != 0 ? 1 : 0;
        l->desc.power = p.child_real_or("power", -1);
        return l;
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
        // GaussianFilter::GaussianFilter (gaussian_filter.cpp:45-56)
        if (f->radius <= 0) throw ObjectConstructionException("invalid radius value: " + std::to_string(f->radius));
        if (f->alpha <= 0) throw ObjectConstructionException("invalid alpha value: " + std::to_string(f->alpha));
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
template std::shared_ptr<Texture3D> CreatingContext::create<Texture3D>(const ConfigGroup &);
template std::shared_ptr<Geometry> CreatingContext::create<Geometry>(const ConfigGroup &);
template std::shared_ptr<Material> CreatingContext::create<Material>(const ConfigGroup &);
template std::shared_ptr<Medium> CreatingContext::create<Medium>(const ConfigGroup &);
template std::shared_ptr<Entity> CreatingContext::create<Entity>(const ConfigGroup &);
template std::shared_ptr<EnvirLight> CreatingContext::create<EnvirLight>(const ConfigGroup &);
template std::shared_ptr<Aggregate> CreatingContext::create<Aggregate>(const ConfigGroup &);

}  // namespace atrc
