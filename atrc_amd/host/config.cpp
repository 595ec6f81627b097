// Configuration tree and a small JSON reader.
// JSON conversion follows src/factory/src/utility/config_cvt.cpp:39-105;
// typed getters and transforms follow src/tracer/src/utility/config.cpp:11-314.
#include "config.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <sstream>

namespace atrc {

bool ends_with(const std::string &s, const std::string &suffix) {
    return s.size() >= suffix.size() && s.compare(s.size() - suffix.size(), suffix.size(), suffix) == 0;
}

// ------------------------------------------------------------------ Transform3
Transform3::Transform3() {
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) m[i][j] = i == j ? 1 : 0;
}
Transform3 Transform3::translate(double x, double y, double z) {
    Transform3 t;
    t.m[0][3] = x;
    t.m[1][3] = y;
    t.m[2][3] = z;
    return t;
}
Transform3 Transform3::scale(double x, double y, double z) {
    Transform3 t;
    t.m[0][0] = x;
    t.m[1][1] = y;
    t.m[2][2] = z;
    return t;
}
// rotation by `rad` about the axis (Rodrigues)
Transform3 Transform3::rotate(double ax, double ay, double az, double rad) {
    const double l = std::sqrt(ax * ax + ay * ay + az * az);
    if (!(l > 0)) throw ConfigException("rotate: zero axis");
    ax /= l;
    ay /= l;
    az /= l;
    const double c = std::cos(rad), s = std::sin(rad), t = 1 - c;
    Transform3 r;
    r.m[0][0] = t * ax * ax + c;      r.m[0][1] = t * ax * ay - s * az; r.m[0][2] = t * ax * az + s * ay;
    r.m[1][0] = t * ax * ay + s * az; r.m[1][1] = t * ay * ay + c;      r.m[1][2] = t * ay * az - s * ax;
    r.m[2][0] = t * ax * az - s * ay; r.m[2][1] = t * ay * az + s * ax; r.m[2][2] = t * az * az + c;
    return r;
}
Transform3 Transform3::operator*(const Transform3 &r) const {
    Transform3 o;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double acc = 0;
            for (int k = 0; k < 4; ++k) acc += m[i][k] * r.m[k][j];
            o.m[i][j] = acc;
        }
    return o;
}
Transform3 Transform3::inverse() const {
    // affine inverse: [A t]^-1 = [A^-1  -A^-1 t]
    const double(&a)[4][4] = m;
    const double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
                       a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                       a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    if (!(std::fabs(det) > 0)) throw ConfigException("singular transform");
    Transform3 o;
    const double id = 1 / det;
    o.m[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) * id;
    o.m[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) * id;
    o.m[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) * id;
    o.m[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) * id;
    o.m[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) * id;
    o.m[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) * id;
    o.m[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) * id;
    o.m[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) * id;
    o.m[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) * id;
    for (int i = 0; i < 3; ++i)
        o.m[i][3] = -(o.m[i][0] * a[0][3] + o.m[i][1] * a[1][3] + o.m[i][2] * a[2][3]);
    return o;
}
Vec3 Transform3::apply_to_point(const Vec3 &p) const {
    return {float(m[0][0] * p.x + m[0][1] * p.y + m[0][2] * p.z + m[0][3]),
            float(m[1][0] * p.x + m[1][1] * p.y + m[1][2] * p.z + m[1][3]),
            float(m[2][0] * p.x + m[2][1] * p.y + m[2][2] * p.z + m[2][3])};
}
Vec3 Transform3::apply_to_vector(const Vec3 &v) const {
    return {float(m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z), float(m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z),
            float(m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z)};
}
void Transform3::to_rows12(float out[12]) const {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 4; ++j) out[4 * i + j] = float(m[i][j]);
}

// ------------------------------------------------------------------ nodes
const ConfigGroup &ConfigNode::as_group() const {
    if (!is_group()) throw ConfigException("config node is not a group");
    return static_cast<const ConfigGroup &>(*this);
}
const ConfigArray &ConfigNode::as_array() const {
    if (!is_array()) throw ConfigException("config node is not an array");
    return static_cast<const ConfigArray &>(*this);
}
const ConfigValue &ConfigNode::as_value() const {
    if (!is_value()) throw ConfigException("config node is not a value");
    return static_cast<const ConfigValue &>(*this);
}

int ConfigValue::as_int() const {
    try {
        size_t pos = 0;
        const int v = std::stoi(str_, &pos);
        return v;
    } catch (...) {
        throw ConfigException("invalid int value: " + str_);
    }
}
float ConfigValue::as_real() const {
    try {
        return std::stof(str_);
    } catch (...) {
        throw ConfigException("invalid real value: " + str_);
    }
}
std::string ConfigValue::to_string(int) const { return "\"" + str_ + "\""; }

const ConfigNode &ConfigArray::at(size_t i) const {
    if (i >= elems_.size()) throw ConfigException("config array index out of range");
    return *elems_[i];
}
Transform3 ConfigArray::at_transform3(size_t i) const { return to_transform3(at(i)); }
std::string ConfigArray::to_string(int indent) const {
    std::string s = "[";
    for (size_t i = 0; i < elems_.size(); ++i) s += (i ? ", " : "") + elems_[i]->to_string(indent + 1);
    return s + "]";
}

void ConfigGroup::insert_child(const std::string &name, std::shared_ptr<ConfigNode> n) {
    for (auto &c : children_)
        if (c.first == name) {
            c.second = std::move(n);
            return;
        }
    children_.emplace_back(name, std::move(n));
}
const ConfigNode *ConfigGroup::find_child(const std::string &name) const {
    for (auto &c : children_)
        if (c.first == name) return c.second.get();
    return nullptr;
}
const ConfigGroup *ConfigGroup::find_child_group(const std::string &name) const {
    auto n = find_child(name);
    return n && n->is_group() ? &n->as_group() : nullptr;
}
const ConfigArray *ConfigGroup::find_child_array(const std::string &name) const {
    auto n = find_child(name);
    return n && n->is_array() ? &n->as_array() : nullptr;
}
const ConfigValue *ConfigGroup::find_child_value(const std::string &name) const {
    auto n = find_child(name);
    return n && n->is_value() ? &n->as_value() : nullptr;
}
const ConfigNode &ConfigGroup::child(const std::string &name) const {
    auto n = find_child(name);
    if (!n) throw ConfigException("config child not found: " + name);
    return *n;
}
const ConfigGroup &ConfigGroup::child_group(const std::string &name) const { return child(name).as_group(); }
const ConfigArray &ConfigGroup::child_array(const std::string &name) const { return child(name).as_array(); }
const std::string &ConfigGroup::child_str(const std::string &name) const { return child(name).as_value().as_str(); }
std::string ConfigGroup::child_str_or(const std::string &name, const std::string &def) const {
    auto n = find_child_value(name);
    return n ? n->as_str() : def;
}
int ConfigGroup::child_int(const std::string &name) const { return child(name).as_value().as_int(); }
int ConfigGroup::child_int_or(const std::string &name, int def) const {
    auto n = find_child_value(name);
    return n ? n->as_int() : def;
}
float ConfigGroup::child_real(const std::string &name) const { return child(name).as_value().as_real(); }
float ConfigGroup::child_real_or(const std::string &name, float def) const {
    auto n = find_child_value(name);
    return n ? n->as_real() : def;
}
Vec2 ConfigGroup::child_vec2(const std::string &name) const {
    const auto &a = child_array(name);  // config.cpp:11-34
    if (a.size() == 1) return {a.at_real(0), a.at_real(0)};
    if (a.size() == 2) return {a.at_real(0), a.at_real(1)};
    throw ConfigException("invalid array size for vec2: " + name);
}
Vec2 ConfigGroup::child_vec2_or(const std::string &name, Vec2 def) const {
    return find_child(name) ? child_vec2(name) : def;
}
Vec3 ConfigGroup::child_vec3(const std::string &name) const { return to_vec3(child(name)); }
Vec3 ConfigGroup::child_vec3_or(const std::string &name, Vec3 def) const {
    return find_child(name) ? child_vec3(name) : def;
}
Transform3 ConfigGroup::child_transform3(const std::string &name) const {
    auto n = find_child(name);
    return n ? to_transform3(*n) : Transform3();
}
std::string ConfigGroup::to_string(int indent) const {
    std::string s = "{";
    for (size_t i = 0; i < children_.size(); ++i)
        s += (i ? ", " : "") + std::string("\"") + children_[i].first + "\": " + children_[i].second->to_string(indent + 1);
    return s + "}";
}

Vec3 to_vec3(const ConfigNode &node) {  // config.cpp:36-59
    const auto &a = node.as_array();
    if (a.size() == 1) return {a.at_real(0), a.at_real(0), a.at_real(0)};
    if (a.size() == 3) return {a.at_real(0), a.at_real(1), a.at_real(2)};
    throw ConfigException("invalid array size for vec3 (" + std::to_string(a.size()) + ")");
}

static double angle_of(const ConfigGroup &g) {
    if (auto p = g.find_child("rad")) return p->as_value().as_real();
    return g.child_real("deg") / 180 * 3.14159265358979323846f;
}

// to_basic_transform3 / to_transform3 (config.cpp:195-314): an array is the
// product of its elements, so the last one is applied first.
Transform3 to_transform3(const ConfigNode &node) {
    if (node.is_array()) {
        const auto &arr = node.as_array();
        Transform3 r;
        for (size_t i = 0; i < arr.size(); ++i) r = r * arr.at_transform3(i);
        return r;
    }
    const auto &g = node.as_group();
    const std::string &type = g.child_str("type");
    if (ends_with(type, "//")) return Transform3();
    if (type == "translate") {
        const Vec3 o = g.child_vec3("offset");
        return Transform3::translate(o.x, o.y, o.z);
    }
    if (type == "rotate") {
        const Vec3 a = g.child_vec3("axis");
        return Transform3::rotate(a.x, a.y, a.z, angle_of(g));
    }
    if (type == "rotate_x") return Transform3::rotate(1, 0, 0, angle_of(g));
    if (type == "rotate_y") return Transform3::rotate(0, 1, 0, angle_of(g));
    if (type == "rotate_z") return Transform3::rotate(0, 0, 1, angle_of(g));
    if (type == "scale") {
        const ConfigNode &r = g.child("ratio");
        if (r.is_array()) {
            const Vec3 s = g.child_vec3("ratio");
            return Transform3::scale(s.x, s.y, s.z);
        }
        const float s = g.child_real("ratio");
        return Transform3::scale(s, s, s);
    }
    if (type == "mat3") {
        Transform3 t;
        if (auto arr = g.find_child_array("row_major")) {
            if (arr->size() != 9) throw ConfigException("invalid mat3 elem count");
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) t.m[i][j] = arr->at_real(size_t(3 * i + j));
            return t;
        }
        if (auto arr = g.find_child_array("col_major")) {
            if (arr->size() != 9) throw ConfigException("invalid mat3 elem count");
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) t.m[j][i] = arr->at_real(size_t(3 * i + j));
            return t;
        }
        throw ConfigException("row_major/col_major expected");
    }
    throw ConfigException("unknown transform3 type (type = " + type + ")");
}

// ------------------------------------------------------------------ JSON reader
namespace {

struct Json {
    enum Kind { Null, Bool, Int, Float, String, Array, Object } kind = Null;
    bool b = false;
    long long i = 0;
    double f = 0;
    std::string s;
    std::vector<Json> arr;
    std::vector<std::pair<std::string, Json>> obj;
};

struct Parser {
    const std::string &t;
    size_t p = 0;
    explicit Parser(const std::string &text) : t(text) {}
    [[noreturn]] void fail(const std::string &m) const {
        throw ConfigException("JSON parse error at offset " + std::to_string(p) + ": " + m);
    }
    void ws() {
        for (;;) {
            while (p < t.size() && (t[p] == ' ' || t[p] == '\n' || t[p] == '\r' || t[p] == '\t')) ++p;
            // tolerate // line comments (Atrc scene files sometimes carry them)
            if (p + 1 < t.size() && t[p] == '/' && t[p + 1] == '/') {
                while (p < t.size() && t[p] != '\n') ++p;
                continue;
            }
            break;
        }
    }
    Json value() {
        ws();
        if (p >= t.size()) fail("unexpected end");
        const char c = t[p];
        if (c == '{') return object();
        if (c == '[') return array();
        if (c == '"') { Json j; j.kind = Json::String; j.s = str(); return j; }
        if (t.compare(p, 4, "true") == 0) { p += 4; Json j; j.kind = Json::Bool; j.b = true; return j; }
        if (t.compare(p, 5, "false") == 0) { p += 5; Json j; j.kind = Json::Bool; j.b = false; return j; }
        if (t.compare(p, 4, "null") == 0) { p += 4; return Json(); }
        return number();
    }
    std::string str() {
        if (t[p] != '"') fail("string expected");
        ++p;
        std::string out;
        while (p < t.size() && t[p] != '"') {
            char c = t[p++];
            if (c == '\\') {
                if (p >= t.size()) fail("bad escape");
                const char e = t[p++];
                switch (e) {
                case 'n': out += '\n'; break;
                case 't': out += '\t'; break;
                case 'r': out += '\r'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'u': {
                    if (p + 4 > t.size()) fail("bad \\u escape");
                    const unsigned cp = unsigned(std::stoul(t.substr(p, 4), nullptr, 16));
                    p += 4;
                    if (cp < 0x80) out += char(cp);
                    else if (cp < 0x800) { out += char(0xC0 | (cp >> 6)); out += char(0x80 | (cp & 0x3F)); }
                    else { out += char(0xE0 | (cp >> 12)); out += char(0x80 | ((cp >> 6) & 0x3F)); out += char(0x80 | (cp & 0x3F)); }
                    break;
                }
                default: out += e; break;
                }
            } else {
                out += c;
            }
        }
        if (p >= t.size()) fail("unterminated string");
        ++p;
        return out;
    }
    Json number() {
        const size_t b = p;
        bool is_float = false;
        if (t[p] == '-' || t[p] == '+') ++p;
        while (p < t.size() && (isdigit(static_cast<unsigned char>(t[p])) || t[p] == '.' || t[p] == 'e' ||
                                t[p] == 'E' || t[p] == '-' || t[p] == '+')) {
            if (t[p] == '.' || t[p] == 'e' || t[p] == 'E') is_float = true;
            ++p;
        }
        if (p == b) fail("value expected");
        const std::string s = t.substr(b, p - b);
        Json j;
        try {
            if (is_float) { j.kind = Json::Float; j.f = std::stod(s); }
            else { j.kind = Json::Int; j.i = std::stoll(s); }
        } catch (...) {
            fail("bad number " + s);
        }
        return j;
    }
    Json array() {
        Json j;
        j.kind = Json::Array;
        ++p;
        ws();
        if (p < t.size() && t[p] == ']') { ++p; return j; }
        for (;;) {
            j.arr.push_back(value());
            ws();
            if (p < t.size() && t[p] == ',') { ++p; continue; }
            if (p < t.size() && t[p] == ']') { ++p; return j; }
            fail("',' or ']' expected");
        }
    }
    Json object() {
        Json j;
        j.kind = Json::Object;
        ++p;
        ws();
        if (p < t.size() && t[p] == '}') { ++p; return j; }
        for (;;) {
            ws();
            std::string k = str();
            ws();
            if (p >= t.size() || t[p] != ':') fail("':' expected");
            ++p;
            Json v = value();
            j.obj.emplace_back(std::move(k), std::move(v));
            ws();
            if (p < t.size() && t[p] == ',') { ++p; continue; }
            if (p < t.size() && t[p] == '}') { ++p; return j; }
            fail("',' or '}' expected");
        }
    }
};

// config_cvt.cpp:40-49: bools -> "1"/"0", floats -> std::to_string(float),
// integers -> std::to_string(int), strings verbatim.
std::shared_ptr<ConfigNode> to_node(const Json &j) {
    switch (j.kind) {
    case Json::Object: {
        auto g = std::make_shared<ConfigGroup>();
        for (auto &kv : j.obj) g->insert_child(kv.first, to_node(kv.second));
        return g;
    }
    case Json::Array: {
        auto a = std::make_shared<ConfigArray>();
        for (auto &e : j.arr) a->push_back(to_node(e));
        return a;
    }
    case Json::Bool: return std::make_shared<ConfigValue>(j.b ? "1" : "0");
    case Json::Float: return std::make_shared<ConfigValue>(std::to_string(float(j.f)));
    case Json::Int: return std::make_shared<ConfigValue>(std::to_string(int(j.i)));
    case Json::String: return std::make_shared<ConfigValue>(j.s);
    default: throw ConfigException("null is not a valid config value");
    }
}

}  // namespace

std::shared_ptr<ConfigGroup> json_to_config(const std::string &text) {
    Parser ps(text);
    Json j = ps.value();
    ps.ws();
    if (ps.p != text.size()) ps.fail("trailing characters");
    if (j.kind != Json::Object) throw ConfigException("top-level JSON value must be an object");
    return std::static_pointer_cast<ConfigGroup>(to_node(j));
}

}  // namespace atrc
