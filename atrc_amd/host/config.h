// Atrc's configuration tree: JSON text -> ConfigGroup / ConfigArray / ConfigValue
// of strings with typed getters.
//
// Mirrors src/tracer/include/agz/tracer/utility/config.h:11-204 and the JSON
// conversion of src/factory/src/utility/config_cvt.cpp:39-105, including its
// quirk: a JSON float is stored as std::to_string(float) (6 decimals), so values
// below 5e-7 in magnitude read back as 0.
#pragma once

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace atrc {

class ConfigException : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

// Atrc's ObjectConstructionException (common.h:130-135)
class ObjectConstructionException : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

struct Vec2 { float x = 0, y = 0; };
struct Vec3 { float x = 0, y = 0, z = 0; };

// Row-major 4x4 affine transform (last row 0 0 0 1).
struct Transform3 {
    double m[4][4];
    Transform3();
    static Transform3 translate(double x, double y, double z);
    static Transform3 rotate(double ax, double ay, double az, double rad);
    static Transform3 scale(double x, double y, double z);
    Transform3 operator*(const Transform3 &r) const;
    Transform3 inverse() const;
    Vec3 apply_to_point(const Vec3 &p) const;
    Vec3 apply_to_vector(const Vec3 &v) const;
    void to_rows12(float out[12]) const;
};

class ConfigGroup;
class ConfigArray;
class ConfigValue;

class ConfigNode {
public:
    virtual ~ConfigNode() = default;
    virtual bool is_group() const { return false; }
    virtual bool is_array() const { return false; }
    virtual bool is_value() const { return false; }
    const ConfigGroup &as_group() const;
    const ConfigArray &as_array() const;
    const ConfigValue &as_value() const;
    virtual std::string to_string(int indent = 0) const = 0;
};

class ConfigValue : public ConfigNode {
    std::string str_;
public:
    explicit ConfigValue(std::string s) : str_(std::move(s)) {}
    bool is_value() const override { return true; }
    const std::string &as_str() const { return str_; }
    int as_int() const;
    float as_real() const;
    std::string to_string(int indent = 0) const override;
};

class ConfigArray : public ConfigNode {
    std::vector<std::shared_ptr<ConfigNode>> elems_;
public:
    bool is_array() const override { return true; }
    size_t size() const { return elems_.size(); }
    const ConfigNode &at(size_t i) const;
    const ConfigGroup &at_group(size_t i) const { return at(i).as_group(); }
    float at_real(size_t i) const { return at(i).as_value().as_real(); }
    int at_int(size_t i) const { return at(i).as_value().as_int(); }
    const std::string &at_str(size_t i) const { return at(i).as_value().as_str(); }
    Transform3 at_transform3(size_t i) const;
    void push_back(std::shared_ptr<ConfigNode> n) { elems_.push_back(std::move(n)); }
    std::string to_string(int indent = 0) const override;
};

class ConfigGroup : public ConfigNode {
    std::vector<std::pair<std::string, std::shared_ptr<ConfigNode>>> children_;
public:
    bool is_group() const override { return true; }
    void insert_child(const std::string &name, std::shared_ptr<ConfigNode> n);
    const ConfigNode *find_child(const std::string &name) const;
    const ConfigGroup *find_child_group(const std::string &name) const;
    const ConfigArray *find_child_array(const std::string &name) const;
    const ConfigValue *find_child_value(const std::string &name) const;
    const ConfigNode &child(const std::string &name) const;
    const ConfigGroup &child_group(const std::string &name) const;
    const ConfigArray &child_array(const std::string &name) const;
    const std::string &child_str(const std::string &name) const;
    std::string child_str_or(const std::string &name, const std::string &def) const;
    int child_int(const std::string &name) const;
    int child_int_or(const std::string &name, int def) const;
    float child_real(const std::string &name) const;
    float child_real_or(const std::string &name, float def) const;
    Vec2 child_vec2(const std::string &name) const;
    Vec2 child_vec2_or(const std::string &name, Vec2 def) const;
    Vec3 child_vec3(const std::string &name) const;
    Vec3 child_vec3_or(const std::string &name, Vec3 def) const;
    Vec3 child_spectrum(const std::string &name) const { return child_vec3(name); }
    Vec3 child_spectrum_or(const std::string &name, Vec3 def) const { return child_vec3_or(name, def); }
    Transform3 child_transform3(const std::string &name) const;  // missing -> identity
    const std::vector<std::pair<std::string, std::shared_ptr<ConfigNode>>> &children() const { return children_; }
    std::string to_string(int indent = 0) const override;
};

// JSON text -> configuration tree (top level must be an object).
std::shared_ptr<ConfigGroup> json_to_config(const std::string &json_text);

Vec3 to_vec3(const ConfigNode &node);
Transform3 to_transform3(const ConfigNode &node);

bool ends_with(const std::string &s, const std::string &suffix);

}  // namespace atrc
