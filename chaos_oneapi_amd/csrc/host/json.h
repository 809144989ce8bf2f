// Minimal JSON value with a reader and a writer, for config.json input and the
// Data/<Tool> <model>.json output.  It covers what the reference uses of
// nlohmann/json (include/json/json.hpp): object lookup with at() that throws
// on a missing key, get_to() into scalars and vectors, sorted object keys and
// indented output where doubles print as the shortest round-trip form.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace chaos {

class Json {
public:
    enum class Type { Null, Bool, Int, Double, String, Array, Object };

    Json() = default;
    Json(std::nullptr_t) {}
    Json(bool b) : type_(Type::Bool), b_(b) {}
    Json(int v) : type_(Type::Int), i_(v) {}
    Json(int64_t v) : type_(Type::Int), i_(v) {}
    Json(double v) : type_(Type::Double), d_(v) {}
    Json(const char* s) : type_(Type::String), s_(s) {}
    Json(const std::string& s) : type_(Type::String), s_(s) {}
    Json(const std::vector<double>& v);
    Json(const std::vector<int>& v);

    static Json array();
    static Json object();
    static Json parse(const std::string& text);

    Type type() const { return type_; }
    bool is_object() const { return type_ == Type::Object; }
    bool is_array() const { return type_ == Type::Array; }
    bool is_number() const { return type_ == Type::Int || type_ == Type::Double; }
    bool is_string() const { return type_ == Type::String; }
    bool is_bool() const { return type_ == Type::Bool; }
    bool is_null() const { return type_ == Type::Null; }

    // object access; at() throws std::out_of_range like nlohmann's
    const Json& at(const std::string& key) const;
    bool contains(const std::string& key) const;
    Json& operator[](const std::string& key);  // creates (turns null into an object)
    const std::map<std::string, Json>& items() const;

    // array access
    size_t size() const;
    const Json& at(size_t i) const;
    void push_back(const Json& v);

    // conversions (throw std::invalid_argument on a type mismatch)
    double as_double() const;
    int64_t as_int() const;
    bool as_bool() const;
    const std::string& as_string() const;
    std::vector<double> as_double_vector() const;

    void get_to(double& v) const { v = as_double(); }
    void get_to(int32_t& v) const { v = (int32_t)as_int(); }
    void get_to(int64_t& v) const { v = as_int(); }
    void get_to(bool& v) const { v = as_bool(); }
    void get_to(std::string& v) const { v = as_string(); }
    void get_to(std::vector<double>& v) const { v = as_double_vector(); }

    bool operator==(const std::string& s) const { return type_ == Type::String && s_ == s; }
    bool operator!=(const std::string& s) const { return !(*this == s); }

    // indent < 0: compact; else pretty with `indent` spaces per level
    std::string dump(int indent = -1) const;

private:
    void dump_to(std::string& out, int indent, int level) const;

    Type type_ = Type::Null;
    bool b_ = false;
    int64_t i_ = 0;
    double d_ = 0.0;
    std::string s_;
    std::vector<Json> a_;
    std::map<std::string, Json> o_;
};

// shortest decimal that reads back to the same double, printf %g style
// exponents, integral values with a trailing ".0"
std::string format_double(double v);

}  // namespace chaos
