// Minimal JSON value and parser for the renderer configuration.
//
// Covers what the reference's configs use through nlohmann::json
// (src/main/main.cu:5-17 parses with comments allowed): objects, arrays,
// strings, numbers, booleans, null, and // or /* */ comments.  The accessor
// names follow nlohmann's (at, value, find/contains, operator[], is_*), so the
// pass code reads like the reference's from_json functions.
#pragma once
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace srt {

class json_error : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

class Json {
public:
    enum class Type { Null, Bool, Number, String, Array, Object };

    Json() = default;
    Json(bool b) : type_(Type::Bool), b_(b) {}
    Json(double d) : type_(Type::Number), d_(d) {}
    Json(int i) : type_(Type::Number), d_(i) {}
    Json(const char* s) : type_(Type::String), s_(s) {}
    Json(std::string s) : type_(Type::String), s_(std::move(s)) {}

    static Json array() { Json j; j.type_ = Type::Array; return j; }
    static Json object() { Json j; j.type_ = Type::Object; return j; }
    static Json parse(const std::string& text);        // throws json_error
    static Json load(const std::string& path);         // throws json_error

    Type type() const { return type_; }
    bool is_null() const { return type_ == Type::Null; }
    bool is_bool() const { return type_ == Type::Bool; }
    bool is_number() const { return type_ == Type::Number; }
    bool is_string() const { return type_ == Type::String; }
    bool is_array() const { return type_ == Type::Array; }
    bool is_object() const { return type_ == Type::Object; }

    // object access
    bool contains(const std::string& k) const { return is_object() && obj_.count(k) > 0; }
    const Json& at(const std::string& k) const;        // throws json_error when missing
    const Json& operator[](const std::string& k) const; // null when missing
    Json& operator[](const std::string& k);             // inserts
    // array access
    size_t size() const { return is_array() ? arr_.size() : (is_object() ? obj_.size() : 0); }
    const Json& at(size_t i) const;
    const Json& operator[](size_t i) const { return at(i); }
    void push_back(Json v) { if (type_ == Type::Null) type_ = Type::Array; arr_.push_back(std::move(v)); }
    const std::vector<Json>& items() const { return arr_; }
    const std::map<std::string, Json>& members() const { return obj_; }

    // conversions (throw json_error on type mismatch)
    bool get_bool() const;
    double get_number() const;
    int get_int() const { return (int)get_number(); }
    float get_float() const { return (float)get_number(); }
    const std::string& get_string() const;

    // value(key, default): default when missing or null (nlohmann semantics)
    int value(const std::string& k, int def) const;
    float value(const std::string& k, float def) const;
    double value(const std::string& k, double def) const;
    bool value(const std::string& k, bool def) const;
    std::string value(const std::string& k, const char* def) const;
    std::string value(const std::string& k, const std::string& def) const;

    std::string dump() const;

private:
    Type type_ = Type::Null;
    bool b_ = false;
    double d_ = 0.0;
    std::string s_;
    std::vector<Json> arr_;
    std::map<std::string, Json> obj_;
    friend class JsonParser;
};

}  // namespace srt
