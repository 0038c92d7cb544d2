// Small JSON value type for the session protocol, model files and results.
// Objects keep insertion order; numbers are FP64 (integral literals also keep
// an exact int64); non-finite numbers serialise as null.
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace nyk {

class Json {
  public:
    enum class Type { Null, Bool, Number, String, Array, Object };

    Json() = default;
    Json(std::nullptr_t) {}
    Json(bool b) : type_(Type::Bool), bool_(b) {}
    Json(double d) : type_(Type::Number), num_(d) {}
    Json(int v) : Json((long long)v) {}
    Json(long v) : Json((long long)v) {}
    Json(long long v) : type_(Type::Number), num_(double(v)), int_(int64_t(v)), integral_(true) {}
    Json(unsigned long v) : Json((long long)v) {}
    Json(unsigned long long v) : Json((long long)v) {}
    Json(const char *s) : type_(Type::String), str_(s) {}
    Json(std::string s) : type_(Type::String), str_(std::move(s)) {}
    Json(const std::vector<double> &v);
    Json(const std::vector<std::string> &v);

    static Json array() {
        Json j;
        j.type_ = Type::Array;
        return j;
    }
    static Json object() {
        Json j;
        j.type_ = Type::Object;
        return j;
    }
    // Parses one JSON document; nesting deeper than max_depth is rejected.
    static Json parse(const std::string &text, int max_depth = 64);

    Type type() const { return type_; }
    bool is_null() const { return type_ == Type::Null; }
    bool is_bool() const { return type_ == Type::Bool; }
    bool is_number() const { return type_ == Type::Number; }
    bool is_string() const { return type_ == Type::String; }
    bool is_array() const { return type_ == Type::Array; }
    bool is_object() const { return type_ == Type::Object; }
    bool is_integral() const { return type_ == Type::Number && integral_; }

    bool boolean() const;
    double number() const;
    int64_t integer() const;
    const std::string &str() const;

    size_t size() const;
    const Json &operator[](size_t i) const;
    void push_back(Json v);

    Json &operator[](const std::string &key);  // inserts null when missing
    const Json &at(const std::string &key) const;  // throws when missing
    const Json *find(const std::string &key) const;
    bool contains(const std::string &key) const { return find(key) != nullptr; }
    const std::vector<std::pair<std::string, Json>> &items() const;

    std::vector<double> numbers() const;  // array of numbers

    std::string dump(int indent = -1) const;

  private:
    void write(std::string &out, int indent, int level) const;
    Type type_ = Type::Null;
    bool bool_ = false;
    double num_ = 0;
    int64_t int_ = 0;
    bool integral_ = false;
    std::string str_;
    std::vector<Json> arr_;
    std::vector<std::pair<std::string, Json>> obj_;
    friend class JsonParser;
};

}  // namespace nyk
