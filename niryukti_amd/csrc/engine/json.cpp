#include "json.hpp"

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>

namespace nyk {

Json::Json(const std::vector<double> &v) : type_(Type::Array) {
    arr_.reserve(v.size());
    for (double d : v) arr_.emplace_back(d);
}
Json::Json(const std::vector<std::string> &v) : type_(Type::Array) {
    arr_.reserve(v.size());
    for (const auto &s : v) arr_.emplace_back(s);
}

bool Json::boolean() const {
    if (type_ != Type::Bool) throw std::runtime_error("JSON value is not a boolean");
    return bool_;
}
double Json::number() const {
    if (type_ != Type::Number) throw std::runtime_error("JSON value is not a number");
    return num_;
}
int64_t Json::integer() const {
    if (type_ != Type::Number) throw std::runtime_error("JSON value is not a number");
    if (integral_) return int_;
    if (std::floor(num_) != num_ || !std::isfinite(num_)) throw std::runtime_error("JSON number is not an integer");
    return int64_t(num_);
}
const std::string &Json::str() const {
    if (type_ != Type::String) throw std::runtime_error("JSON value is not a string");
    return str_;
}
size_t Json::size() const {
    if (type_ == Type::Array) return arr_.size();
    if (type_ == Type::Object) return obj_.size();
    throw std::runtime_error("JSON value has no size");
}
const Json &Json::operator[](size_t i) const {
    if (type_ != Type::Array || i >= arr_.size()) throw std::runtime_error("JSON array index out of range");
    return arr_[i];
}
void Json::push_back(Json v) {
    if (type_ == Type::Null) type_ = Type::Array;
    if (type_ != Type::Array) throw std::runtime_error("JSON value is not an array");
    arr_.push_back(std::move(v));
}
Json &Json::operator[](const std::string &key) {
    if (type_ == Type::Null) type_ = Type::Object;
    if (type_ != Type::Object) throw std::runtime_error("JSON value is not an object");
    for (auto &kv : obj_)
        if (kv.first == key) return kv.second;
    obj_.emplace_back(key, Json());
    return obj_.back().second;
}
const Json *Json::find(const std::string &key) const {
    if (type_ != Type::Object) return nullptr;
    for (const auto &kv : obj_)
        if (kv.first == key) return &kv.second;
    return nullptr;
}
const Json &Json::at(const std::string &key) const {
    const Json *j = find(key);
    if (!j) throw std::runtime_error("Missing JSON key: " + key);
    return *j;
}
const std::vector<std::pair<std::string, Json>> &Json::items() const {
    if (type_ != Type::Object) throw std::runtime_error("JSON value is not an object");
    return obj_;
}
std::vector<double> Json::numbers() const {
    if (type_ != Type::Array) throw std::runtime_error("Expected a JSON array of numbers");
    std::vector<double> out;
    out.reserve(arr_.size());
    for (const auto &v : arr_) out.push_back(v.number());
    return out;
}

namespace {
void escape(std::string &out, const std::string &s) {
    out += '"';
    for (unsigned char ch : s) {
        switch (ch) {
        case '"': out += "\\\""; break;
        case '\\': out += "\\\\"; break;
        case '\n': out += "\\n"; break;
        case '\r': out += "\\r"; break;
        case '\t': out += "\\t"; break;
        case '\b': out += "\\b"; break;
        case '\f': out += "\\f"; break;
        default:
            if (ch < 0x20) {
                char buf[8];
                std::snprintf(buf, sizeof buf, "\\u%04x", ch);
                out += buf;
            } else
                out += char(ch);
        }
    }
    out += '"';
}
void newline(std::string &out, int indent, int level) {
    if (indent < 0) return;
    out += '\n';
    out.append(size_t(indent) * size_t(level), ' ');
}
}  // namespace

void Json::write(std::string &out, int indent, int level) const {
    switch (type_) {
    case Type::Null: out += "null"; break;
    case Type::Bool: out += bool_ ? "true" : "false"; break;
    case Type::Number: {
        if (!std::isfinite(num_)) {
            out += "null";
            break;
        }
        char buf[40];
        if (integral_)
            std::snprintf(buf, sizeof buf, "%lld", (long long)int_);
        else
            std::snprintf(buf, sizeof buf, "%.17g", num_);
        out += buf;
        break;
    }
    case Type::String: escape(out, str_); break;
    case Type::Array:
        out += '[';
        for (size_t i = 0; i < arr_.size(); ++i) {
            if (i) out += ',';
            newline(out, indent, level + 1);
            arr_[i].write(out, indent, level + 1);
        }
        if (!arr_.empty()) newline(out, indent, level);
        out += ']';
        break;
    case Type::Object:
        out += '{';
        for (size_t i = 0; i < obj_.size(); ++i) {
            if (i) out += ',';
            newline(out, indent, level + 1);
            escape(out, obj_[i].first);
            out += indent >= 0 ? ": " : ":";
            obj_[i].second.write(out, indent, level + 1);
        }
        if (!obj_.empty()) newline(out, indent, level);
        out += '}';
        break;
    }
}

std::string Json::dump(int indent) const {
    std::string out;
    write(out, indent, 0);
    return out;
}

class JsonParser {
  public:
    JsonParser(const std::string &t, int max_depth) : s_(t), max_depth_(max_depth) {}
    Json document() {
        Json v = value(0);
        skip();
        if (p_ != s_.size()) fail("trailing characters");
        return v;
    }

  private:
    [[noreturn]] void fail(const char *what) const {
        throw std::runtime_error(std::string("Invalid JSON (") + what + ") at offset " + std::to_string(p_));
    }
    void skip() {
        while (p_ < s_.size() && (s_[p_] == ' ' || s_[p_] == '\t' || s_[p_] == '\n' || s_[p_] == '\r')) ++p_;
    }
    bool consume(const char *word) {
        size_t n = std::char_traits<char>::length(word);
        if (s_.compare(p_, n, word) == 0) {
            p_ += n;
            return true;
        }
        return false;
    }
    Json value(int depth) {
        if (depth > max_depth_) fail("nesting too deep");
        skip();
        if (p_ >= s_.size()) fail("unexpected end");
        char c = s_[p_];
        if (c == '{') return object(depth);
        if (c == '[') return array(depth);
        if (c == '"') return Json(string());
        if (consume("true")) return Json(true);
        if (consume("false")) return Json(false);
        if (consume("null")) return Json();
        if (c == '-' || (c >= '0' && c <= '9')) return number();
        fail("unexpected character");
    }
    Json object(int depth) {
        Json j = Json::object();
        ++p_;
        skip();
        if (p_ < s_.size() && s_[p_] == '}') {
            ++p_;
            return j;
        }
        while (true) {
            skip();
            if (p_ >= s_.size() || s_[p_] != '"') fail("expected key");
            std::string key = string();
            skip();
            if (p_ >= s_.size() || s_[p_] != ':') fail("expected ':'");
            ++p_;
            if (j.find(key)) fail("duplicate key");
            j.obj_.emplace_back(std::move(key), value(depth + 1));
            skip();
            if (p_ < s_.size() && s_[p_] == ',') {
                ++p_;
                continue;
            }
            if (p_ < s_.size() && s_[p_] == '}') {
                ++p_;
                return j;
            }
            fail("expected ',' or '}'");
        }
    }
    Json array(int depth) {
        Json j = Json::array();
        ++p_;
        skip();
        if (p_ < s_.size() && s_[p_] == ']') {
            ++p_;
            return j;
        }
        while (true) {
            j.arr_.push_back(value(depth + 1));
            skip();
            if (p_ < s_.size() && s_[p_] == ',') {
                ++p_;
                continue;
            }
            if (p_ < s_.size() && s_[p_] == ']') {
                ++p_;
                return j;
            }
            fail("expected ',' or ']'");
        }
    }
    static void utf8(std::string &out, unsigned cp) {
        if (cp < 0x80)
            out += char(cp);
        else if (cp < 0x800) {
            out += char(0xC0 | (cp >> 6));
            out += char(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            out += char(0xE0 | (cp >> 12));
            out += char(0x80 | ((cp >> 6) & 0x3F));
            out += char(0x80 | (cp & 0x3F));
        } else {
            out += char(0xF0 | (cp >> 18));
            out += char(0x80 | ((cp >> 12) & 0x3F));
            out += char(0x80 | ((cp >> 6) & 0x3F));
            out += char(0x80 | (cp & 0x3F));
        }
    }
    unsigned hex4() {
        if (p_ + 4 > s_.size()) fail("short \\u escape");
        unsigned v = 0;
        for (int i = 0; i < 4; ++i) {
            char h = s_[p_++];
            v <<= 4;
            if (h >= '0' && h <= '9') v |= unsigned(h - '0');
            else if (h >= 'a' && h <= 'f') v |= unsigned(h - 'a' + 10);
            else if (h >= 'A' && h <= 'F') v |= unsigned(h - 'A' + 10);
            else fail("bad \\u escape");
        }
        return v;
    }
    std::string string() {
        ++p_;  // opening quote
        std::string out;
        while (true) {
            if (p_ >= s_.size()) fail("unterminated string");
            char c = s_[p_++];
            if (c == '"') return out;
            if (static_cast<unsigned char>(c) < 0x20) fail("control character in string");
            if (c != '\\') {
                out += c;
                continue;
            }
            if (p_ >= s_.size()) fail("unterminated escape");
            char e = s_[p_++];
            switch (e) {
            case '"': out += '"'; break;
            case '\\': out += '\\'; break;
            case '/': out += '/'; break;
            case 'b': out += '\b'; break;
            case 'f': out += '\f'; break;
            case 'n': out += '\n'; break;
            case 'r': out += '\r'; break;
            case 't': out += '\t'; break;
            case 'u': {
                unsigned cp = hex4();
                if (cp >= 0xD800 && cp < 0xDC00 && p_ + 6 <= s_.size() && s_[p_] == '\\' && s_[p_ + 1] == 'u') {
                    p_ += 2;
                    unsigned lo = hex4();
                    cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                }
                utf8(out, cp);
                break;
            }
            default: fail("bad escape");
            }
        }
    }
    Json number() {
        size_t start = p_;
        if (s_[p_] == '-') ++p_;
        bool integral = true;
        while (p_ < s_.size()) {
            char c = s_[p_];
            if (c >= '0' && c <= '9') {
                ++p_;
            } else if (c == '.' || c == 'e' || c == 'E' || c == '+' || c == '-') {
                integral = false;
                ++p_;
            } else
                break;
        }
        std::string token = s_.substr(start, p_ - start);
        char *end = nullptr;
        double d = std::strtod(token.c_str(), &end);
        if (!end || *end != '\0' || token == "-") fail("malformed number");
        if (!std::isfinite(d)) fail("number out of range");
        if (integral) {
            errno = 0;
            long long v = std::strtoll(token.c_str(), &end, 10);
            if (errno == 0 && *end == '\0') return Json(v);
        }
        return Json(d);
    }
    const std::string &s_;
    size_t p_ = 0;
    int max_depth_;
};

Json Json::parse(const std::string &text, int max_depth) { return JsonParser(text, max_depth).document(); }

}  // namespace nyk
