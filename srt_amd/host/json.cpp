#include "json.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace srt {

class JsonParser {
public:
    explicit JsonParser(const std::string& t) : s(t) {}

    Json parse_document() {
        Json v = parse_value();
        skip_ws();
        if (i != s.size()) fail("trailing characters");
        return v;
    }

private:
    const std::string& s;
    size_t i = 0;

    [[noreturn]] void fail(const char* what) {
        size_t line = 1, col = 1;
        for (size_t k = 0; k < i && k < s.size(); k++) {
            if (s[k] == '\n') { line++; col = 1; } else col++;
        }
        throw json_error(std::string("JSON parse error: ") + what + " at line " + std::to_string(line) + ", column " +
                         std::to_string(col));
    }

    void skip_ws() {
        for (;;) {
            while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r')) i++;
            if (i + 1 < s.size() && s[i] == '/' && s[i + 1] == '/') {
                while (i < s.size() && s[i] != '\n') i++;
                continue;
            }
            if (i + 1 < s.size() && s[i] == '/' && s[i + 1] == '*') {
                size_t e = s.find("*/", i + 2);
                if (e == std::string::npos) fail("unterminated comment");
                i = e + 2;
                continue;
            }
            break;
        }
    }

    Json parse_value() {
        skip_ws();
        if (i >= s.size()) fail("unexpected end of input");
        char c = s[i];
        if (c == '{') return parse_object();
        if (c == '[') return parse_array();
        if (c == '"') return Json(parse_string());
        if (c == 't' && s.compare(i, 4, "true") == 0) { i += 4; return Json(true); }
        if (c == 'f' && s.compare(i, 5, "false") == 0) { i += 5; return Json(false); }
        if (c == 'n' && s.compare(i, 4, "null") == 0) { i += 4; return Json(); }
        if (c == '-' || (c >= '0' && c <= '9')) return parse_number();
        fail("unexpected character");
    }

    Json parse_object() {
        Json o = Json::object();
        i++;
        skip_ws();
        if (i < s.size() && s[i] == '}') { i++; return o; }
        for (;;) {
            skip_ws();
            if (i >= s.size() || s[i] != '"') fail("expected a key string");
            std::string k = parse_string();
            skip_ws();
            if (i >= s.size() || s[i] != ':') fail("expected ':'");
            i++;
            o.obj_[k] = parse_value();
            skip_ws();
            if (i < s.size() && s[i] == ',') { i++; continue; }
            if (i < s.size() && s[i] == '}') { i++; return o; }
            fail("expected ',' or '}'");
        }
    }

    Json parse_array() {
        Json a = Json::array();
        i++;
        skip_ws();
        if (i < s.size() && s[i] == ']') { i++; return a; }
        for (;;) {
            a.arr_.push_back(parse_value());
            skip_ws();
            if (i < s.size() && s[i] == ',') { i++; continue; }
            if (i < s.size() && s[i] == ']') { i++; return a; }
            fail("expected ',' or ']'");
        }
    }

    static void put_utf8(std::string& o, unsigned cp) {
        if (cp < 0x80) o += (char)cp;
        else if (cp < 0x800) { o += (char)(0xC0 | (cp >> 6)); o += (char)(0x80 | (cp & 0x3F)); }
        else if (cp < 0x10000) {
            o += (char)(0xE0 | (cp >> 12)); o += (char)(0x80 | ((cp >> 6) & 0x3F)); o += (char)(0x80 | (cp & 0x3F));
        } else {
            o += (char)(0xF0 | (cp >> 18)); o += (char)(0x80 | ((cp >> 12) & 0x3F));
            o += (char)(0x80 | ((cp >> 6) & 0x3F)); o += (char)(0x80 | (cp & 0x3F));
        }
    }

    unsigned hex4() {
        if (i + 4 > s.size()) fail("short \\u escape");
        unsigned v = 0;
        for (int k = 0; k < 4; k++) {
            char c = s[i++];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (unsigned)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (unsigned)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (unsigned)(c - 'A' + 10);
            else fail("bad \\u escape");
        }
        return v;
    }

    std::string parse_string() {
        i++;  // opening quote
        std::string o;
        while (i < s.size()) {
            char c = s[i++];
            if (c == '"') return o;
            if (c != '\\') { o += c; continue; }
            if (i >= s.size()) break;
            char e = s[i++];
            switch (e) {
            case '"': o += '"'; break;
            case '\\': o += '\\'; break;
            case '/': o += '/'; break;
            case 'b': o += '\b'; break;
            case 'f': o += '\f'; break;
            case 'n': o += '\n'; break;
            case 'r': o += '\r'; break;
            case 't': o += '\t'; break;
            case 'u': {
                unsigned cp = hex4();
                if (cp >= 0xD800 && cp < 0xDC00 && i + 1 < s.size() && s[i] == '\\' && s[i + 1] == 'u') {
                    i += 2;
                    unsigned lo = hex4();
                    cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                }
                put_utf8(o, cp);
                break;
            }
            default: fail("bad escape");
            }
        }
        fail("unterminated string");
    }

    Json parse_number() {
        size_t b = i;
        if (s[i] == '-') i++;
        while (i < s.size() && ((s[i] >= '0' && s[i] <= '9') || s[i] == '.' || s[i] == 'e' || s[i] == 'E' || s[i] == '+' ||
                                s[i] == '-'))
            i++;
        std::string t = s.substr(b, i - b);
        char* end = nullptr;
        double d = std::strtod(t.c_str(), &end);
        if (!end || *end != '\0') fail("bad number");
        return Json(d);
    }
};

Json Json::parse(const std::string& text) { return JsonParser(text).parse_document(); }

Json Json::load(const std::string& path) {
    std::ifstream f(path);
    if (!f.is_open()) throw json_error("Failed to open file: " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return parse(ss.str());
}

static const Json& null_json() {
    static const Json n;
    return n;
}

const Json& Json::at(const std::string& k) const {
    if (!is_object()) throw json_error("not an object (looking up '" + k + "')");
    auto it = obj_.find(k);
    if (it == obj_.end()) throw json_error("key '" + k + "' not found");
    return it->second;
}
const Json& Json::operator[](const std::string& k) const {
    if (!is_object()) return null_json();
    auto it = obj_.find(k);
    return it == obj_.end() ? null_json() : it->second;
}
Json& Json::operator[](const std::string& k) {
    if (type_ == Type::Null) type_ = Type::Object;
    if (!is_object()) throw json_error("not an object (assigning '" + k + "')");
    return obj_[k];
}
const Json& Json::at(size_t idx) const {
    if (!is_array() || idx >= arr_.size()) throw json_error("array index out of range");
    return arr_[idx];
}
bool Json::get_bool() const {
    if (!is_bool()) throw json_error("expected a boolean");
    return b_;
}
double Json::get_number() const {
    if (!is_number()) throw json_error("expected a number");
    return d_;
}
const std::string& Json::get_string() const {
    if (!is_string()) throw json_error("expected a string");
    return s_;
}

int Json::value(const std::string& k, int def) const {
    const Json& v = (*this)[k];
    return v.is_null() ? def : v.get_int();
}
float Json::value(const std::string& k, float def) const {
    const Json& v = (*this)[k];
    return v.is_null() ? def : v.get_float();
}
double Json::value(const std::string& k, double def) const {
    const Json& v = (*this)[k];
    return v.is_null() ? def : v.get_number();
}
bool Json::value(const std::string& k, bool def) const {
    const Json& v = (*this)[k];
    return v.is_null() ? def : v.get_bool();
}
std::string Json::value(const std::string& k, const char* def) const {
    const Json& v = (*this)[k];
    return v.is_null() ? std::string(def) : v.get_string();
}
std::string Json::value(const std::string& k, const std::string& def) const {
    const Json& v = (*this)[k];
    return v.is_null() ? def : v.get_string();
}

static void dump_string(std::string& o, const std::string& s) {
    o += '"';
    for (char c : s) {
        switch (c) {
        case '"': o += "\\\""; break;
        case '\\': o += "\\\\"; break;
        case '\n': o += "\\n"; break;
        case '\t': o += "\\t"; break;
        case '\r': o += "\\r"; break;
        default:
            if ((unsigned char)c < 0x20) {
                char b[8];
                std::snprintf(b, sizeof(b), "\\u%04x", c);
                o += b;
            } else o += c;
        }
    }
    o += '"';
}

std::string Json::dump() const {
    std::string o;
    switch (type_) {
    case Type::Null: return "null";
    case Type::Bool: return b_ ? "true" : "false";
    case Type::Number: {
        if (std::isfinite(d_) && d_ == std::floor(d_) && std::fabs(d_) < 1e15) {
            char b[32];
            std::snprintf(b, sizeof(b), "%.0f", d_);
            return b;
        }
        char b[40];
        std::snprintf(b, sizeof(b), "%.9g", d_);
        return b;
    }
    case Type::String: dump_string(o, s_); return o;
    case Type::Array: {
        o += '[';
        for (size_t k = 0; k < arr_.size(); k++) { if (k) o += ','; o += arr_[k].dump(); }
        o += ']';
        return o;
    }
    case Type::Object: {
        o += '{';
        bool first = true;
        for (auto& kv : obj_) {
            if (!first) o += ',';
            first = false;
            dump_string(o, kv.first);
            o += ':';
            o += kv.second.dump();
        }
        o += '}';
        return o;
    }
    }
    return "null";
}

}  // namespace srt
