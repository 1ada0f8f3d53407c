// json_value.h -- a small JSON reader for the settings front-end.
//
// The reference reads its settings with nlohmann::json and converts the tree
// into nested std::any maps (reference json_helper.h:17-82): objects become
// maps, arrays vectors, strings strings, numbers double (with a '.', 'e' or
// 'E') or 64-bit integers, booleans bool, null null.  This reader builds the
// same tree shape directly; the typed accessors below keep the reference's
// rules (json_helper.h:95-154, 256-288):
//   * num():   any number or a boolean, converted to the requested type;
//   * str(), boolean(), object(): the exact JSON type, no conversion;
//   * path():  dotted keys with an optional [index] per segment.
// Every failure throws std::runtime_error whose message names the key.
#pragma once
#include <cctype>
#include <cerrno>
#include <cstdint>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace jsonv {

struct Value {
    enum Kind { Null, Bool, Int, Real, String, Array, Object };
    Kind kind = Null;
    bool b = false;
    long long i = 0;
    double d = 0.0;
    std::string s;
    std::vector<Value> a;
    std::map<std::string, Value> o;

    bool is_number() const { return kind == Int || kind == Real; }
};

class Parser {
public:
    explicit Parser(const std::string &text) : t_(text) {}

    Value document() {
        Value v = value();
        ws();
        if (p_ != t_.size()) fail("trailing characters");
        return v;
    }

private:
    const std::string &t_;
    size_t p_ = 0;

    [[noreturn]] void fail(const std::string &what) const {
        size_t line = 1, col = 1;
        for (size_t k = 0; k < p_ && k < t_.size(); ++k) {
            if (t_[k] == '\n') {
                ++line;
                col = 1;
            } else {
                ++col;
            }
        }
        throw std::runtime_error("JSON parse error at line " + std::to_string(line) + ", column " +
                                 std::to_string(col) + ": " + what);
    }

    void ws() {
        while (p_ < t_.size() && (t_[p_] == ' ' || t_[p_] == '\t' || t_[p_] == '\n' || t_[p_] == '\r')) ++p_;
    }

    bool eat(const char *lit) {
        size_t n = 0;
        while (lit[n]) ++n;
        if (t_.compare(p_, n, lit) == 0) {
            p_ += n;
            return true;
        }
        return false;
    }

    Value value() {
        ws();
        if (p_ >= t_.size()) fail("unexpected end of input");
        const char c = t_[p_];
        Value v;
        if (c == '{') return object();
        if (c == '[') return array();
        if (c == '"') {
            v.kind = Value::String;
            v.s = string();
            return v;
        }
        if (eat("true")) {
            v.kind = Value::Bool;
            v.b = true;
            return v;
        }
        if (eat("false")) {
            v.kind = Value::Bool;
            return v;
        }
        if (eat("null")) return v;
        if (c == '-' || std::isdigit((unsigned char)c)) return number();
        fail(std::string("unexpected character '") + c + "'");
    }

    Value object() {
        Value v;
        v.kind = Value::Object;
        ++p_;  // '{'
        ws();
        if (p_ < t_.size() && t_[p_] == '}') {
            ++p_;
            return v;
        }
        while (true) {
            ws();
            if (p_ >= t_.size() || t_[p_] != '"') fail("expected a key string");
            std::string key = string();
            ws();
            if (p_ >= t_.size() || t_[p_] != ':') fail("expected ':'");
            ++p_;
            v.o[key] = value();  // a repeated key keeps the last value, like nlohmann
            ws();
            if (p_ < t_.size() && t_[p_] == ',') {
                ++p_;
                continue;
            }
            if (p_ < t_.size() && t_[p_] == '}') {
                ++p_;
                return v;
            }
            fail("expected ',' or '}'");
        }
    }

    Value array() {
        Value v;
        v.kind = Value::Array;
        ++p_;  // '['
        ws();
        if (p_ < t_.size() && t_[p_] == ']') {
            ++p_;
            return v;
        }
        while (true) {
            v.a.push_back(value());
            ws();
            if (p_ < t_.size() && t_[p_] == ',') {
                ++p_;
                continue;
            }
            if (p_ < t_.size() && t_[p_] == ']') {
                ++p_;
                return v;
            }
            fail("expected ',' or ']'");
        }
    }

    static void put_utf8(std::string &out, uint32_t cp) {
        if (cp < 0x80) {
            out += (char)cp;
        } else if (cp < 0x800) {
            out += (char)(0xC0 | (cp >> 6));
            out += (char)(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            out += (char)(0xE0 | (cp >> 12));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        } else {
            out += (char)(0xF0 | (cp >> 18));
            out += (char)(0x80 | ((cp >> 12) & 0x3F));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        }
    }

    uint32_t hex4() {
        if (p_ + 4 > t_.size()) fail("short \\u escape");
        uint32_t cp = 0;
        for (int k = 0; k < 4; ++k) {
            const char h = t_[p_++];
            cp <<= 4;
            if (h >= '0' && h <= '9')
                cp |= (uint32_t)(h - '0');
            else if (h >= 'a' && h <= 'f')
                cp |= (uint32_t)(h - 'a' + 10);
            else if (h >= 'A' && h <= 'F')
                cp |= (uint32_t)(h - 'A' + 10);
            else
                fail("bad \\u escape");
        }
        return cp;
    }

    std::string string() {
        ++p_;  // opening quote
        std::string out;
        while (true) {
            if (p_ >= t_.size()) fail("unterminated string");
            const char c = t_[p_++];
            if (c == '"') return out;
            if ((unsigned char)c < 0x20) fail("control character in string");
            if (c != '\\') {
                out += c;
                continue;
            }
            if (p_ >= t_.size()) fail("unterminated escape");
            const char e = t_[p_++];
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
                    uint32_t cp = hex4();
                    if (cp >= 0xD800 && cp < 0xDC00) {  // surrogate pair
                        if (!eat("\\u")) fail("lone high surrogate");
                        const uint32_t lo = hex4();
                        if (lo < 0xDC00 || lo >= 0xE000) fail("bad low surrogate");
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    put_utf8(out, cp);
                    break;
                }
                default: fail(std::string("bad escape \\") + e);
            }
        }
    }

    Value number() {
        const size_t start = p_;
        bool real = false;
        if (t_[p_] == '-') ++p_;
        if (p_ >= t_.size() || !std::isdigit((unsigned char)t_[p_])) fail("bad number");
        if (t_[p_] == '0') {
            ++p_;
        } else {
            while (p_ < t_.size() && std::isdigit((unsigned char)t_[p_])) ++p_;
        }
        if (p_ < t_.size() && t_[p_] == '.') {
            real = true;
            ++p_;
            if (p_ >= t_.size() || !std::isdigit((unsigned char)t_[p_])) fail("bad fraction");
            while (p_ < t_.size() && std::isdigit((unsigned char)t_[p_])) ++p_;
        }
        if (p_ < t_.size() && (t_[p_] == 'e' || t_[p_] == 'E')) {
            real = true;
            ++p_;
            if (p_ < t_.size() && (t_[p_] == '+' || t_[p_] == '-')) ++p_;
            if (p_ >= t_.size() || !std::isdigit((unsigned char)t_[p_])) fail("bad exponent");
            while (p_ < t_.size() && std::isdigit((unsigned char)t_[p_])) ++p_;
        }
        const std::string tok = t_.substr(start, p_ - start);
        Value v;
        if (!real) {
            errno = 0;
            char *end = nullptr;
            const long long x = std::strtoll(tok.c_str(), &end, 10);
            if (errno == 0 && end && *end == 0) {
                v.kind = Value::Int;
                v.i = x;
                return v;
            }
            // integers beyond 64 bits read as doubles
        }
        v.kind = Value::Real;
        v.d = std::strtod(tok.c_str(), nullptr);
        return v;
    }
};

inline Value parse(const std::string &text) { return Parser(text).document(); }

// ---- typed access with the reference's conversion rules ----

inline const Value &must_get(const Value &m, const std::string &key) {
    if (m.kind != Value::Object) throw std::runtime_error("not an object when looking up key: " + key);
    auto it = m.o.find(key);
    if (it == m.o.end()) throw std::runtime_error("Missing key: " + key);
    return it->second;
}

// dotted path with an optional [index] per segment (json_helper.h:180-254)
inline const Value &must_get_path(const Value &root, const std::string &path) {
    std::vector<std::string> parts;
    std::string cur;
    for (char c : path) {
        if (c == '.') {
            if (!cur.empty()) parts.push_back(cur);
            cur.clear();
        } else {
            cur += c;
        }
    }
    if (!cur.empty()) parts.push_back(cur);
    if (parts.empty()) throw std::runtime_error("Empty path");
    const Value *v = &root;
    for (const std::string &seg : parts) {
        std::string key = seg;
        long idx = -1;
        const size_t lb = seg.find('[');
        if (lb != std::string::npos) {
            const size_t rb = seg.find(']', lb + 1);
            if (rb == std::string::npos) throw std::runtime_error("Malformed path segment: " + seg);
            const std::string digits = seg.substr(lb + 1, rb - lb - 1);
            if (digits.empty()) throw std::runtime_error("Empty index in path segment: " + seg);
            for (char c : digits)
                if (!std::isdigit((unsigned char)c)) throw std::runtime_error("Non-numeric index in path segment: " + seg);
            key = seg.substr(0, lb);
            idx = std::stol(digits);
        }
        if (v->kind != Value::Object) throw std::runtime_error("Not an object at path segment: " + seg);
        auto it = v->o.find(key);
        if (it == v->o.end()) throw std::runtime_error("Missing key in path: " + key);
        v = &it->second;
        if (idx >= 0) {
            if (v->kind != Value::Array) throw std::runtime_error("Not an array at path segment: " + seg);
            if ((size_t)idx >= v->a.size()) throw std::runtime_error("Index out of range in path segment: " + seg);
            v = &v->a[(size_t)idx];
        }
    }
    return *v;
}

template <typename T>
T as_num(const Value &v, const std::string &what) {
    switch (v.kind) {
        case Value::Int: return static_cast<T>(v.i);
        case Value::Real: return static_cast<T>(v.d);
        case Value::Bool: return static_cast<T>(v.b);
        default: throw std::runtime_error("not a number: " + what);
    }
}

template <typename T>
T num(const Value &m, const std::string &path) {
    return as_num<T>(must_get_path(m, path), path);
}

inline bool boolean(const Value &m, const std::string &path) {
    const Value &v = must_get_path(m, path);
    if (v.kind != Value::Bool) throw std::runtime_error("not a boolean: " + path);
    return v.b;
}

inline const std::string &str(const Value &m, const std::string &path) {
    const Value &v = must_get_path(m, path);
    if (v.kind != Value::String) throw std::runtime_error("not a string: " + path);
    return v.s;
}

inline const Value &object(const Value &m, const std::string &path) {
    const Value &v = must_get_path(m, path);
    if (v.kind != Value::Object) throw std::runtime_error("not an object: " + path);
    return v;
}

template <typename T>
std::vector<T> num_array(const Value &m, const std::string &path) {
    const Value &v = must_get_path(m, path);
    if (v.kind != Value::Array) throw std::runtime_error("not an array: " + path);
    std::vector<T> out;
    for (const Value &e : v.a) out.push_back(as_num<T>(e, path));
    return out;
}

}  // namespace jsonv
