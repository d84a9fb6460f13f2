// Minimal JSON value for the graph host: null, bool, integer, float, string, array and object
// (keys kept sorted, as graph descriptions are stored), a strict parser, a serializer and
// RFC 6901 JSON pointers. Integers and floats stay distinct so that "8" set from the command
// line stores as the integer 8 and round-trips unchanged.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace mrt::json {

class Value;
using Object = std::map<std::string, Value>;
using Array = std::vector<Value>;

class Value {
  public:
    enum class Type { Null, Bool, Int, Float, String, Array, Object };

    Value() = default;
    Value(std::nullptr_t) {}
    Value(bool b) : t_(Type::Bool), b_(b) {}
    Value(int v) : t_(Type::Int), i_(v) {}
    Value(int64_t v) : t_(Type::Int), i_(v) {}
    Value(uint32_t v) : t_(Type::Int), i_(int64_t(v)) {}
    Value(double v) : t_(Type::Float), f_(v) {}
    Value(const char* s) : t_(Type::String), s_(s) {}
    Value(std::string s) : t_(Type::String), s_(std::move(s)) {}
    Value(Array a) : t_(Type::Array), a_(std::make_shared<Array>(std::move(a))) {}
    Value(Object o) : t_(Type::Object), o_(std::make_shared<Object>(std::move(o))) {}
    Value(const Value& v) { *this = v; }
    Value& operator=(const Value& v) {
        if (this == &v) return *this;
        t_ = v.t_;
        b_ = v.b_;
        i_ = v.i_;
        f_ = v.f_;
        s_ = v.s_;
        a_ = v.a_ ? std::make_shared<Array>(*v.a_) : nullptr;
        o_ = v.o_ ? std::make_shared<Object>(*v.o_) : nullptr;
        return *this;
    }
    Value(Value&&) = default;
    Value& operator=(Value&&) = default;

    static Value object() { return Value(Object{}); }
    static Value array() { return Value(Array{}); }

    Type type() const { return t_; }
    bool is_null() const { return t_ == Type::Null; }
    bool is_bool() const { return t_ == Type::Bool; }
    bool is_number() const { return t_ == Type::Int || t_ == Type::Float; }
    bool is_int() const { return t_ == Type::Int; }
    bool is_string() const { return t_ == Type::String; }
    bool is_array() const { return t_ == Type::Array; }
    bool is_object() const { return t_ == Type::Object; }

    bool as_bool() const {
        if (t_ != Type::Bool) throw std::runtime_error("json: not a boolean");
        return b_;
    }
    int64_t as_int() const {
        if (t_ == Type::Int) return i_;
        if (t_ == Type::Float && std::floor(f_) == f_) return int64_t(f_);
        throw std::runtime_error("json: not an integer");
    }
    double as_double() const {
        if (t_ == Type::Int) return double(i_);
        if (t_ == Type::Float) return f_;
        throw std::runtime_error("json: not a number");
    }
    const std::string& as_string() const {
        if (t_ != Type::String) throw std::runtime_error("json: not a string");
        return s_;
    }
    Array& as_array() {
        if (t_ != Type::Array) throw std::runtime_error("json: not an array");
        return *a_;
    }
    const Array& as_array() const {
        if (t_ != Type::Array) throw std::runtime_error("json: not an array");
        return *a_;
    }
    Object& as_object() {
        if (t_ != Type::Object) throw std::runtime_error("json: not an object");
        return *o_;
    }
    const Object& as_object() const {
        if (t_ != Type::Object) throw std::runtime_error("json: not an object");
        return *o_;
    }

    // Object member access; a null value becomes an object (as assignment through it implies).
    Value& operator[](const std::string& key) {
        if (t_ == Type::Null) *this = object();
        return as_object()[key];
    }
    const Value* find(const std::string& key) const {
        if (t_ != Type::Object) return nullptr;
        auto it = o_->find(key);
        return it == o_->end() ? nullptr : &it->second;
    }
    bool contains(const std::string& key) const { return find(key) != nullptr; }
    size_t size() const {
        if (t_ == Type::Array) return a_->size();
        if (t_ == Type::Object) return o_->size();
        return t_ == Type::Null ? 0 : 1;
    }
    bool empty() const { return size() == 0; }
    bool erase(const std::string& key) { return t_ == Type::Object && o_->erase(key) > 0; }
    void push_back(Value v) {
        if (t_ == Type::Null) *this = array();
        as_array().push_back(std::move(v));
    }

    // value of member `key`, or `def` when absent
    Value value(const std::string& key, Value def) const {
        const Value* v = find(key);
        return v ? *v : def;
    }

    friend bool operator==(const Value& a, const Value& b) {
        if (a.is_number() && b.is_number()) {
            if (a.t_ == Type::Int && b.t_ == Type::Int) return a.i_ == b.i_;
            return a.as_double() == b.as_double();
        }
        if (a.t_ != b.t_) return false;
        switch (a.t_) {
        case Type::Null: return true;
        case Type::Bool: return a.b_ == b.b_;
        case Type::String: return a.s_ == b.s_;
        case Type::Array: return *a.a_ == *b.a_;
        case Type::Object: return *a.o_ == *b.o_;
        default: return false;
        }
    }
    friend bool operator!=(const Value& a, const Value& b) { return !(a == b); }

  private:
    Type t_ = Type::Null;
    bool b_ = false;
    int64_t i_ = 0;
    double f_ = 0.0;
    std::string s_;
    std::shared_ptr<Array> a_;
    std::shared_ptr<Object> o_;
};

// ---------------------------------------------------------------------------------------------
// parsing

class Parser {
  public:
    explicit Parser(const std::string& text) : s_(text) {}
    Value parse() {
        ws();
        Value v = value();
        ws();
        if (p_ != s_.size()) fail("trailing characters");
        return v;
    }

  private:
    [[noreturn]] void fail(const char* what) const {
        throw std::runtime_error(std::string("json parse error at offset ") + std::to_string(p_) + ": " + what);
    }
    void ws() {
        while (p_ < s_.size() && (s_[p_] == ' ' || s_[p_] == '\t' || s_[p_] == '\n' || s_[p_] == '\r')) p_++;
    }
    bool lit(const char* w) {
        size_t n = std::char_traits<char>::length(w);
        if (s_.compare(p_, n, w) == 0) {
            p_ += n;
            return true;
        }
        return false;
    }
    Value value() {
        if (p_ >= s_.size()) fail("unexpected end");
        const char c = s_[p_];
        if (c == '{') return object();
        if (c == '[') return array();
        if (c == '"') return Value(string());
        if (lit("true")) return Value(true);
        if (lit("false")) return Value(false);
        if (lit("null")) return Value();
        if (c == '-' || (c >= '0' && c <= '9')) return number();
        fail("unexpected character");
    }
    Value object() {
        p_++;
        Object o;
        ws();
        if (p_ < s_.size() && s_[p_] == '}') {
            p_++;
            return Value(std::move(o));
        }
        while (true) {
            ws();
            if (p_ >= s_.size() || s_[p_] != '"') fail("expected key");
            std::string k = string();
            ws();
            if (p_ >= s_.size() || s_[p_] != ':') fail("expected ':'");
            p_++;
            ws();
            o[k] = value();
            ws();
            if (p_ < s_.size() && s_[p_] == ',') {
                p_++;
                continue;
            }
            if (p_ < s_.size() && s_[p_] == '}') {
                p_++;
                return Value(std::move(o));
            }
            fail("expected ',' or '}'");
        }
    }
    Value array() {
        p_++;
        Array a;
        ws();
        if (p_ < s_.size() && s_[p_] == ']') {
            p_++;
            return Value(std::move(a));
        }
        while (true) {
            ws();
            a.push_back(value());
            ws();
            if (p_ < s_.size() && s_[p_] == ',') {
                p_++;
                continue;
            }
            if (p_ < s_.size() && s_[p_] == ']') {
                p_++;
                return Value(std::move(a));
            }
            fail("expected ',' or ']'");
        }
    }
    static void utf8(std::string& out, uint32_t cp) {
        if (cp < 0x80) {
            out += char(cp);
        } else if (cp < 0x800) {
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
    uint32_t hex4() {
        if (p_ + 4 > s_.size()) fail("short unicode escape");
        uint32_t v = 0;
        for (int k = 0; k < 4; k++) {
            const char h = s_[p_++];
            v <<= 4;
            if (h >= '0' && h <= '9') v |= uint32_t(h - '0');
            else if (h >= 'a' && h <= 'f') v |= uint32_t(h - 'a' + 10);
            else if (h >= 'A' && h <= 'F') v |= uint32_t(h - 'A' + 10);
            else fail("bad unicode escape");
        }
        return v;
    }
    std::string string() {
        p_++;
        std::string out;
        while (true) {
            if (p_ >= s_.size()) fail("unterminated string");
            const char c = s_[p_++];
            if (c == '"') return out;
            if (static_cast<unsigned char>(c) < 0x20) fail("control character in string");
            if (c != '\\') {
                out += c;
                continue;
            }
            if (p_ >= s_.size()) fail("bad escape");
            const char e = s_[p_++];
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
                if (cp >= 0xD800 && cp < 0xDC00 && p_ + 6 <= s_.size() && s_[p_] == '\\' && s_[p_ + 1] == 'u') {
                    p_ += 2;
                    const uint32_t lo = hex4();
                    cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                }
                utf8(out, cp);
                break;
            }
            default: fail("bad escape");
            }
        }
    }
    Value number() {
        const size_t start = p_;
        if (s_[p_] == '-') p_++;
        bool is_float = false;
        while (p_ < s_.size() && s_[p_] >= '0' && s_[p_] <= '9') p_++;
        if (p_ < s_.size() && s_[p_] == '.') {
            is_float = true;
            p_++;
            while (p_ < s_.size() && s_[p_] >= '0' && s_[p_] <= '9') p_++;
        }
        if (p_ < s_.size() && (s_[p_] == 'e' || s_[p_] == 'E')) {
            is_float = true;
            p_++;
            if (p_ < s_.size() && (s_[p_] == '+' || s_[p_] == '-')) p_++;
            while (p_ < s_.size() && s_[p_] >= '0' && s_[p_] <= '9') p_++;
        }
        const std::string t = s_.substr(start, p_ - start);
        if (t == "-" || t.empty()) fail("bad number");
        if (!is_float) {
            try {
                return Value(int64_t(std::stoll(t)));
            } catch (...) {
                return Value(std::stod(t));
            }
        }
        return Value(std::stod(t));
    }

    const std::string& s_;
    size_t p_ = 0;
};

inline Value parse(const std::string& text) { return Parser(text).parse(); }

// Parses or returns false (no exception): the graph's "is it JSON" test.
inline bool try_parse(const std::string& text, Value& out) {
    try {
        out = parse(text);
        return true;
    } catch (...) {
        return false;
    }
}

// ---------------------------------------------------------------------------------------------
// serialization

inline void dump_string(std::string& out, const std::string& s) {
    out += '"';
    for (const char c : s) {
        switch (c) {
        case '"': out += "\\\""; break;
        case '\\': out += "\\\\"; break;
        case '\n': out += "\\n"; break;
        case '\r': out += "\\r"; break;
        case '\t': out += "\\t"; break;
        case '\b': out += "\\b"; break;
        case '\f': out += "\\f"; break;
        default:
            if (static_cast<unsigned char>(c) < 0x20) {
                char buf[8];
                std::snprintf(buf, sizeof buf, "\\u%04x", unsigned(static_cast<unsigned char>(c)));
                out += buf;
            } else {
                out += c;
            }
        }
    }
    out += '"';
}

inline std::string format_double(double v) {
    if (!std::isfinite(v)) return "null";
    char buf[32];
    for (int prec = 15; prec <= 17; prec++) { // shortest of 15..17 digits that round-trips
        std::snprintf(buf, sizeof buf, "%.*g", prec, v);
        if (std::stod(buf) == v) break;
    }
    std::string s = buf;
    if (s.find_first_of(".eE") == std::string::npos) s += ".0";
    return s;
}

inline void dump(std::string& out, const Value& v, int indent, int depth) {
    const std::string nl = indent >= 0 ? "\n" : "";
    auto pad = [&](int d) { return indent >= 0 ? std::string(size_t(indent * d), ' ') : std::string(); };
    switch (v.type()) {
    case Value::Type::Null: out += "null"; break;
    case Value::Type::Bool: out += v.as_bool() ? "true" : "false"; break;
    case Value::Type::Int: out += std::to_string(v.as_int()); break;
    case Value::Type::Float: out += format_double(v.as_double()); break;
    case Value::Type::String: dump_string(out, v.as_string()); break;
    case Value::Type::Array: {
        const Array& a = v.as_array();
        if (a.empty()) {
            out += "[]";
            break;
        }
        out += "[" + nl;
        for (size_t k = 0; k < a.size(); k++) {
            out += pad(depth + 1);
            dump(out, a[k], indent, depth + 1);
            out += (k + 1 < a.size() ? "," : "") + nl;
        }
        out += pad(depth) + "]";
        break;
    }
    case Value::Type::Object: {
        const Object& o = v.as_object();
        if (o.empty()) {
            out += "{}";
            break;
        }
        out += "{" + nl;
        size_t k = 0;
        for (const auto& [key, val] : o) {
            out += pad(depth + 1);
            dump_string(out, key);
            out += indent >= 0 ? ": " : ":";
            dump(out, val, indent, depth + 1);
            out += (++k < o.size() ? "," : "") + nl;
        }
        out += pad(depth) + "}";
        break;
    }
    }
}

inline std::string dump(const Value& v, int indent = -1) {
    std::string out;
    dump(out, v, indent, 0);
    return out;
}

// ---------------------------------------------------------------------------------------------
// JSON pointers (RFC 6901)

inline std::vector<std::string> pointer_tokens(const std::string& ptr) {
    std::vector<std::string> toks;
    if (ptr.empty()) return toks;
    if (ptr[0] != '/') throw std::runtime_error("json pointer must start with '/': " + ptr);
    size_t p = 1;
    while (true) {
        const size_t q = ptr.find('/', p);
        std::string t = ptr.substr(p, q == std::string::npos ? std::string::npos : q - p);
        std::string u;
        for (size_t k = 0; k < t.size(); k++) {
            if (t[k] == '~' && k + 1 < t.size() && (t[k + 1] == '0' || t[k + 1] == '1')) {
                u += t[k + 1] == '0' ? '~' : '/';
                k++;
            } else {
                u += t[k];
            }
        }
        toks.push_back(u);
        if (q == std::string::npos) break;
        p = q + 1;
    }
    return toks;
}

inline bool array_index(const std::string& t, size_t& idx) {
    if (t.empty() || (t.size() > 1 && t[0] == '0')) return false;
    for (char c : t)
        if (c < '0' || c > '9') return false;
    idx = size_t(std::stoull(t));
    return true;
}

// The value at `ptr`, or nullptr.
inline const Value* pointer_find(const Value& root, const std::string& ptr) {
    const Value* cur = &root;
    for (const std::string& t : pointer_tokens(ptr)) {
        if (cur->is_object()) {
            cur = cur->find(t);
            if (!cur) return nullptr;
        } else if (cur->is_array()) {
            size_t i;
            if (!array_index(t, i) || i >= cur->as_array().size()) return nullptr;
            cur = &cur->as_array()[i];
        } else {
            return nullptr;
        }
    }
    return cur;
}

// The value at `ptr`, creating objects (and array slots, for numeric tokens into arrays) on the
// way, as assignment through a pointer does.
inline Value& pointer_ref(Value& root, const std::string& ptr) {
    Value* cur = &root;
    for (const std::string& t : pointer_tokens(ptr)) {
        size_t i;
        if (cur->is_array() && (t == "-" || array_index(t, i))) {
            Array& a = cur->as_array();
            if (t == "-") i = a.size();
            if (i >= a.size()) a.resize(i + 1);
            cur = &a[i];
        } else {
            if (!cur->is_object() && !cur->is_null()) *cur = Value::object();
            cur = &(*cur)[t];
        }
    }
    return *cur;
}

} // namespace mrt::json
