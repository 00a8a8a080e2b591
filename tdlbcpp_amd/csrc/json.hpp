// Minimal JSON value with parse and dump, enough for the reference's input and
// AllParams files (tdlbcpp uses nlohmann::json; objects there are key-sorted
// maps, which std::map reproduces).  Numbers keep their integer / unsigned /
// floating kind; get<T>() converts between numeric kinds like nlohmann's
// get<>() and throws on a type mismatch with bool, string, array or object.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace tdlb {
namespace json {

struct error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

class Value {
public:
    enum Kind { Null, Bool, Int, UInt, Float, String, Array, Object };

    Value() : k_(Null) {}
    Value(std::nullptr_t) : k_(Null) {}
    Value(bool b) : k_(Bool), b_(b) {}
    Value(int v) : k_(Int), i_(v) {}
    Value(long v) : k_(Int), i_(v) {}
    Value(long long v) : k_(Int), i_(v) {}
    Value(unsigned v) : k_(UInt), u_(v) {}
    Value(unsigned long v) : k_(UInt), u_(v) {}
    Value(unsigned long long v) : k_(UInt), u_(v) {}
    Value(double v) : k_(Float), d_(v) {}
    Value(float v) : k_(Float), d_(v) {}
    Value(const char* s) : k_(String), s_(s) {}
    Value(const std::string& s) : k_(String), s_(s) {}

    static Value object() { Value v; v.k_ = Object; return v; }
    static Value array() { Value v; v.k_ = Array; return v; }

    Kind kind() const { return k_; }
    bool is_object() const { return k_ == Object; }
    bool is_number() const { return k_ == Int || k_ == UInt || k_ == Float; }
    bool empty() const { return k_ == Null || (k_ == Object && o_.empty()) || (k_ == Array && a_.empty()); }

    // object access: operator[] creates (like nlohmann on a non-const object)
    Value& operator[](const std::string& key) {
        if (k_ == Null) k_ = Object;
        if (k_ != Object) throw error("[json] operator[] on a non-object with key '" + key + "'");
        return o_[key];
    }
    const Value& at(const std::string& key) const {
        if (k_ != Object) throw error("[json] at('" + key + "') on a non-object");
        auto it = o_.find(key);
        if (it == o_.end()) throw error("[json] key '" + key + "' not found");
        return it->second;
    }
    bool contains(const std::string& key) const { return k_ == Object && o_.count(key) > 0; }
    template <typename T>
    T value(const std::string& key, T dflt) const {
        return contains(key) ? at(key).get<T>() : dflt;
    }
    const std::map<std::string, Value>& items() const { return o_; }
    void push_back(const Value& v) {
        if (k_ == Null) k_ = Array;
        if (k_ != Array) throw error("[json] push_back on a non-array");
        a_.push_back(v);
    }
    const std::vector<Value>& elements() const { return a_; }

    template <typename T>
    T get() const {
        if constexpr (std::is_same<T, bool>::value) {
            if (k_ != Bool) throw error("[json] type must be boolean");
            return b_;
        } else if constexpr (std::is_same<T, std::string>::value) {
            if (k_ != String) throw error("[json] type must be string");
            return s_;
        } else if constexpr (std::is_same<T, Value>::value) {
            return *this;
        } else {
            static_assert(std::is_arithmetic<T>::value, "unsupported get<T>");
            switch (k_) {
                case Int: return static_cast<T>(i_);
                case UInt: return static_cast<T>(u_);
                case Float: return static_cast<T>(d_);
                default: throw error("[json] type must be number");
            }
        }
    }

    // ---------------------------------------------------------------- dump
    std::string dump(int indent = -1) const {
        std::string out;
        dump_to(out, indent, 0);
        return out;
    }

    // ---------------------------------------------------------------- parse
    static Value parse(const std::string& text) {
        size_t p = 0;
        Value v = parse_value(text, p);
        skip_ws(text, p);
        if (p != text.size()) throw error("[json] trailing characters at offset " + std::to_string(p));
        return v;
    }

private:
    Kind k_;
    bool b_ = false;
    int64_t i_ = 0;
    uint64_t u_ = 0;
    double d_ = 0.0;
    std::string s_;
    std::vector<Value> a_;
    std::map<std::string, Value> o_;

    static void put_string(std::string& out, const std::string& s) {
        out += '"';
        for (unsigned char c : s) {
            switch (c) {
                case '"': out += "\\\""; break;
                case '\\': out += "\\\\"; break;
                case '\n': out += "\\n"; break;
                case '\t': out += "\\t"; break;
                case '\r': out += "\\r"; break;
                case '\b': out += "\\b"; break;
                case '\f': out += "\\f"; break;
                default:
                    if (c < 0x20) {
                        char buf[8];
                        std::snprintf(buf, sizeof buf, "\\u%04x", c);
                        out += buf;
                    } else {
                        out += (char)c;
                    }
            }
        }
        out += '"';
    }

    // shortest decimal that reads back to the same double; integral values keep ".0"
    static void put_double(std::string& out, double d) {
        if (!std::isfinite(d)) { out += "null"; return; }
        char buf[40];
        for (int prec = 1; prec <= 17; prec++) {
            std::snprintf(buf, sizeof buf, "%.*g", prec, d);
            if (std::strtod(buf, nullptr) == d) break;
        }
        std::string s(buf);
        if (s.find_first_of(".eEn") == std::string::npos) s += ".0";
        out += s;
    }

    void dump_to(std::string& out, int indent, int level) const {
        auto nl = [&](int l) {
            if (indent >= 0) {
                out += '\n';
                out.append((size_t)(indent * l), ' ');
            }
        };
        switch (k_) {
            case Null: out += "null"; break;
            case Bool: out += b_ ? "true" : "false"; break;
            case Int: out += std::to_string(i_); break;
            case UInt: out += std::to_string(u_); break;
            case Float: put_double(out, d_); break;
            case String: put_string(out, s_); break;
            case Array: {
                out += '[';
                for (size_t n = 0; n < a_.size(); n++) {
                    if (n) out += ',';
                    nl(level + 1);
                    a_[n].dump_to(out, indent, level + 1);
                }
                if (!a_.empty()) nl(level);
                out += ']';
                break;
            }
            case Object: {
                out += '{';
                size_t n = 0;
                for (auto& kv : o_) {
                    if (n++) out += ',';
                    nl(level + 1);
                    put_string(out, kv.first);
                    out += indent >= 0 ? ": " : ":";
                    kv.second.dump_to(out, indent, level + 1);
                }
                if (!o_.empty()) nl(level);
                out += '}';
                break;
            }
        }
    }

    static void skip_ws(const std::string& t, size_t& p) {
        while (p < t.size() && (t[p] == ' ' || t[p] == '\n' || t[p] == '\r' || t[p] == '\t')) p++;
    }

    static std::string parse_string(const std::string& t, size_t& p) {
        if (t[p] != '"') throw error("[json] expected string at offset " + std::to_string(p));
        p++;
        std::string s;
        while (p < t.size() && t[p] != '"') {
            char c = t[p++];
            if (c == '\\') {
                if (p >= t.size()) break;
                char e = t[p++];
                switch (e) {
                    case 'n': s += '\n'; break;
                    case 't': s += '\t'; break;
                    case 'r': s += '\r'; break;
                    case 'b': s += '\b'; break;
                    case 'f': s += '\f'; break;
                    case 'u': {
                        if (p + 4 > t.size()) throw error("[json] bad \\u escape");
                        unsigned cp = (unsigned)std::strtoul(t.substr(p, 4).c_str(), nullptr, 16);
                        p += 4;
                        if (cp < 0x80) s += (char)cp;
                        else if (cp < 0x800) { s += (char)(0xC0 | (cp >> 6)); s += (char)(0x80 | (cp & 0x3F)); }
                        else { s += (char)(0xE0 | (cp >> 12)); s += (char)(0x80 | ((cp >> 6) & 0x3F));
                               s += (char)(0x80 | (cp & 0x3F)); }
                        break;
                    }
                    default: s += e;
                }
            } else {
                s += c;
            }
        }
        if (p >= t.size()) throw error("[json] unterminated string");
        p++;
        return s;
    }

    static Value parse_value(const std::string& t, size_t& p) {
        skip_ws(t, p);
        if (p >= t.size()) throw error("[json] unexpected end of input");
        const char c = t[p];
        if (c == '{') {
            Value v = object();
            p++;
            skip_ws(t, p);
            if (p < t.size() && t[p] == '}') { p++; return v; }
            while (true) {
                skip_ws(t, p);
                std::string key = parse_string(t, p);
                skip_ws(t, p);
                if (p >= t.size() || t[p] != ':') throw error("[json] expected ':' at offset " + std::to_string(p));
                p++;
                v.o_[key] = parse_value(t, p);
                skip_ws(t, p);
                if (p < t.size() && t[p] == ',') { p++; continue; }
                if (p < t.size() && t[p] == '}') { p++; return v; }
                throw error("[json] expected ',' or '}' at offset " + std::to_string(p));
            }
        }
        if (c == '[') {
            Value v = array();
            p++;
            skip_ws(t, p);
            if (p < t.size() && t[p] == ']') { p++; return v; }
            while (true) {
                v.a_.push_back(parse_value(t, p));
                skip_ws(t, p);
                if (p < t.size() && t[p] == ',') { p++; continue; }
                if (p < t.size() && t[p] == ']') { p++; return v; }
                throw error("[json] expected ',' or ']' at offset " + std::to_string(p));
            }
        }
        if (c == '"') return Value(parse_string(t, p));
        if (t.compare(p, 4, "true") == 0) { p += 4; return Value(true); }
        if (t.compare(p, 5, "false") == 0) { p += 5; return Value(false); }
        if (t.compare(p, 4, "null") == 0) { p += 4; return Value(); }
        // number
        size_t q = p;
        if (t[q] == '-') q++;
        bool is_float = false;
        while (q < t.size() && (std::isdigit((unsigned char)t[q]) || t[q] == '.' || t[q] == 'e' || t[q] == 'E' ||
                                t[q] == '+' || t[q] == '-')) {
            if (t[q] == '.' || t[q] == 'e' || t[q] == 'E') is_float = true;
            q++;
        }
        if (q == p) throw error("[json] unexpected character at offset " + std::to_string(p));
        const std::string num = t.substr(p, q - p);
        p = q;
        if (is_float) return Value(std::strtod(num.c_str(), nullptr));
        if (num[0] == '-') return Value((long long)std::strtoll(num.c_str(), nullptr, 10));
        return Value((unsigned long long)std::strtoull(num.c_str(), nullptr, 10));
    }
};

}  // namespace json
}  // namespace tdlb
