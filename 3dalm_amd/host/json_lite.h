// json_lite.h -- a small JSON reader for the case files (setup.json, mesh.json,
// windturbine.json).  Numbers are converted with strtod, so a decimal in a case
// file becomes the same double the reference's reader produces.
#pragma once

#include <cctype>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace jl {

struct Value {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0;
    std::string str;
    std::vector<Value> arr;
    std::vector<std::pair<std::string, Value>> obj;

    bool is_number() const { return kind == Number; }
    bool is_object() const { return kind == Object; }
    bool is_array() const { return kind == Array; }
    bool has(const std::string& k) const {
        for (auto& kv : obj)
            if (kv.first == k) return true;
        return false;
    }
    const Value& operator[](const std::string& k) const {
        for (auto& kv : obj)
            if (kv.first == k) return kv.second;
        throw std::runtime_error("missing key \"" + k + "\"");
    }
    const Value& operator[](size_t i) const {
        if (kind != Array || i >= arr.size()) throw std::runtime_error("array index out of range");
        return arr[i];
    }
    size_t size() const { return kind == Array ? arr.size() : obj.size(); }
    double d() const {
        if (kind != Number) throw std::runtime_error("expected a number");
        return num;
    }
    long long i() const { return (long long)d(); }
    const std::string& s() const {
        if (kind != String) throw std::runtime_error("expected a string");
        return str;
    }
    std::vector<double> dvec() const {
        std::vector<double> v;
        for (auto& e : arr) v.push_back(e.d());
        return v;
    }
};

class Parser {
  public:
    explicit Parser(const std::string& t) : t_(t) {}
    Value parse() {
        Value v = value();
        ws();
        if (p_ != t_.size()) fail("trailing characters");
        return v;
    }

  private:
    const std::string& t_;
    size_t p_ = 0;

    [[noreturn]] void fail(const char* m) {
        throw std::runtime_error(std::string("json: ") + m + " at offset " + std::to_string(p_));
    }
    void ws() {
        while (p_ < t_.size() && std::isspace((unsigned char)t_[p_])) ++p_;
    }
    bool lit(const char* w) {
        size_t n = std::char_traits<char>::length(w);
        if (t_.compare(p_, n, w) == 0) {
            p_ += n;
            return true;
        }
        return false;
    }
    Value value() {
        ws();
        if (p_ >= t_.size()) fail("unexpected end");
        char c = t_[p_];
        Value v;
        if (c == '{') {
            v.kind = Value::Object;
            ++p_;
            ws();
            if (t_[p_] == '}') {
                ++p_;
                return v;
            }
            for (;;) {
                ws();
                std::string k = string();
                ws();
                if (t_[p_++] != ':') fail("expected ':'");
                v.obj.emplace_back(k, value());
                ws();
                char e = t_[p_++];
                if (e == '}') break;
                if (e != ',') fail("expected ',' or '}'");
            }
        } else if (c == '[') {
            v.kind = Value::Array;
            ++p_;
            ws();
            if (t_[p_] == ']') {
                ++p_;
                return v;
            }
            for (;;) {
                v.arr.push_back(value());
                ws();
                char e = t_[p_++];
                if (e == ']') break;
                if (e != ',') fail("expected ',' or ']'");
            }
        } else if (c == '"') {
            v.kind = Value::String;
            v.str = string();
        } else if (lit("true")) {
            v.kind = Value::Bool;
            v.b = true;
        } else if (lit("false")) {
            v.kind = Value::Bool;
        } else if (lit("null")) {
            v.kind = Value::Null;
        } else {
            const char* s = t_.c_str() + p_;
            char* end = nullptr;
            v.num = std::strtod(s, &end);
            if (end == s) fail("bad value");
            v.kind = Value::Number;
            p_ += (size_t)(end - s);
        }
        return v;
    }
    std::string string() {
        if (t_[p_] != '"') fail("expected string");
        ++p_;
        std::string out;
        while (p_ < t_.size() && t_[p_] != '"') {
            char c = t_[p_++];
            if (c == '\\') {
                char e = t_[p_++];
                switch (e) {
                case 'n': out += '\n'; break;
                case 't': out += '\t'; break;
                case 'r': out += '\r'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'u': p_ += 4; out += '?'; break;
                default: out += e; break;
                }
            } else {
                out += c;
            }
        }
        if (p_ >= t_.size()) fail("unterminated string");
        ++p_;
        return out;
    }
};

inline Value parse(const std::string& text) { return Parser(text).parse(); }

}  // namespace jl
