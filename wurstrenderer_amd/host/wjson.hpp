// wjson.hpp -- minimal JSON reader for scene files (objects, arrays, numbers,
// strings, booleans, null).  Key lookups follow the reference's
// Util::Json::GetValue semantics (src/common/util/json.h:44-49): a missing key
// yields the caller's default.
#pragma once
#include <array>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace wj {

struct Value {
    enum Type { Null, Bool, Number, String, Array, Object };
    Type type = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Value> arr;
    std::vector<std::pair<std::string, Value>> obj;

    const Value *find(const std::string &key) const {
        if (type != Object) return nullptr;
        for (const auto &kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
    bool has(const std::string &key) const { return find(key) != nullptr; }
    const Value &at(const std::string &key) const {
        const Value *v = find(key);
        if (!v) throw std::runtime_error("missing key \"" + key + "\"");
        return *v;
    }
    double number(const std::string &key, double def) const {
        const Value *v = find(key);
        if (!v) return def;
        if (v->type == Bool) return v->b ? 1.0 : 0.0;
        if (v->type != Number) throw std::runtime_error("key \"" + key + "\" is not a number");
        return v->num;
    }
    bool boolean(const std::string &key, bool def) const {
        const Value *v = find(key);
        if (!v) return def;
        if (v->type == Bool) return v->b;
        if (v->type == Number) return v->num != 0.0;
        throw std::runtime_error("key \"" + key + "\" is not a boolean");
    }
    std::string string(const std::string &key, const std::string &def) const {
        const Value *v = find(key);
        if (!v) return def;
        if (v->type != String) throw std::runtime_error("key \"" + key + "\" is not a string");
        return v->str;
    }
    // Vec3FromJson: a single number broadcasts, otherwise exactly 3 numbers
    // (src/common/util/json.h:9-15).
    std::array<double, 3> vec3(const std::string &key) const {
        const Value &v = at(key);
        if (v.type == Number) return {v.num, v.num, v.num};
        if (v.type == Array && v.arr.size() == 1 && v.arr[0].type == Number) return {v.arr[0].num, v.arr[0].num, v.arr[0].num};
        if (v.type == Array && v.arr.size() == 3) {
            std::array<double, 3> r;
            for (int i = 0; i < 3; i++) {
                if (v.arr[i].type != Number) throw std::runtime_error("can't parse " + key);
                r[i] = v.arr[i].num;
            }
            return r;
        }
        throw std::runtime_error("can't parse " + key);
    }
    // Ivec2FromJson: exactly 2 numbers (src/common/util/json.h:37-42)
    std::array<int, 2> ivec2(const std::string &key) const {
        const Value &v = at(key);
        if (v.type != Array || v.arr.size() != 2 || v.arr[0].type != Number || v.arr[1].type != Number)
            throw std::runtime_error("can't parse " + key);
        return {int(v.arr[0].num), int(v.arr[1].num)};
    }
};

class Parser {
  public:
    explicit Parser(const std::string &text) : s_(text) {}
    Value parse() {
        Value v = value();
        ws();
        if (i_ != s_.size()) fail("trailing characters");
        return v;
    }

  private:
    const std::string &s_;
    size_t i_ = 0;

    [[noreturn]] void fail(const std::string &what) const {
        throw std::runtime_error("json: " + what + " at offset " + std::to_string(i_));
    }
    void ws() {
        while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\t' || s_[i_] == '\n' || s_[i_] == '\r')) i_++;
    }
    bool lit(const char *w) {
        size_t n = 0;
        while (w[n]) n++;
        if (s_.compare(i_, n, w) == 0) { i_ += n; return true; }
        return false;
    }
    Value value() {
        ws();
        if (i_ >= s_.size()) fail("unexpected end");
        Value v;
        const char c = s_[i_];
        if (c == '{') {
            v.type = Value::Object;
            i_++;
            ws();
            if (i_ < s_.size() && s_[i_] == '}') { i_++; return v; }
            for (;;) {
                ws();
                if (i_ >= s_.size() || s_[i_] != '"') fail("expected a key");
                std::string k = string_lit();
                ws();
                if (i_ >= s_.size() || s_[i_] != ':') fail("expected ':'");
                i_++;
                v.obj.emplace_back(std::move(k), value());
                ws();
                if (i_ < s_.size() && s_[i_] == ',') { i_++; continue; }
                if (i_ < s_.size() && s_[i_] == '}') { i_++; return v; }
                fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.type = Value::Array;
            i_++;
            ws();
            if (i_ < s_.size() && s_[i_] == ']') { i_++; return v; }
            for (;;) {
                v.arr.push_back(value());
                ws();
                if (i_ < s_.size() && s_[i_] == ',') { i_++; continue; }
                if (i_ < s_.size() && s_[i_] == ']') { i_++; return v; }
                fail("expected ',' or ']'");
            }
        }
        if (c == '"') { v.type = Value::String; v.str = string_lit(); return v; }
        if (lit("true")) { v.type = Value::Bool; v.b = true; return v; }
        if (lit("false")) { v.type = Value::Bool; v.b = false; return v; }
        if (lit("null")) return v;
        if (c == '-' || (c >= '0' && c <= '9')) {
            const char *begin = s_.c_str() + i_;
            char *end = nullptr;
            v.type = Value::Number;
            v.num = std::strtod(begin, &end);
            if (end == begin) fail("bad number");
            i_ += size_t(end - begin);
            return v;
        }
        fail("unexpected character");
    }
    std::string string_lit() {
        i_++;  // opening quote
        std::string out;
        while (i_ < s_.size() && s_[i_] != '"') {
            char c = s_[i_++];
            if (c == '\\') {
                if (i_ >= s_.size()) fail("bad escape");
                const char e = s_[i_++];
                switch (e) {
                case 'n': out += '\n'; break;
                case 't': out += '\t'; break;
                case 'r': out += '\r'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'u': {
                    if (i_ + 4 > s_.size()) fail("bad \\u escape");
                    const unsigned cp = unsigned(std::strtoul(s_.substr(i_, 4).c_str(), nullptr, 16));
                    i_ += 4;
                    if (cp < 0x80) out += char(cp);
                    else if (cp < 0x800) { out += char(0xC0 | (cp >> 6)); out += char(0x80 | (cp & 0x3F)); }
                    else { out += char(0xE0 | (cp >> 12)); out += char(0x80 | ((cp >> 6) & 0x3F)); out += char(0x80 | (cp & 0x3F)); }
                    break;
                }
                default: out += e;
                }
            } else {
                out += c;
            }
        }
        if (i_ >= s_.size()) fail("unterminated string");
        i_++;
        return out;
    }
};

inline Value parse(const std::string &text) { return Parser(text).parse(); }

}  // namespace wj
