// Minimal JSON reader for the TDVMC config schema (objects, arrays, strings,
// numbers, true/false/null).  Integers and reals are kept apart because the
// reference's ConfigItem::setValue compares a value against the integer 0
// (src/ConfigItem.cpp:45-47): an explicit integer 0 leaves the previous value,
// a real 0.0 or a missing key sets 0 / empty.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace tdvmc {

struct JsonValue {
    enum Type { Null, Bool, Int, Real, String, Array, Object } type = Null;
    bool b = false;
    int64_t i = 0;
    double d = 0.0;
    std::string s;
    std::vector<JsonValue> arr;
    std::map<std::string, JsonValue> obj;

    bool is_null() const { return type == Null; }
    bool is_number() const { return type == Int || type == Real; }
    double as_double() const { return type == Int ? (double)i : (type == Real ? d : (type == Bool ? (b ? 1.0 : 0.0) : 0.0)); }
    int64_t as_int() const { return type == Int ? i : (type == Real ? (int64_t)d : (type == Bool ? (b ? 1 : 0) : 0)); }
    const std::string& as_string() const { return s; }
    const JsonValue* get(const std::string& key) const
    {
        if (type != Object) return nullptr;
        auto it = obj.find(key);
        return it == obj.end() ? nullptr : &it->second;
    }
};

class JsonParser {
public:
    explicit JsonParser(const std::string& text) : t_(text), p_(0) {}
    JsonValue parse()
    {
        JsonValue v = value();
        ws();
        if (p_ != t_.size()) fail("trailing characters");
        return v;
    }

private:
    const std::string& t_;
    size_t p_;

    [[noreturn]] void fail(const std::string& what)
    {
        throw std::runtime_error("JSON parse error at offset " + std::to_string(p_) + ": " + what);
    }
    void ws()
    {
        while (p_ < t_.size() && (t_[p_] == ' ' || t_[p_] == '\t' || t_[p_] == '\n' || t_[p_] == '\r')) p_++;
    }
    bool eat(char c)
    {
        ws();
        if (p_ < t_.size() && t_[p_] == c) {
            p_++;
            return true;
        }
        return false;
    }
    void expect(char c)
    {
        if (!eat(c)) fail(std::string("expected '") + c + "'");
    }
    JsonValue value()
    {
        ws();
        if (p_ >= t_.size()) fail("unexpected end");
        char c = t_[p_];
        if (c == '{') return object();
        if (c == '[') return array();
        if (c == '"') {
            JsonValue v;
            v.type = JsonValue::String;
            v.s = str();
            return v;
        }
        if (t_.compare(p_, 4, "true") == 0) {
            p_ += 4;
            JsonValue v;
            v.type = JsonValue::Bool;
            v.b = true;
            return v;
        }
        if (t_.compare(p_, 5, "false") == 0) {
            p_ += 5;
            JsonValue v;
            v.type = JsonValue::Bool;
            return v;
        }
        if (t_.compare(p_, 4, "null") == 0) {
            p_ += 4;
            return JsonValue();
        }
        return number();
    }
    JsonValue object()
    {
        JsonValue v;
        v.type = JsonValue::Object;
        expect('{');
        if (eat('}')) return v;
        do {
            ws();
            std::string k = str();
            expect(':');
            v.obj[k] = value();
        } while (eat(','));
        expect('}');
        return v;
    }
    JsonValue array()
    {
        JsonValue v;
        v.type = JsonValue::Array;
        expect('[');
        if (eat(']')) return v;
        do {
            v.arr.push_back(value());
        } while (eat(','));
        expect(']');
        return v;
    }
    std::string str()
    {
        if (p_ >= t_.size() || t_[p_] != '"') fail("expected string");
        p_++;
        std::string out;
        while (p_ < t_.size() && t_[p_] != '"') {
            char c = t_[p_++];
            if (c == '\\') {
                if (p_ >= t_.size()) fail("bad escape");
                char e = t_[p_++];
                switch (e) {
                case 'n': out += '\n'; break;
                case 't': out += '\t'; break;
                case 'r': out += '\r'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'u': {
                    if (p_ + 4 > t_.size()) fail("bad \\u escape");
                    unsigned cp = std::stoul(t_.substr(p_, 4), nullptr, 16);
                    p_ += 4;
                    if (cp < 0x80) out += (char)cp;
                    else if (cp < 0x800) {
                        out += (char)(0xC0 | (cp >> 6));
                        out += (char)(0x80 | (cp & 0x3F));
                    } else {
                        out += (char)(0xE0 | (cp >> 12));
                        out += (char)(0x80 | ((cp >> 6) & 0x3F));
                        out += (char)(0x80 | (cp & 0x3F));
                    }
                    break;
                }
                default: out += e;
                }
            } else {
                out += c;
            }
        }
        if (p_ >= t_.size()) fail("unterminated string");
        p_++;
        return out;
    }
    JsonValue number()
    {
        size_t start = p_;
        bool real = false;
        if (p_ < t_.size() && (t_[p_] == '-' || t_[p_] == '+')) p_++;
        while (p_ < t_.size()) {
            char c = t_[p_];
            if (c >= '0' && c <= '9') {
                p_++;
            } else if (c == '.' || c == 'e' || c == 'E' || c == '-' || c == '+') {
                real = true;
                p_++;
            } else {
                break;
            }
        }
        if (start == p_) fail("bad value");
        std::string tok = t_.substr(start, p_ - start);
        JsonValue v;
        try {
            if (real) {
                v.type = JsonValue::Real;
                v.d = std::stod(tok);
            } else {
                v.type = JsonValue::Int;
                v.i = std::stoll(tok);
            }
        } catch (...) {
            fail("bad number '" + tok + "'");
        }
        return v;
    }
};

}  // namespace tdvmc
