// Minimal JSON reader for the guide's configuration files (objects, arrays,
// numbers, strings, booleans, null).  Only what the nn/*.json configuration
// files of the reference use (common/configs/nn/*.json) is needed.
#pragma once
#include <cctype>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace dfg {

struct Json {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Json> arr;
    std::map<std::string, Json> obj;

    bool is_object() const { return kind == Object; }
    bool is_array() const { return kind == Array; }
    bool contains(const std::string& k) const { return kind == Object && obj.count(k) != 0; }
    const Json& at(const std::string& k) const {
        auto it = obj.find(k);
        if (kind != Object || it == obj.end()) throw std::runtime_error("json: missing key '" + k + "'");
        return it->second;
    }
    double number_or(const std::string& k, double def) const {
        if (!contains(k)) return def;
        const Json& v = at(k);
        if (v.kind == Number) return v.num;
        if (v.kind == Bool) return v.b ? 1.0 : 0.0;
        throw std::runtime_error("json: key '" + k + "' is not a number");
    }
    std::string string_or(const std::string& k, const std::string& def) const {
        if (!contains(k)) return def;
        const Json& v = at(k);
        if (v.kind != String) throw std::runtime_error("json: key '" + k + "' is not a string");
        return v.str;
    }
    bool bool_or(const std::string& k, bool def) const {
        if (!contains(k)) return def;
        const Json& v = at(k);
        if (v.kind == Bool) return v.b;
        if (v.kind == Number) return v.num != 0.0;
        throw std::runtime_error("json: key '" + k + "' is not a boolean");
    }
};

class JsonParser {
public:
    explicit JsonParser(const std::string& s) : s_(s) {}
    Json parse() {
        Json v = value();
        ws();
        if (p_ != s_.size()) fail("trailing characters");
        return v;
    }

private:
    const std::string& s_;
    size_t p_ = 0;

    [[noreturn]] void fail(const char* what) {
        throw std::runtime_error(std::string("json: ") + what + " at offset " + std::to_string(p_));
    }
    void ws() {
        while (p_ < s_.size()) {
            char c = s_[p_];
            if (c == ' ' || c == '\t' || c == '\n' || c == '\r') { ++p_; continue; }
            if (c == '/' && p_ + 1 < s_.size() && s_[p_ + 1] == '/') {  // tolerate // comments
                while (p_ < s_.size() && s_[p_] != '\n') ++p_;
                continue;
            }
            break;
        }
    }
    bool lit(const char* w) {
        size_t n = std::char_traits<char>::length(w);
        if (s_.compare(p_, n, w) == 0) { p_ += n; return true; }
        return false;
    }
    Json value() {
        ws();
        if (p_ >= s_.size()) fail("unexpected end");
        char c = s_[p_];
        Json v;
        if (c == '{') {
            v.kind = Json::Object; ++p_; ws();
            if (p_ < s_.size() && s_[p_] == '}') { ++p_; return v; }
            for (;;) {
                ws();
                if (p_ >= s_.size() || s_[p_] != '"') fail("expected key");
                std::string k = string_lit();
                ws();
                if (p_ >= s_.size() || s_[p_] != ':') fail("expected ':'");
                ++p_;
                v.obj[k] = value();
                ws();
                if (p_ < s_.size() && s_[p_] == ',') { ++p_; continue; }
                if (p_ < s_.size() && s_[p_] == '}') { ++p_; break; }
                fail("expected ',' or '}'");
            }
        } else if (c == '[') {
            v.kind = Json::Array; ++p_; ws();
            if (p_ < s_.size() && s_[p_] == ']') { ++p_; return v; }
            for (;;) {
                v.arr.push_back(value());
                ws();
                if (p_ < s_.size() && s_[p_] == ',') { ++p_; continue; }
                if (p_ < s_.size() && s_[p_] == ']') { ++p_; break; }
                fail("expected ',' or ']'");
            }
        } else if (c == '"') {
            v.kind = Json::String; v.str = string_lit();
        } else if (lit("true")) { v.kind = Json::Bool; v.b = true; }
        else if (lit("false")) { v.kind = Json::Bool; v.b = false; }
        else if (lit("null")) { v.kind = Json::Null; }
        else if (c == '-' || c == '+' || c == '.' || std::isdigit((unsigned char)c)) {
            const char* b = s_.c_str() + p_;
            char* e = nullptr;
            v.num = std::strtod(b, &e);
            if (e == b) fail("bad number");
            v.kind = Json::Number;
            p_ += (size_t)(e - b);
        } else {
            fail("unexpected character");
        }
        return v;
    }
    std::string string_lit() {
        ++p_;  // opening quote
        std::string out;
        while (p_ < s_.size() && s_[p_] != '"') {
            char c = s_[p_++];
            if (c == '\\' && p_ < s_.size()) {
                char e = s_[p_++];
                switch (e) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'u': p_ += 4; out += '?'; break;  // config strings are ASCII
                    default: out += e;
                }
            } else {
                out += c;
            }
        }
        if (p_ >= s_.size()) fail("unterminated string");
        ++p_;
        return out;
    }
};

inline Json parse_json(const std::string& s) { return JsonParser(s).parse(); }

}  // namespace dfg
