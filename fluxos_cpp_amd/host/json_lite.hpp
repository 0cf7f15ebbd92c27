// json_lite.hpp -- a small JSON reader for the modset master file.  The
// reference parses it with comments allowed (main.cpp:103-109), so // and
// /* */ comments are skipped.  Values: null, bool, number, string, array,
// object (keys kept in file order).
#pragma once

#include <cstdlib>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace fxjson {

struct Value {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Value> arr;
    std::vector<std::pair<std::string, Value>> obj;

    const Value *find(const std::string &key) const
    {
        if (kind != Object) return nullptr;
        for (auto &kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
    // lookup by path of keys; nullptr when any step is missing
    const Value *at(std::initializer_list<const char *> path) const
    {
        const Value *v = this;
        for (const char *k : path) {
            v = v->find(k);
            if (!v) return nullptr;
        }
        return v;
    }
};

class Parser {
public:
    explicit Parser(const std::string &text) : s_(text) {}

    Value parse()
    {
        Value v = value();
        ws();
        if (i_ != s_.size()) fail("trailing characters");
        return v;
    }

private:
    const std::string &s_;
    size_t i_ = 0;

    [[noreturn]] void fail(const char *what)
    {
        size_t line = 1;
        for (size_t k = 0; k < i_ && k < s_.size(); k++) line += s_[k] == '\n';
        throw std::runtime_error(std::string("modset JSON: ") + what + " at line " + std::to_string(line));
    }

    void ws()
    {
        for (;;) {
            while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\t' || s_[i_] == '\n' || s_[i_] == '\r')) i_++;
            if (i_ + 1 < s_.size() && s_[i_] == '/' && s_[i_ + 1] == '/') {
                while (i_ < s_.size() && s_[i_] != '\n') i_++;
            } else if (i_ + 1 < s_.size() && s_[i_] == '/' && s_[i_ + 1] == '*') {
                const size_t e = s_.find("*/", i_ + 2);
                if (e == std::string::npos) fail("unterminated comment");
                i_ = e + 2;
            } else {
                return;
            }
        }
    }

    bool lit(const char *w)
    {
        const size_t n = std::char_traits<char>::length(w);
        if (s_.compare(i_, n, w) == 0) { i_ += n; return true; }
        return false;
    }

    std::string string_()
    {
        if (s_[i_] != '"') fail("expected a string");
        i_++;
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
                    const unsigned cp = (unsigned)std::strtoul(s_.substr(i_, 4).c_str(), nullptr, 16);
                    i_ += 4;
                    if (cp < 0x80) out += (char)cp;
                    else if (cp < 0x800) { out += (char)(0xC0 | (cp >> 6)); out += (char)(0x80 | (cp & 0x3F)); }
                    else { out += (char)(0xE0 | (cp >> 12)); out += (char)(0x80 | ((cp >> 6) & 0x3F)); out += (char)(0x80 | (cp & 0x3F)); }
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

    Value value()
    {
        ws();
        if (i_ >= s_.size()) fail("unexpected end");
        Value v;
        const char c = s_[i_];
        if (c == '{') {
            v.kind = Value::Object;
            i_++;
            ws();
            if (s_[i_] == '}') { i_++; return v; }
            for (;;) {
                ws();
                std::string k = string_();
                ws();
                if (s_[i_] != ':') fail("expected ':'");
                i_++;
                v.obj.emplace_back(std::move(k), value());
                ws();
                if (s_[i_] == ',') { i_++; continue; }
                if (s_[i_] == '}') { i_++; return v; }
                fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.kind = Value::Array;
            i_++;
            ws();
            if (s_[i_] == ']') { i_++; return v; }
            for (;;) {
                v.arr.push_back(value());
                ws();
                if (s_[i_] == ',') { i_++; continue; }
                if (s_[i_] == ']') { i_++; return v; }
                fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.kind = Value::String;
            v.str = string_();
            return v;
        }
        if (lit("true")) { v.kind = Value::Bool; v.b = true; return v; }
        if (lit("false")) { v.kind = Value::Bool; v.b = false; return v; }
        if (lit("null")) return v;
        char *end = nullptr;
        v.num = std::strtod(s_.c_str() + i_, &end);
        if (end == s_.c_str() + i_) fail("unexpected character");
        i_ = (size_t)(end - s_.c_str());
        v.kind = Value::Number;
        return v;
    }
};

inline Value parse(const std::string &text) { return Parser(text).parse(); }

} // namespace fxjson
