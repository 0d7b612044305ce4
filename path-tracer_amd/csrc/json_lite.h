// json_lite.h - a small DOM JSON reader for the glTF front end (RFC 8259:
// objects, arrays, strings with escapes incl. surrogate pairs, numbers
// parsed with strtod, true/false/null). Objects keep member order.
#pragma once

#include <cstdint>
#include <cstdlib>
#include <string>
#include <utility>
#include <vector>

namespace ptcore {

struct Json {
    enum Type { NUL, BOOL, NUM, STR, ARR, OBJ } type = NUL;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Json> arr;
    std::vector<std::pair<std::string, Json>> obj;

    const Json* get(const char* key) const {
        if (type != OBJ) return nullptr;
        for (const auto& kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
    bool is_num() const { return type == NUM; }
    double num_or(double d) const { return type == NUM ? num : d; }
    // member `key` as a number, or `d` when absent / not a number
    double num_at(const char* key, double d) const {
        const Json* v = get(key);
        return v ? v->num_or(d) : d;
    }
    // member `key` as a non-negative integer index, or -1
    long index_at(const char* key) const {
        const Json* v = get(key);
        if (!v || v->type != NUM || v->num < 0 || v->num != (double)(long)v->num) return -1;
        return (long)v->num;
    }
};

class JsonParser {
  public:
    JsonParser(const char* s, size_t n) : p_(s), end_(s + n) {}
    bool parse(Json& out, std::string& err) {
        if (!value(out, 0)) {
            err = "JSON: " + err_;
            return false;
        }
        ws();
        if (p_ != end_) {
            err = "JSON: trailing characters";
            return false;
        }
        return true;
    }

  private:
    const char* p_;
    const char* end_;
    std::string err_;

    bool fail(const char* m) {
        err_ = m;
        return false;
    }
    void ws() {
        while (p_ < end_ && (*p_ == ' ' || *p_ == '\t' || *p_ == '\n' || *p_ == '\r')) ++p_;
    }
    bool lit(const char* w) {
        size_t n = 0;
        while (w[n]) ++n;
        if ((size_t)(end_ - p_) < n) return false;
        for (size_t i = 0; i < n; ++i)
            if (p_[i] != w[i]) return false;
        p_ += n;
        return true;
    }
    static void utf8(std::string& s, uint32_t cp) {
        if (cp < 0x80) s.push_back((char)cp);
        else if (cp < 0x800) {
            s.push_back((char)(0xc0 | (cp >> 6)));
            s.push_back((char)(0x80 | (cp & 63)));
        } else if (cp < 0x10000) {
            s.push_back((char)(0xe0 | (cp >> 12)));
            s.push_back((char)(0x80 | ((cp >> 6) & 63)));
            s.push_back((char)(0x80 | (cp & 63)));
        } else {
            s.push_back((char)(0xf0 | (cp >> 18)));
            s.push_back((char)(0x80 | ((cp >> 12) & 63)));
            s.push_back((char)(0x80 | ((cp >> 6) & 63)));
            s.push_back((char)(0x80 | (cp & 63)));
        }
    }
    bool hex4(uint32_t& v) {
        if (end_ - p_ < 4) return false;
        v = 0;
        for (int i = 0; i < 4; ++i) {
            char c = *p_++;
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (uint32_t)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (uint32_t)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (uint32_t)(c - 'A' + 10);
            else return false;
        }
        return true;
    }
    bool string(std::string& s) {
        ++p_;  // opening quote
        s.clear();
        while (p_ < end_ && *p_ != '"') {
            char c = *p_++;
            if ((unsigned char)c < 0x20) return fail("control character in string");
            if (c != '\\') {
                s.push_back(c);
                continue;
            }
            if (p_ >= end_) return fail("bad escape");
            char e = *p_++;
            switch (e) {
                case '"': s.push_back('"'); break;
                case '\\': s.push_back('\\'); break;
                case '/': s.push_back('/'); break;
                case 'b': s.push_back('\b'); break;
                case 'f': s.push_back('\f'); break;
                case 'n': s.push_back('\n'); break;
                case 'r': s.push_back('\r'); break;
                case 't': s.push_back('\t'); break;
                case 'u': {
                    uint32_t cp;
                    if (!hex4(cp)) return fail("bad \\u escape");
                    if (cp >= 0xd800 && cp < 0xdc00) {
                        uint32_t lo;
                        if (!lit("\\u") || !hex4(lo) || lo < 0xdc00 || lo >= 0xe000) return fail("bad surrogate pair");
                        cp = 0x10000 + ((cp - 0xd800) << 10) + (lo - 0xdc00);
                    }
                    utf8(s, cp);
                    break;
                }
                default: return fail("bad escape");
            }
        }
        if (p_ >= end_) return fail("unterminated string");
        ++p_;
        return true;
    }
    bool number(Json& v) {
        const char* b = p_;
        if (p_ < end_ && *p_ == '-') ++p_;
        while (p_ < end_ && ((*p_ >= '0' && *p_ <= '9') || *p_ == '.' || *p_ == 'e' || *p_ == 'E' || *p_ == '+' ||
                             *p_ == '-'))
            ++p_;
        std::string t(b, p_);
        char* e = nullptr;
        v.num = std::strtod(t.c_str(), &e);
        if (t.empty() || e != t.c_str() + t.size()) return fail("bad number");
        v.type = Json::NUM;
        return true;
    }
    bool value(Json& v, int depth) {
        if (depth > 512) return fail("nesting too deep");
        ws();
        if (p_ >= end_) return fail("unexpected end");
        char c = *p_;
        if (c == '{') {
            ++p_;
            v.type = Json::OBJ;
            ws();
            if (p_ < end_ && *p_ == '}') {
                ++p_;
                return true;
            }
            for (;;) {
                ws();
                if (p_ >= end_ || *p_ != '"') return fail("expected member name");
                std::string k;
                if (!string(k)) return false;
                ws();
                if (p_ >= end_ || *p_ != ':') return fail("expected ':'");
                ++p_;
                v.obj.emplace_back(std::move(k), Json());
                if (!value(v.obj.back().second, depth + 1)) return false;
                ws();
                if (p_ < end_ && *p_ == ',') {
                    ++p_;
                    continue;
                }
                if (p_ < end_ && *p_ == '}') {
                    ++p_;
                    return true;
                }
                return fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            ++p_;
            v.type = Json::ARR;
            ws();
            if (p_ < end_ && *p_ == ']') {
                ++p_;
                return true;
            }
            for (;;) {
                v.arr.emplace_back();
                if (!value(v.arr.back(), depth + 1)) return false;
                ws();
                if (p_ < end_ && *p_ == ',') {
                    ++p_;
                    continue;
                }
                if (p_ < end_ && *p_ == ']') {
                    ++p_;
                    return true;
                }
                return fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.type = Json::STR;
            return string(v.str);
        }
        if (lit("true")) {
            v.type = Json::BOOL;
            v.b = true;
            return true;
        }
        if (lit("false")) {
            v.type = Json::BOOL;
            return true;
        }
        if (lit("null")) return true;
        return number(v);
    }
};

}  // namespace ptcore
