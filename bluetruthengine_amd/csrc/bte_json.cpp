// bte_json.cpp — recursive-descent JSON reader (RFC 8259 grammar: objects,
// arrays, strings with escapes, numbers, true / false / null).
#include "bte_json.h"

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace bte {

const JValue* JValue::get(const char* key) const {
    if (type != Object) return nullptr;
    for (const auto& kv : obj)
        if (kv.first == key) return &kv.second;
    return nullptr;
}

double JValue::number(const char* key, double def) const {
    const JValue* v = get(key);
    return (v && v->type == Number) ? v->num : def;
}

int JValue::integer(const char* key, int def) const {
    const JValue* v = get(key);
    // non-numbers and values outside int (or not finite) read as the default
    if (!v || v->type != Number || !(v->num >= -2147483648.0 && v->num <= 2147483647.0)) return def;
    return (int)v->num;
}

std::string JValue::string(const char* key) const {
    const JValue* v = get(key);
    return (v && v->type == String) ? v->str : std::string();
}

std::vector<double> JValue::numbers(const char* key) const {
    std::vector<double> out;
    const JValue* v = get(key);
    if (!v || v->type != Array) return out;
    for (const JValue& e : v->arr) out.push_back(e.type == Number ? e.num : 0.0);
    return out;
}

namespace {

struct Parser {
    const char* s;
    size_t n, i = 0;
    std::string err;
    int depth = 0;
    Parser(const char* text, size_t size) : s(text), n(size) {}

    bool fail(const char* what) {
        char buf[160];
        std::snprintf(buf, sizeof buf, "JSON: %s at byte %zu", what, i);
        err = buf;
        return false;
    }
    void ws() {
        while (i < n && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r')) ++i;
    }
    bool literal(const char* word) {
        const size_t k = std::strlen(word);
        if (n - i < k || std::memcmp(s + i, word, k) != 0) return fail("invalid literal");
        i += k;
        return true;
    }
    static void utf8(uint32_t cp, std::string& out) {
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
    bool hex4(uint32_t& v) {
        if (n - i < 4) return fail("truncated \\u escape");
        v = 0;
        for (int k = 0; k < 4; ++k) {
            const char c = s[i++];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (uint32_t)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (uint32_t)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (uint32_t)(c - 'A' + 10);
            else return fail("bad \\u escape");
        }
        return true;
    }
    bool string(std::string& out) {
        ++i;   // opening quote
        out.clear();
        while (i < n) {
            const char c = s[i++];
            if (c == '"') return true;
            if ((unsigned char)c < 0x20) return fail("control character in string");
            if (c != '\\') {
                out += c;
                continue;
            }
            if (i >= n) break;
            const char e = s[i++];
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
                    uint32_t cp;
                    if (!hex4(cp)) return false;
                    if (cp >= 0xD800 && cp < 0xDC00 && n - i >= 6 && s[i] == '\\' && s[i + 1] == 'u') {
                        i += 2;
                        uint32_t lo;
                        if (!hex4(lo)) return false;
                        if (lo >= 0xDC00 && lo < 0xE000) cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                        else { utf8(cp, out); cp = lo; }
                    }
                    utf8(cp, out);
                    break;
                }
                default: return fail("bad escape");
            }
        }
        return fail("unterminated string");
    }
    bool number(double& v) {
        const size_t start = i;
        if (i < n && s[i] == '-') ++i;
        if (i >= n || !(s[i] >= '0' && s[i] <= '9')) return fail("bad number");
        if (s[i] == '0') ++i;
        else while (i < n && s[i] >= '0' && s[i] <= '9') ++i;
        if (i < n && s[i] == '.') {
            ++i;
            if (i >= n || !(s[i] >= '0' && s[i] <= '9')) return fail("bad number fraction");
            while (i < n && s[i] >= '0' && s[i] <= '9') ++i;
        }
        if (i < n && (s[i] == 'e' || s[i] == 'E')) {
            ++i;
            if (i < n && (s[i] == '+' || s[i] == '-')) ++i;
            if (i >= n || !(s[i] >= '0' && s[i] <= '9')) return fail("bad number exponent");
            while (i < n && s[i] >= '0' && s[i] <= '9') ++i;
        }
        const std::string tok(s + start, i - start);
        v = std::strtod(tok.c_str(), nullptr);
        return true;
    }
    bool value(JValue& v) {
        if (++depth > 512) return fail("nesting too deep");
        ws();
        if (i >= n) return fail("unexpected end");
        bool ok = true;
        const char c = s[i];
        if (c == '{') {
            v.type = JValue::Object;
            ++i;
            ws();
            if (i < n && s[i] == '}') {
                ++i;
            } else {
                for (;;) {
                    ws();
                    if (i >= n || s[i] != '"') { ok = fail("expected a member name"); break; }
                    std::string key;
                    if (!(ok = string(key))) break;
                    ws();
                    if (i >= n || s[i] != ':') { ok = fail("expected ':'"); break; }
                    ++i;
                    v.obj.emplace_back(std::move(key), JValue());
                    if (!(ok = value(v.obj.back().second))) break;
                    ws();
                    if (i < n && s[i] == ',') { ++i; continue; }
                    if (i < n && s[i] == '}') { ++i; break; }
                    ok = fail("expected ',' or '}'");
                    break;
                }
            }
        } else if (c == '[') {
            v.type = JValue::Array;
            ++i;
            ws();
            if (i < n && s[i] == ']') {
                ++i;
            } else {
                for (;;) {
                    v.arr.emplace_back();
                    if (!(ok = value(v.arr.back()))) break;
                    ws();
                    if (i < n && s[i] == ',') { ++i; continue; }
                    if (i < n && s[i] == ']') { ++i; break; }
                    ok = fail("expected ',' or ']'");
                    break;
                }
            }
        } else if (c == '"') {
            v.type = JValue::String;
            ok = string(v.str);
        } else if (c == 't') {
            v.type = JValue::Bool;
            v.b = true;
            ok = literal("true");
        } else if (c == 'f') {
            v.type = JValue::Bool;
            ok = literal("false");
        } else if (c == 'n') {
            ok = literal("null");
        } else {
            v.type = JValue::Number;
            ok = number(v.num);
        }
        --depth;
        return ok;
    }
};

}  // namespace

bool parse_json(const char* text, size_t size, JValue& out, std::string& err) {
    Parser p(text, size);
    out = JValue();
    // a UTF-8 byte-order mark is tolerated
    if (size >= 3 && (unsigned char)text[0] == 0xEF && (unsigned char)text[1] == 0xBB && (unsigned char)text[2] == 0xBF)
        p.i = 3;
    if (!p.value(out)) {
        err = p.err;
        return false;
    }
    p.ws();
    if (p.i != size) {
        p.fail("trailing characters");
        err = p.err;
        return false;
    }
    return true;
}

}  // namespace bte
