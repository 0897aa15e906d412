#include "json.h"

#include <cstdlib>
#include <cstring>

namespace rt {
namespace {

const Json& null_json() {
    static const Json n;
    return n;
}

struct Parser {
    const std::string& s;
    size_t p = 0;

    [[noreturn]] void fail(const char* msg) {
        throw JsonError(std::string("JSON parse error at byte ") + std::to_string(p) + ": " + msg);
    }
    void ws() {
        while (p < s.size() && (s[p] == ' ' || s[p] == '\t' || s[p] == '\n' || s[p] == '\r')) ++p;
    }
    bool lit(const char* w) {
        size_t n = std::strlen(w);
        if (s.compare(p, n, w) == 0) { p += n; return true; }
        return false;
    }
    Json value() {
        ws();
        if (p >= s.size()) fail("unexpected end of input");
        char c = s[p];
        Json v;
        if (c == '{') {
            v.kind = Json::Object;
            ++p;
            ws();
            if (p < s.size() && s[p] == '}') { ++p; return v; }
            while (true) {
                ws();
                if (p >= s.size() || s[p] != '"') fail("expected object key");
                std::string k = string();
                ws();
                if (p >= s.size() || s[p] != ':') fail("expected ':'");
                ++p;
                Json item = value();
                bool replaced = false;
                for (auto& kv : v.obj)
                    if (kv.first == k) { kv.second = std::move(item); replaced = true; break; }
                if (!replaced) v.obj.emplace_back(std::move(k), std::move(item));
                ws();
                if (p < s.size() && s[p] == ',') { ++p; continue; }
                if (p < s.size() && s[p] == '}') { ++p; break; }
                fail("expected ',' or '}'");
            }
        } else if (c == '[') {
            v.kind = Json::Array;
            ++p;
            ws();
            if (p < s.size() && s[p] == ']') { ++p; return v; }
            while (true) {
                v.arr.push_back(value());
                ws();
                if (p < s.size() && s[p] == ',') { ++p; continue; }
                if (p < s.size() && s[p] == ']') { ++p; break; }
                fail("expected ',' or ']'");
            }
        } else if (c == '"') {
            v.kind = Json::String;
            v.str = string();
        } else if (lit("true")) {
            v.kind = Json::Bool;
            v.b = true;
        } else if (lit("false")) {
            v.kind = Json::Bool;
            v.b = false;
        } else if (lit("null")) {
            v.kind = Json::Null;
        } else if (c == '-' || (c >= '0' && c <= '9')) {
            const char* begin = s.c_str() + p;
            char* end = nullptr;
            v.kind = Json::Number;
            v.num = std::strtod(begin, &end);
            if (end == begin) fail("bad number");
            p += (size_t)(end - begin);
        } else {
            fail("unexpected character");
        }
        return v;
    }
    void put_utf8(std::string& out, unsigned cp) {
        if (cp < 0x80) out += (char)cp;
        else if (cp < 0x800) { out += (char)(0xC0 | (cp >> 6)); out += (char)(0x80 | (cp & 0x3F)); }
        else if (cp < 0x10000) {
            out += (char)(0xE0 | (cp >> 12)); out += (char)(0x80 | ((cp >> 6) & 0x3F)); out += (char)(0x80 | (cp & 0x3F));
        } else {
            out += (char)(0xF0 | (cp >> 18)); out += (char)(0x80 | ((cp >> 12) & 0x3F));
            out += (char)(0x80 | ((cp >> 6) & 0x3F)); out += (char)(0x80 | (cp & 0x3F));
        }
    }
    unsigned hex4() {
        if (p + 4 > s.size()) fail("bad \\u escape");
        unsigned v = 0;
        for (int i = 0; i < 4; ++i) {
            char h = s[p++];
            v <<= 4;
            if (h >= '0' && h <= '9') v |= (unsigned)(h - '0');
            else if (h >= 'a' && h <= 'f') v |= (unsigned)(h - 'a' + 10);
            else if (h >= 'A' && h <= 'F') v |= (unsigned)(h - 'A' + 10);
            else fail("bad hex digit");
        }
        return v;
    }
    std::string string() {
        ++p;  // opening quote
        std::string out;
        while (true) {
            if (p >= s.size()) fail("unterminated string");
            char c = s[p++];
            if (c == '"') break;
            if (c != '\\') { out += c; continue; }
            if (p >= s.size()) fail("bad escape");
            char e = s[p++];
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
                    unsigned cp = hex4();
                    if (cp >= 0xD800 && cp < 0xDC00 && p + 6 <= s.size() && s[p] == '\\' && s[p + 1] == 'u') {
                        p += 2;
                        unsigned lo = hex4();
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    put_utf8(out, cp);
                    break;
                }
                default: fail("bad escape");
            }
        }
        return out;
    }
};

}  // namespace

Json Json::parse(const std::string& text) {
    Parser ps{text};
    // A leading UTF-8 byte-order mark is skipped, as nlohmann's lexer does
    // (reference: ThirdParty/Json/json.hpp:8858-8884); some of the reference's
    // scene files start with one.
    if (text.compare(0, 3, "\xEF\xBB\xBF") == 0) ps.p = 3;
    Json v = ps.value();
    ps.ws();
    if (ps.p != text.size()) ps.fail("trailing characters");
    return v;
}

bool Json::contains(const std::string& key) const {
    if (kind != Object) return false;
    for (const auto& kv : obj)
        if (kv.first == key) return true;
    return false;
}

const Json& Json::operator[](const std::string& key) const {
    if (kind == Object)
        for (const auto& kv : obj)
            if (kv.first == key) return kv.second;
    return null_json();
}

const Json& Json::operator[](size_t i) const {
    if (kind == Array && i < arr.size()) return arr[i];
    return null_json();
}

double Json::as_number(const char* what) const {
    if (kind == Number) return num;
    if (kind == Bool) return b ? 1.0 : 0.0;
    throw JsonError(std::string("expected a number for '") + what + "'");
}

int Json::as_int(const char* what) const { return (int)as_number(what); }

bool Json::as_bool(const char* what) const {
    if (kind == Bool) return b;
    if (kind == Number) return num != 0.0;
    throw JsonError(std::string("expected a boolean for '") + what + "'");
}

const std::string& Json::as_string(const char* what) const {
    if (kind == String) return str;
    throw JsonError(std::string("expected a string for '") + what + "'");
}

}  // namespace rt
