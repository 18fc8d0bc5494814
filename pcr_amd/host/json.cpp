// json.cpp -- JSON-with-comments reader (see json.h).
#include "json.h"

#include <cerrno>
#include <cstdlib>
#include <cstring>

namespace pcr::json {

namespace {

struct Parser {
    const std::string &s;
    size_t i = 0;
    explicit Parser(const std::string &t) : s(t) {}

    [[noreturn]] void error(const std::string &msg) const {
        int line = 1, col = 1;
        for (size_t k = 0; k < i && k < s.size(); k++) {
            if (s[k] == '\n') { line++; col = 1; }
            else col++;
        }
        throw ParseError(msg, line, col);
    }
    void skip() {
        for (;;) {
            while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r')) i++;
            if (i + 1 < s.size() && s[i] == '/' && s[i + 1] == '/') {
                while (i < s.size() && s[i] != '\n') i++;
                continue;
            }
            if (i + 1 < s.size() && s[i] == '/' && s[i + 1] == '*') {
                size_t e = s.find("*/", i + 2);
                if (e == std::string::npos) error("unterminated comment");
                i = e + 2;
                continue;
            }
            return;
        }
    }
    bool lit(const char *w) {
        size_t n = std::strlen(w);
        if (s.compare(i, n, w) == 0) {
            i += n;
            return true;
        }
        return false;
    }
    std::string string_body() {
        // at the opening quote
        i++;
        std::string out;
        while (i < s.size() && s[i] != '"') {
            char c = s[i++];
            if (c == '\\') {
                if (i >= s.size()) error("bad escape");
                char e = s[i++];
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
                    if (i + 4 > s.size()) error("bad \\u escape");
                    unsigned cp = (unsigned)std::strtoul(s.substr(i, 4).c_str(), nullptr, 16);
                    i += 4;
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
                default: error("bad escape");
                }
            } else {
                out += c;
            }
        }
        if (i >= s.size()) error("unterminated string");
        i++;
        return out;
    }
    Value value() {
        skip();
        if (i >= s.size()) error("unexpected end of input");
        Value v;
        char c = s[i];
        if (c == '{') {
            v.type = Value::Object;
            i++;
            skip();
            if (i < s.size() && s[i] == '}') {
                i++;
                return v;
            }
            for (;;) {
                skip();
                if (i >= s.size() || s[i] != '"') error("expected object key");
                std::string k = string_body();
                skip();
                if (i >= s.size() || s[i] != ':') error("expected ':'");
                i++;
                v.obj[k] = value();
                skip();
                if (i < s.size() && s[i] == ',') { i++; continue; }
                if (i < s.size() && s[i] == '}') { i++; return v; }
                error("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.type = Value::Array;
            i++;
            skip();
            if (i < s.size() && s[i] == ']') {
                i++;
                return v;
            }
            for (;;) {
                v.arr.push_back(value());
                skip();
                if (i < s.size() && s[i] == ',') { i++; continue; }
                if (i < s.size() && s[i] == ']') { i++; return v; }
                error("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.type = Value::String;
            v.str = string_body();
            return v;
        }
        if (lit("true")) { v.type = Value::Bool; v.b = true; return v; }
        if (lit("false")) { v.type = Value::Bool; v.b = false; return v; }
        if (lit("null")) return v;
        if (c == '-' || (c >= '0' && c <= '9')) {
            size_t st = i;
            if (s[i] == '-') i++;
            while (i < s.size() && s[i] >= '0' && s[i] <= '9') i++;
            bool integer = true;
            if (i < s.size() && s[i] == '.') {
                integer = false;
                i++;
                while (i < s.size() && s[i] >= '0' && s[i] <= '9') i++;
            }
            if (i < s.size() && (s[i] == 'e' || s[i] == 'E')) {
                integer = false;
                i++;
                if (i < s.size() && (s[i] == '+' || s[i] == '-')) i++;
                while (i < s.size() && s[i] >= '0' && s[i] <= '9') i++;
            }
            std::string tok = s.substr(st, i - st);
            v.type = Value::Number;
            v.is_int = integer;
            if (integer) {
                errno = 0;
                long long iv = std::strtoll(tok.c_str(), nullptr, 10);
                v.num = (errno == 0) ? (double)iv : std::strtod(tok.c_str(), nullptr);
            } else {
                v.num = std::strtod(tok.c_str(), nullptr);
            }
            return v;
        }
        error(std::string("unexpected character '") + c + "'");
    }
};

}  // namespace

Value parse(const std::string &text) {
    Parser p(text);
    Value v = p.value();
    p.skip();
    if (p.i != text.size()) p.error("trailing characters after the document");
    return v;
}

}  // namespace pcr::json
