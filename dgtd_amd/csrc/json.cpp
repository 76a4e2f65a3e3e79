// Recursive-descent JSON reader (see json.hpp).
#include "json.hpp"

#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace dgtd {

namespace {

struct Parser {
    const std::string& s;
    size_t i = 0;
    int line = 1;

    [[noreturn]] void fail(const std::string& what) const
    {
        throw std::runtime_error("JSON parse error at line " + std::to_string(line) + ": " + what);
    }
    void ws()
    {
        while (i < s.size()) {
            const char c = s[i];
            if (c == '\n') ++line;
            if (c == ' ' || c == '\t' || c == '\r' || c == '\n')
                ++i;
            else
                break;
        }
    }
    char peek()
    {
        ws();
        if (i >= s.size()) fail("unexpected end of input");
        return s[i];
    }
    void expect(char c)
    {
        if (peek() != c) fail(std::string("expected '") + c + "'");
        ++i;
    }
    bool literal(const char* w)
    {
        const size_t n = std::char_traits<char>::length(w);
        if (s.compare(i, n, w) == 0) {
            i += n;
            return true;
        }
        return false;
    }
    static void put_utf8(std::string& out, unsigned cp)
    {
        if (cp < 0x80) {
            out += (char)cp;
        } else if (cp < 0x800) {
            out += (char)(0xC0 | (cp >> 6));
            out += (char)(0x80 | (cp & 0x3F));
        } else {
            out += (char)(0xE0 | (cp >> 12));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        }
    }
    std::string string_body()
    {
        expect('"');
        std::string out;
        while (true) {
            if (i >= s.size()) fail("unterminated string");
            const char c = s[i++];
            if (c == '"') return out;
            if (c != '\\') {
                out += c;
                continue;
            }
            if (i >= s.size()) fail("unterminated escape");
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
                    if (i + 4 > s.size()) fail("short \\u escape");
                    put_utf8(out, (unsigned)std::strtoul(s.substr(i, 4).c_str(), nullptr, 16));
                    i += 4;
                    break;
                }
                default: fail("bad escape");
            }
        }
    }
    Json value()
    {
        Json v;
        const char c = peek();
        if (c == '{') {
            ++i;
            v.kind = Json::kObject;
            if (peek() == '}') {
                ++i;
                return v;
            }
            while (true) {
                std::string k = string_body();
                expect(':');
                v.obj.emplace_back(std::move(k), value());
                const char d = peek();
                ++i;
                if (d == '}') return v;
                if (d != ',') fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            ++i;
            v.kind = Json::kArray;
            if (peek() == ']') {
                ++i;
                return v;
            }
            while (true) {
                v.arr.push_back(value());
                const char d = peek();
                ++i;
                if (d == ']') return v;
                if (d != ',') fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.kind = Json::kString;
            v.str = string_body();
            return v;
        }
        if (literal("true")) {
            v.kind = Json::kBool;
            v.b = true;
            return v;
        }
        if (literal("false")) {
            v.kind = Json::kBool;
            return v;
        }
        if (literal("null")) return v;
        // number
        const char* start = s.c_str() + i;
        char* end = nullptr;
        v.num = std::strtod(start, &end);
        if (end == start) fail("unexpected character");
        i += (size_t)(end - start);
        v.kind = Json::kNumber;
        return v;
    }
};

}  // namespace

bool Json::contains(const std::string& k) const
{
    if (kind != kObject) return false;
    for (const auto& kv : obj)
        if (kv.first == k) return true;
    return false;
}

const Json& Json::at(const std::string& k) const
{
    if (kind == kObject)
        for (const auto& kv : obj)
            if (kv.first == k) return kv.second;
    throw std::runtime_error("JSON: missing key '" + k + "'");
}

const Json& Json::at(size_t idx) const
{
    if (kind != kArray || idx >= arr.size()) throw std::runtime_error("JSON: array index out of range");
    return arr[idx];
}

double Json::number() const
{
    if (kind == kNumber) return num;
    if (kind == kBool) return b ? 1.0 : 0.0;
    throw std::runtime_error("JSON: value is not a number");
}

int Json::integer() const
{
    const double v = number();
    if (v != std::floor(v)) throw std::runtime_error("JSON: value is not an integer");
    return (int)v;
}

const std::string& Json::string() const
{
    if (kind != kString) throw std::runtime_error("JSON: value is not a string");
    return str;
}

std::vector<double> Json::numbers() const
{
    std::vector<double> out;
    if (kind == kNumber) {
        out.push_back(num);
        return out;
    }
    if (kind != kArray) throw std::runtime_error("JSON: value is not an array of numbers");
    for (const auto& v : arr) out.push_back(v.number());
    return out;
}

Json Json::parse(const std::string& text)
{
    Parser p{text};
    Json v = p.value();
    p.ws();
    if (p.i != text.size()) p.fail("trailing characters");
    return v;
}

Json Json::parse_file(const std::string& path)
{
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return parse(ss.str());
}

}  // namespace dgtd
