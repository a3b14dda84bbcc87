// json.cpp -- reader and writer for fpsjson::Value (see json.hpp).
#include "../../include/fps_json.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace fpsjson {

namespace {
const Value& nullValue()
{
    static const Value v;
    return v;
}
}  // namespace

double Value::asDouble() const
{
    switch (type_) {
    case Int: return static_cast<double>(i_);
    case Real: return d_;
    case Bool: return b_ ? 1.0 : 0.0;
    case String: return std::strtod(s_.c_str(), nullptr);
    default: return 0.0;
    }
}

int64_t Value::asInt() const
{
    switch (type_) {
    case Int: return i_;
    case Real: return static_cast<int64_t>(d_);
    case Bool: return b_ ? 1 : 0;
    case String: return std::strtoll(s_.c_str(), nullptr, 10);
    default: return 0;
    }
}

bool Value::asBool() const
{
    switch (type_) {
    case Bool: return b_;
    case Int: return i_ != 0;
    case Real: return d_ != 0.0;
    case String: return s_ == "true";
    default: return false;
    }
}

std::string Value::asString() const
{
    switch (type_) {
    case String: return s_;
    case Bool: return b_ ? "true" : "false";
    case Int: return std::to_string(i_);
    case Real: {
        char buf[64];
        std::snprintf(buf, sizeof buf, "%.17g", d_);
        return buf;
    }
    default: return "";
    }
}

size_t Value::size() const
{
    if (type_ == Array) return arr_.size();
    if (type_ == Object) return obj_.size();
    return 0;
}

bool Value::isMember(const std::string& k) const { return type_ == Object && obj_.count(k) != 0; }

const Value& Value::operator[](const std::string& k) const
{
    if (type_ != Object) return nullValue();
    auto it = obj_.find(k);
    return it == obj_.end() ? nullValue() : it->second;
}

const Value& Value::operator[](size_t i) const
{
    if (type_ != Array || i >= arr_.size()) return nullValue();
    return arr_[i];
}

Value& Value::operator[](const std::string& k)
{
    if (type_ == Null) type_ = Object;
    return obj_[k];
}

Value& Value::operator[](size_t i)
{
    if (type_ == Null) type_ = Array;
    if (i >= arr_.size()) arr_.resize(i + 1);
    return arr_[i];
}

Value Value::get(const std::string& k, const Value& dflt) const
{
    if (type_ != Object) return dflt;
    auto it = obj_.find(k);
    return it == obj_.end() ? dflt : it->second;
}

void Value::append(Value v)
{
    if (type_ == Null) type_ = Array;
    arr_.push_back(std::move(v));
}

bool Value::operator==(const Value& o) const
{
    if (isNumber() && o.isNumber()) return asDouble() == o.asDouble();
    if (type_ != o.type_) return false;
    switch (type_) {
    case Null: return true;
    case Bool: return b_ == o.b_;
    case String: return s_ == o.s_;
    case Array: return arr_ == o.arr_;
    case Object: return obj_ == o.obj_;
    default: return false;
    }
}

// ---- reader -------------------------------------------------------------------------
namespace {

struct Reader {
    const std::string& t;
    size_t p = 0;
    std::string err;

    explicit Reader(const std::string& text) : t(text) {}

    bool fail(const std::string& why)
    {
        size_t line = 1, col = 1;
        for (size_t i = 0; i < p && i < t.size(); ++i) {
            if (t[i] == '\n') { ++line; col = 1; } else { ++col; }
        }
        err = std::to_string(line) + ":" + std::to_string(col) + ": " + why;
        return false;
    }

    void ws()
    {
        for (;;) {
            while (p < t.size() && (t[p] == ' ' || t[p] == '\t' || t[p] == '\n' || t[p] == '\r')) ++p;
            // jsoncpp accepts comments (fps.cpp:139 sets collectComments)
            if (p + 1 < t.size() && t[p] == '/' && t[p + 1] == '/') {
                while (p < t.size() && t[p] != '\n') ++p;
                continue;
            }
            if (p + 1 < t.size() && t[p] == '/' && t[p + 1] == '*') {
                size_t e = t.find("*/", p + 2);
                p = (e == std::string::npos) ? t.size() : e + 2;
                continue;
            }
            break;
        }
    }

    bool value(Value& out)
    {
        ws();
        if (p >= t.size()) return fail("unexpected end of input");
        char c = t[p];
        if (c == '{') return object(out);
        if (c == '[') return array(out);
        if (c == '"') {
            std::string s;
            if (!str(s)) return false;
            out = Value(std::move(s));
            return true;
        }
        if (t.compare(p, 4, "true") == 0) { p += 4; out = Value(true); return true; }
        if (t.compare(p, 5, "false") == 0) { p += 5; out = Value(false); return true; }
        if (t.compare(p, 4, "null") == 0) { p += 4; out = Value(); return true; }
        return number(out);
    }

    bool number(Value& out)
    {
        size_t s = p;
        if (p < t.size() && (t[p] == '-' || t[p] == '+')) ++p;
        bool isReal = false;
        while (p < t.size()) {
            char c = t[p];
            if (c >= '0' && c <= '9') { ++p; continue; }
            if (c == '.' || c == 'e' || c == 'E' || ((c == '-' || c == '+') && (t[p - 1] == 'e' || t[p - 1] == 'E'))) {
                isReal = true;
                ++p;
                continue;
            }
            break;
        }
        if (p == s) return fail("invalid value");
        std::string num = t.substr(s, p - s);
        char* end = nullptr;
        if (!isReal) {
            long long v = std::strtoll(num.c_str(), &end, 10);
            if (end && *end == '\0') { out = Value(static_cast<int64_t>(v)); return true; }
        }
        double d = std::strtod(num.c_str(), &end);
        if (!end || *end != '\0') return fail("invalid number '" + num + "'");
        out = Value(d);
        return true;
    }

    static void utf8(std::string& s, unsigned cp)
    {
        if (cp < 0x80) s += static_cast<char>(cp);
        else if (cp < 0x800) { s += static_cast<char>(0xC0 | (cp >> 6)); s += static_cast<char>(0x80 | (cp & 0x3F)); }
        else if (cp < 0x10000) {
            s += static_cast<char>(0xE0 | (cp >> 12));
            s += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
            s += static_cast<char>(0x80 | (cp & 0x3F));
        } else {
            s += static_cast<char>(0xF0 | (cp >> 18));
            s += static_cast<char>(0x80 | ((cp >> 12) & 0x3F));
            s += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
            s += static_cast<char>(0x80 | (cp & 0x3F));
        }
    }

    bool hex4(unsigned& v)
    {
        if (p + 4 > t.size()) return fail("short \\u escape");
        v = 0;
        for (int k = 0; k < 4; ++k) {
            char c = t[p++];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= c - '0';
            else if (c >= 'a' && c <= 'f') v |= c - 'a' + 10;
            else if (c >= 'A' && c <= 'F') v |= c - 'A' + 10;
            else return fail("bad \\u escape");
        }
        return true;
    }

    bool str(std::string& s)
    {
        ++p;  // opening quote
        while (p < t.size()) {
            char c = t[p++];
            if (c == '"') return true;
            if (c != '\\') { s += c; continue; }
            if (p >= t.size()) break;
            char e = t[p++];
            switch (e) {
            case '"': s += '"'; break;
            case '\\': s += '\\'; break;
            case '/': s += '/'; break;
            case 'b': s += '\b'; break;
            case 'f': s += '\f'; break;
            case 'n': s += '\n'; break;
            case 'r': s += '\r'; break;
            case 't': s += '\t'; break;
            case 'u': {
                unsigned cp = 0;
                if (!hex4(cp)) return false;
                if (cp >= 0xD800 && cp < 0xDC00 && p + 6 <= t.size() && t[p] == '\\' && t[p + 1] == 'u') {
                    p += 2;
                    unsigned lo = 0;
                    if (!hex4(lo)) return false;
                    cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                }
                utf8(s, cp);
                break;
            }
            default: return fail("bad escape");
            }
        }
        return fail("unterminated string");
    }

    bool array(Value& out)
    {
        ++p;
        out = Value::array();
        ws();
        if (p < t.size() && t[p] == ']') { ++p; return true; }
        for (;;) {
            Value v;
            if (!value(v)) return false;
            out.append(std::move(v));
            ws();
            if (p >= t.size()) return fail("unterminated array");
            if (t[p] == ',') { ++p; continue; }
            if (t[p] == ']') { ++p; return true; }
            return fail("expected ',' or ']'");
        }
    }

    bool object(Value& out)
    {
        ++p;
        out = Value::object();
        ws();
        if (p < t.size() && t[p] == '}') { ++p; return true; }
        for (;;) {
            ws();
            if (p >= t.size() || t[p] != '"') return fail("expected member name");
            std::string k;
            if (!str(k)) return false;
            ws();
            if (p >= t.size() || t[p] != ':') return fail("expected ':'");
            ++p;
            Value v;
            if (!value(v)) return false;
            out[k] = std::move(v);
            ws();
            if (p >= t.size()) return fail("unterminated object");
            if (t[p] == ',') { ++p; continue; }
            if (t[p] == '}') { ++p; return true; }
            return fail("expected ',' or '}'");
        }
    }
};

void writeString(std::string& o, const std::string& s)
{
    o += '"';
    for (unsigned char c : s) {
        switch (c) {
        case '"': o += "\\\""; break;
        case '\\': o += "\\\\"; break;
        case '\n': o += "\\n"; break;
        case '\r': o += "\\r"; break;
        case '\t': o += "\\t"; break;
        case '\b': o += "\\b"; break;
        case '\f': o += "\\f"; break;
        default:
            if (c < 0x20) {
                char buf[8];
                std::snprintf(buf, sizeof buf, "\\u%04x", c);
                o += buf;
            } else {
                o += static_cast<char>(c);
            }
        }
    }
    o += '"';
}

void writeReal(std::string& o, double d)
{
    if (!std::isfinite(d)) {  // jsoncpp writes these as non-JSON tokens; keep the file parseable
        o += std::isnan(d) ? "null" : (d > 0 ? "1e+9999" : "-1e+9999");
        return;
    }
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.17g", d);
    std::string s(buf);
    if (s.find_first_of(".eEn") == std::string::npos) s += ".0";
    o += s;
}

void writeValue(std::string& o, const Value& v, const std::string& unit, int depth)
{
    auto nl = [&](int d) {
        o += '\n';
        for (int k = 0; k < d; ++k) o += unit;
    };
    switch (v.type()) {
    case Value::Null: o += "null"; break;
    case Value::Bool: o += v.asBool() ? "true" : "false"; break;
    case Value::Int: o += std::to_string(v.asInt()); break;
    case Value::Real: writeReal(o, v.asDouble()); break;
    case Value::String: writeString(o, v.asString()); break;
    case Value::Array: {
        if (v.size() == 0) { o += "[]"; break; }
        o += '[';
        bool first = true;
        for (const auto& it : v.items()) {
            if (!first) o += ',';
            first = false;
            nl(depth + 1);
            writeValue(o, it, unit, depth + 1);
        }
        nl(depth);
        o += ']';
        break;
    }
    case Value::Object: {
        if (v.size() == 0) { o += "{}"; break; }
        o += '{';
        bool first = true;
        for (const auto& kv : v.members()) {
            if (!first) o += ',';
            first = false;
            nl(depth + 1);
            writeString(o, kv.first);
            o += " : ";
            writeValue(o, kv.second, unit, depth + 1);
        }
        nl(depth);
        o += '}';
        break;
    }
    }
}

}  // namespace

bool parse(const std::string& text, Value& out, std::string* err)
{
    Reader r(text);
    Value v;
    bool ok = r.value(v);
    if (ok) {
        r.ws();
        if (r.p != text.size()) ok = r.fail("trailing characters");
    }
    if (!ok) {
        if (err) *err = r.err;
        return false;
    }
    out = std::move(v);
    return true;
}

bool parseFile(const std::string& path, Value& out, std::string* err)
{
    std::ifstream f(path, std::ios::binary);
    if (!f.good()) {
        if (err) *err = "cannot open " + path;
        return false;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    return parse(ss.str(), out, err);
}

std::string write(const Value& v, const std::string& indent)
{
    std::string o;
    writeValue(o, v, indent, 0);
    return o;
}

}  // namespace fpsjson
