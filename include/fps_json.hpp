// fps_json.hpp -- a small JSON value, reader and writer for the reference's .json formats (csrc/json.cpp).
//
// The reference uses jsoncpp 1.9.5 (include/json/json.h:96).  This is an independent,
// minimal replacement covering what the simulation and results files need: objects with
// sorted keys (jsoncpp stores members in a std::map, so its output is key-sorted), arrays,
// strings, numbers (integers kept as integers, doubles written with 17 significant digits),
// booleans and null.  Lookups that miss return a shared null value, as jsoncpp's const
// operator[] does.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace fpsjson {

class Value {
public:
    enum Type { Null, Bool, Int, Real, String, Array, Object };

    Value() = default;
    Value(bool b) : type_(Bool), b_(b) {}
    Value(int i) : type_(Int), i_(i) {}
    Value(int64_t i) : type_(Int), i_(i) {}
    Value(unsigned u) : type_(Int), i_(static_cast<int64_t>(u)) {}
    Value(double d) : type_(Real), d_(d) {}
    Value(const char* s) : type_(String), s_(s) {}
    Value(std::string s) : type_(String), s_(std::move(s)) {}
    static Value array() { Value v; v.type_ = Array; return v; }
    static Value object() { Value v; v.type_ = Object; return v; }

    Type type() const { return type_; }
    bool isNull() const { return type_ == Null; }
    bool isBool() const { return type_ == Bool; }
    bool isNumber() const { return type_ == Int || type_ == Real; }
    bool isString() const { return type_ == String; }
    bool isArray() const { return type_ == Array; }
    bool isObject() const { return type_ == Object; }

    double asDouble() const;
    int64_t asInt() const;
    bool asBool() const;
    std::string asString() const;

    size_t size() const;
    bool isMember(const std::string& k) const;
    // const lookups: a miss returns null (jsoncpp semantics)
    const Value& operator[](const std::string& k) const;
    const Value& operator[](size_t i) const;
    // mutable lookups create the member / grow the array
    Value& operator[](const std::string& k);
    Value& operator[](size_t i);
    Value get(const std::string& k, const Value& dflt) const;
    void append(Value v);
    const std::map<std::string, Value>& members() const { return obj_; }
    const std::vector<Value>& items() const { return arr_; }

    bool operator==(const Value& o) const;
    bool operator!=(const Value& o) const { return !(*this == o); }

private:
    Type type_ = Null;
    bool b_ = false;
    int64_t i_ = 0;
    double d_ = 0.0;
    std::string s_;
    std::vector<Value> arr_;
    std::map<std::string, Value> obj_;
};

// Parses text; on failure returns false and sets err (line:column and reason).
bool parse(const std::string& text, Value& out, std::string* err);
bool parseFile(const std::string& path, Value& out, std::string* err);
// Writes with the given indentation unit ("\t" as the reference's writer uses, src/FHSolver.cpp:297).
std::string write(const Value& v, const std::string& indent = "\t");

}  // namespace fpsjson
