// json.h -- a small JSON-with-comments reader for scene files.
//
// Scene files follow the reference's schema 1.0 and are parsed the way the
// reference's loader sees them (code/Scenes/SceneLoader.cpp:475-491, which uses
// nlohmann::json with comments enabled):
//   - // and /* */ comments are skipped,
//   - objects iterate in sorted key order (nlohmann's default std::map),
//   - numbers are held as double (strtod) and narrowed with a cast when a
//     float is requested; integers are exact.
// Errors carry the line:column of the offending byte.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace pcr::json {

struct ParseError : std::runtime_error {
    int line, col;
    ParseError(const std::string &m, int l, int c) : std::runtime_error(m), line(l), col(c) {}
};

class Value {
public:
    enum Type { Null, Bool, Number, String, Array, Object };
    Type type = Null;
    bool b = false;
    double num = 0.0;
    bool is_int = false;
    std::string str;
    std::vector<Value> arr;
    std::map<std::string, Value> obj;

    bool is_null() const { return type == Null; }
    bool is_bool() const { return type == Bool; }
    bool is_number() const { return type == Number; }
    bool is_string() const { return type == String; }
    bool is_array() const { return type == Array; }
    bool is_object() const { return type == Object; }
    const Value *find(const std::string &k) const {
        if (type != Object) return nullptr;
        auto it = obj.find(k);
        return it == obj.end() ? nullptr : &it->second;
    }
    bool contains(const std::string &k) const { return find(k) != nullptr; }
    float as_float() const { return (float)num; }
};

Value parse(const std::string &text);

}  // namespace pcr::json
