// Minimal JSON DOM for scene files (objects, arrays, numbers, strings,
// booleans, null).  Parse errors throw JsonError with a byte offset.
#pragma once

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace rt {

struct JsonError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

class Json {
public:
    enum Kind { Null, Bool, Number, String, Array, Object };

    Kind kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Json> arr;
    std::vector<std::pair<std::string, Json>> obj;  // keeps file order

    static Json parse(const std::string& text);

    bool is_object() const { return kind == Object; }
    bool is_array() const { return kind == Array; }
    bool is_number() const { return kind == Number; }
    bool is_string() const { return kind == String; }
    bool is_bool() const { return kind == Bool; }
    bool contains(const std::string& key) const;
    // Missing keys yield a shared null value (like reading an absent key).
    const Json& operator[](const std::string& key) const;
    const Json& operator[](size_t i) const;
    size_t size() const { return kind == Array ? arr.size() : (kind == Object ? obj.size() : 0); }

    double as_number(const char* what) const;
    float as_float(const char* what) const { return (float)as_number(what); }
    int as_int(const char* what) const;
    bool as_bool(const char* what) const;
    const std::string& as_string(const char* what) const;
};

}  // namespace rt
