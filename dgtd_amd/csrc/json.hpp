// Minimal JSON reader for case files (RFC 8259 subset the case files use:
// objects, arrays, numbers, strings with the common escapes, true/false/null).
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

namespace dgtd {

struct Json {
    enum Kind { kNull, kBool, kNumber, kString, kArray, kObject };
    Kind kind = kNull;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Json> arr;
    std::vector<std::pair<std::string, Json>> obj;  // keeps file order

    bool is_object() const { return kind == kObject; }
    bool is_array() const { return kind == kArray; }
    bool is_number() const { return kind == kNumber; }
    bool is_string() const { return kind == kString; }
    bool contains(const std::string& k) const;
    // Member / element access; throws std::runtime_error naming the path on a miss.
    const Json& at(const std::string& k) const;
    const Json& at(size_t i) const;
    size_t size() const { return kind == kArray ? arr.size() : kind == kObject ? obj.size() : 0; }
    double number() const;
    int integer() const;
    const std::string& string() const;
    std::vector<double> numbers() const;

    static Json parse(const std::string& text);
    static Json parse_file(const std::string& path);
};

}  // namespace dgtd
