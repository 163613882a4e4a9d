// bte_json.h — a small JSON (RFC 8259) reader for glTF documents.
#pragma once

#include <string>
#include <utility>
#include <vector>

namespace bte {

struct JValue {
    enum Type { Null, Bool, Number, String, Array, Object };
    Type type = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<JValue> arr;
    std::vector<std::pair<std::string, JValue>> obj;   // document order; first match wins on lookup

    bool isObject() const { return type == Object; }
    bool isArray() const { return type == Array; }
    bool isNumber() const { return type == Number; }
    // member `key` of an object (nullptr when absent or not an object)
    const JValue* get(const char* key) const;
    // number member or the default when absent / not a number
    double number(const char* key, double def) const;
    int integer(const char* key, int def) const;
    // string member or "" when absent / not a string
    std::string string(const char* key) const;
    // array member as numbers (empty when absent; non-numbers read as 0)
    std::vector<double> numbers(const char* key) const;
    size_t size() const { return type == Array ? arr.size() : type == Object ? obj.size() : 0; }
};

// Parses a complete JSON text (UTF-8; \u escapes decoded to UTF-8, surrogate
// pairs joined).  Returns false with a message and the byte offset on error.
bool parse_json(const char* text, size_t size, JValue& out, std::string& err);

}  // namespace bte
