// Minimal JSON reader for the solver's JSON configuration files (objects, arrays, strings,
// numbers, true/false/null). Keys keep their file order.
#pragma once
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace dg {

struct JsonValue {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<JsonValue> arr;
    std::vector<std::pair<std::string, JsonValue>> obj;

    const JsonValue* find(const std::string& key) const {
        for (const auto& kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
    // Member access that throws with the key path in the message (missing keys are fatal
    // for the reference too: nlohmann::json throws on a missing key).
    const JsonValue& at(const std::string& key) const {
        const JsonValue* v = find(key);
        if (!v) throw std::runtime_error("JSON key '" + key + "' not found");
        return *v;
    }
    double number() const {
        if (kind != Number) throw std::runtime_error("JSON value is not a number");
        return num;
    }
    const std::string& string() const {
        if (kind != String) throw std::runtime_error("JSON value is not a string");
        return str;
    }
};

JsonValue parseJson(const std::string& text);

}  // namespace dg
