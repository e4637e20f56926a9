// Scene-from-JSON (reference: core/include/SceneParser.h, core/src/SceneParser.cpp).
// LoadFromJson(fn, scene, is_succeed) with the reference's settings
// (width, height, depth, filter).  Unlike the reference, the filter type is
// actually stored (SceneParser.cpp:27-45 assigned it to a shadowing local) and
// depth is exposed for the renderer's max depth.  A small JSON reader of our
// own replaces nlohmann::json.
#pragma once
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "Scene.h"

namespace oxr_json {
struct Value {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Value> arr;
    std::map<std::string, Value> obj;
    const Value* get(const std::string& k) const {
        auto it = obj.find(k);
        return (kind == Object && it != obj.end()) ? &it->second : nullptr;
    }
    double number(const std::string& k, double dflt) const {
        const Value* v = get(k);
        return (v && v->kind == Number) ? v->num : dflt;
    }
    std::string string(const std::string& k, const std::string& dflt) const {
        const Value* v = get(k);
        return (v && v->kind == String) ? v->str : dflt;
    }
};
// Throws std::runtime_error with the byte offset on malformed input.
Value parse(const std::string& text);
}  // namespace oxr_json

class SceneParser {
public:
    SceneParser() = default;
    void LoadFromJson(const std::string& fn, Scene& scene, bool& is_succeed);
    void Parse(const oxr_json::Value& data, Scene& scene);
    void ParseFilter(const oxr_json::Value& data);

    int width = 1024, height = 1024, depth = 4;
    FilterType filterType = FilterType::Gaussian;
};
