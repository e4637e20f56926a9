// Scene-from-JSON (reference: core/include/SceneParser.h, core/src/SceneParser.cpp,
// schema of models/test.json:1-128).
//
// LoadFromJson(fn, scene, is_succeed) is the reference's call: it reads the
// settings (width, height, depth, filter) and leaves the scene alone
// (SceneParser.cpp:11-23), so main.cpp, which builds its scene in code and
// then reads models/test.json for the window size (main.cpp:233-235),
// renders the same scene as with the reference.
//
// LoadSceneFromJson(fn, scene, is_succeed) reads the settings and the whole
// scene the file describes:
//
//   camera    look_from, look_at (default: the centre of the shapes' bounds),
//             up (default +y), medium (a medium name)
//   light[]   infinite: hdr_path, scale (the map's radiance times scale);
//             quad: position, u, v, radiance, double_side, medium;
//             sphere: position, radius, radiance, medium
//   shape[]   mesh: mesh_path (OBJ), translate, rotate (degrees about x, then
//             y, then z), scale (per axis), material (a name), medium_int /
//             medium_ext (names: the mesh's inside / outside medium)
//   material[] type (diffuse, conductor, dielectric, plastic, metal_workflow,
//             thin_dielectric, clearcoated_conductor, diffuse_transmitter;
//             smooth_* and rough_* variants of the first four), name, and per
//             parameter p: p (a constant), p_texture_type "constant" or
//             "image" with p_path (PPM or BMP); the scalars eta, k, int_ior,
//             ext_ior, nonlinear, coat_weight, and normal_path for a normal map
//   medium[]  homogeneous: name, phase "isotropic" or "henyey_greenstein" with
//             g, sigma_a, sigma_s, scale
//
// Paths are relative to the JSON file's folder.  The reference's Material
// types have no smooth BSDFs (Scene.h:59-68): smooth_* maps to the rough type
// at roughness kSmoothRoughness (a near-mirror GGX lobe), and smooth_diffuse
// to Diffuse at roughness 0 (Lambert).  The filter type is actually stored
// (SceneParser.cpp:27-45 assigned it to a shadowing local) and depth is
// exposed for the renderer's max depth.  A missing name, file or unknown type
// throws std::runtime_error.  A small JSON reader of our own replaces
// nlohmann::json.
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
    // a 3-number array, or one number for all three; dflt when absent
    vec3f vec3(const std::string& k, vec3f dflt) const;
};
// Throws std::runtime_error with the byte offset on malformed input.
Value parse(const std::string& text);
// The same from a BSON document (bsonspec.org, version 1.1), as the
// reference reads a file whose name ends in "bson" (SceneParser.h:38-39,
// nlohmann::json::from_bson): double, string, document, array, bool, null,
// int32, uint64 timestamp and int64 elements (numbers become doubles, array
// keys are ignored); any other element type, a size that does not match or
// bytes after the document throw std::runtime_error with the byte offset.
Value parse_bson(const std::string& bytes);
}  // namespace oxr_json

constexpr float kSmoothRoughness = 0.02f;

class SceneParser {
public:
    SceneParser() = default;
    void LoadFromJson(const std::string& fn, Scene& scene, bool& is_succeed);
    void LoadSceneFromJson(const std::string& fn, Scene& scene, bool& is_succeed);
    void ParseSettings(const oxr_json::Value& data);
    void Parse(const oxr_json::Value& data, Scene& scene);
    void ParseFilter(const oxr_json::Value& data);
    void ParseMedia(const oxr_json::Value& data, Scene& scene);
    void ParseMaterials(const oxr_json::Value& data, Scene& scene);
    void ParseShapes(const oxr_json::Value& data, Scene& scene);
    void ParseLights(const oxr_json::Value& data, Scene& scene);
    void ParseCamera(const oxr_json::Value& data, const Scene& scene);

    int width = 1024, height = 1024, depth = 4;
    FilterType filterType = FilterType::Gaussian;
    bool hasCamera = false;  // the file held a camera
    Camera camera;
    std::string baseDir;     // folder of the JSON file; relative paths start here

private:
    std::string path(const std::string& p) const;
    int medium(const std::string& name) const;
    std::map<std::string, int> mediumIds;
    std::map<std::string, Material> materials;
    std::map<std::string, int> textureIds;  // one image per distinct file
};
