#include "SceneParser.h"

#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace oxr_json {
namespace {

struct Reader {
    const std::string& s;
    size_t i = 0;
    [[noreturn]] void fail(const char* what) {
        throw std::runtime_error(std::string("json: ") + what + " at byte " + std::to_string(i));
    }
    void ws() {
        while (i < s.size() && std::isspace((unsigned char)s[i])) i++;
    }
    bool eat(char c) {
        ws();
        if (i < s.size() && s[i] == c) { i++; return true; }
        return false;
    }
    Value value(int depth) {
        if (depth > 64) fail("nesting too deep");
        ws();
        if (i >= s.size()) fail("unexpected end");
        char c = s[i];
        Value v;
        if (c == '{') {
            i++;
            v.kind = Value::Object;
            if (eat('}')) return v;
            do {
                ws();
                if (i >= s.size() || s[i] != '"') fail("expected key");
                std::string k = string();
                if (!eat(':')) fail("expected ':'");
                v.obj[k] = value(depth + 1);
            } while (eat(','));
            if (!eat('}')) fail("expected '}'");
        } else if (c == '[') {
            i++;
            v.kind = Value::Array;
            if (eat(']')) return v;
            do v.arr.push_back(value(depth + 1));
            while (eat(','));
            if (!eat(']')) fail("expected ']'");
        } else if (c == '"') {
            v.kind = Value::String;
            v.str = string();
        } else if (s.compare(i, 4, "true") == 0) {
            v.kind = Value::Bool; v.b = true; i += 4;
        } else if (s.compare(i, 5, "false") == 0) {
            v.kind = Value::Bool; i += 5;
        } else if (s.compare(i, 4, "null") == 0) {
            i += 4;
        } else {
            const char* b = s.c_str() + i;
            char* e = nullptr;
            v.num = std::strtod(b, &e);
            if (e == b) fail("unexpected character");
            v.kind = Value::Number;
            i += (size_t)(e - b);
        }
        return v;
    }
    std::string string() {
        i++;  // opening quote
        std::string out;
        while (i < s.size() && s[i] != '"') {
            if (s[i] == '\\') {
                if (++i >= s.size()) fail("bad escape");
                char e = s[i];
                out += e == 'n' ? '\n' : e == 't' ? '\t' : e == 'r' ? '\r' : e == 'b' ? '\b' : e == 'f' ? '\f' : e;
                if (e == 'u') fail("\\u escapes are not supported");
                i++;
            } else {
                out += s[i++];
            }
        }
        if (i >= s.size()) fail("unterminated string");
        i++;
        return out;
    }
};

}  // namespace

Value parse(const std::string& text) {
    Reader r{text};
    Value v = r.value(0);
    r.ws();
    if (r.i != text.size()) r.fail("trailing characters");
    return v;
}

vec3f Value::vec3(const std::string& k, vec3f dflt) const {
    const Value* v = get(k);
    if (!v) return dflt;
    if (v->kind == Number) return vec3f((float)v->num);
    if (v->kind != Array || v->arr.size() != 3 || v->arr[0].kind != Number || v->arr[1].kind != Number ||
        v->arr[2].kind != Number)
        throw std::runtime_error("json: '" + k + "' is not a number or a 3-number array");
    return vec3f((float)v->arr[0].num, (float)v->arr[1].num, (float)v->arr[2].num);
}

}  // namespace oxr_json

namespace {

const oxr_json::Value* array_of(const oxr_json::Value& data, const char* key) {
    const oxr_json::Value* v = data.get(key);
    if (!v) return nullptr;
    if (v->kind != oxr_json::Value::Array) throw std::runtime_error(std::string("json: '") + key + "' is not an array");
    return v;
}

std::string need_string(const oxr_json::Value& o, const char* key, const char* what) {
    const oxr_json::Value* v = o.get(key);
    if (!v || v->kind != oxr_json::Value::String)
        throw std::runtime_error(std::string("json: ") + what + " without a string '" + key + "'");
    return v->str;
}

// Rotation by rx about x, then ry about y, then rz about z (degrees).
vec3f rotate_xyz(vec3f p, vec3f deg) {
    const float k = 3.14159265358979323846f / 180.0f;
    float c = std::cos(deg.x * k), s = std::sin(deg.x * k);
    p = vec3f(p.x, c * p.y - s * p.z, s * p.y + c * p.z);
    c = std::cos(deg.y * k); s = std::sin(deg.y * k);
    p = vec3f(c * p.x + s * p.z, p.y, -s * p.x + c * p.z);
    c = std::cos(deg.z * k); s = std::sin(deg.z * k);
    return vec3f(c * p.x - s * p.y, s * p.x + c * p.y, p.z);
}

}  // namespace

namespace oxr_json {
namespace {

// Little-endian BSON reader; i is the offset of the next byte.
struct BsonReader {
    const std::string& s;
    size_t i = 0;
    [[noreturn]] void fail(const std::string& what) {
        throw std::runtime_error("bson: " + what + " at byte " + std::to_string(i));
    }
    void need(size_t n) {
        if (n > s.size() - i) fail("unexpected end");
    }
    uint64_t le(int bytes) {
        need((size_t)bytes);
        uint64_t v = 0;
        for (int k = 0; k < bytes; k++) v |= (uint64_t)(uint8_t)s[i + k] << (8 * k);
        i += (size_t)bytes;
        return v;
    }
    std::string cstring() {
        const size_t end = s.find('\0', i);
        if (end == std::string::npos) fail("unterminated element name");
        std::string r = s.substr(i, end - i);
        i = end + 1;
        return r;
    }
    Value document(bool array, int depth) {
        if (depth > 64) fail("nesting too deep");
        const size_t start = i;
        const uint64_t size = le(4);
        if (size < 5 || size > s.size() - start) fail("bad document size");
        const size_t end = start + (size_t)size;
        Value v;
        v.kind = array ? Value::Array : Value::Object;
        for (;;) {
            if (i >= end) fail("document runs past its size");
            const uint8_t type = (uint8_t)le(1);
            if (type == 0) break;
            const std::string key = cstring();
            Value e = element(type, depth);
            if (i > end) fail("element runs past its document");
            if (array) v.arr.push_back(std::move(e));
            else v.obj[key] = std::move(e);
        }
        if (i != end) fail("document size does not match its elements");
        return v;
    }
    Value element(uint8_t type, int depth) {
        Value e;
        switch (type) {
        case 0x01: {                             // double
            const uint64_t bits = le(8);
            double d;
            std::memcpy(&d, &bits, 8);
            e.kind = Value::Number;
            e.num = d;
            break;
        }
        case 0x02: {                             // string: int32 length with the NUL, bytes, NUL
            const uint64_t n = le(4);
            if (n < 1 || n > 0x7fffffffu) fail("bad string length");
            need((size_t)n);
            if (s[i + n - 1] != '\0') fail("unterminated string");
            e.kind = Value::String;
            e.str = s.substr(i, (size_t)n - 1);
            i += (size_t)n;
            break;
        }
        case 0x03: e = document(false, depth + 1); break;
        case 0x04: e = document(true, depth + 1); break;
        case 0x08: {
            const uint64_t b = le(1);
            if (b > 1) fail("bad boolean");
            e.kind = Value::Bool;
            e.b = b == 1;
            break;
        }
        case 0x0A: break;                        // null
        case 0x10: e.kind = Value::Number; e.num = (double)(int32_t)(uint32_t)le(4); break;
        case 0x11: e.kind = Value::Number; e.num = (double)le(8); break;
        case 0x12: e.kind = Value::Number; e.num = (double)(int64_t)le(8); break;
        default: {
            char hex[8];
            std::snprintf(hex, sizeof hex, "0x%02x", type);
            i -= 1;
            fail(std::string("unsupported element type ") + hex);
        }
        }
        return e;
    }
};

}  // namespace

Value parse_bson(const std::string& bytes) {
    BsonReader r{bytes};
    Value v = r.document(false, 0);
    if (r.i != bytes.size()) r.fail("bytes after the document");
    return v;
}

}  // namespace oxr_json

namespace {

// The file's JSON, or false when it cannot be opened (SceneParser.h:31-35:
// a missing file leaves is_succeed false).  A name ending in "bson" is read
// as BSON (SceneParser.h:13-21, 38-39).
bool read_json(const std::string& fn, oxr_json::Value& out, std::string& baseDir) {
    std::ifstream in(fn, std::ios::binary);
    if (!in.is_open()) return false;
    std::stringstream buf;
    buf << in.rdbuf();
    const bool bson = fn.size() >= 4 && fn.compare(fn.size() - 4, 4, "bson") == 0;
    out = bson ? oxr_json::parse_bson(buf.str()) : oxr_json::parse(buf.str());
    size_t slash = fn.find_last_of('/');
    baseDir = slash == std::string::npos ? std::string() : fn.substr(0, slash + 1);
    return true;
}

}  // namespace

void SceneParser::LoadFromJson(const std::string& fn, Scene& scene, bool& is_succeed) {
    (void)scene;  // the reference's parser reads settings only (SceneParser.cpp:11-23)
    oxr_json::Value data;
    is_succeed = read_json(fn, data, baseDir);
    if (is_succeed) ParseSettings(data);
}

void SceneParser::LoadSceneFromJson(const std::string& fn, Scene& scene, bool& is_succeed) {
    oxr_json::Value data;
    is_succeed = read_json(fn, data, baseDir);
    if (is_succeed) Parse(data, scene);
}

void SceneParser::ParseSettings(const oxr_json::Value& data) {
    width = (int)data.number("width", 1024);
    height = (int)data.number("height", 1024);
    depth = (int)data.number("depth", 4);
    const oxr_json::Value* f = data.get("filter");
    oxr_json::Value empty;
    ParseFilter(f ? *f : empty);
}

void SceneParser::Parse(const oxr_json::Value& data, Scene& scene) {
    ParseSettings(data);
    // media first (materials' shapes and lights name them), then materials,
    // shapes (the bounds a default look_at needs), lights and the camera
    ParseMedia(data, scene);
    ParseMaterials(data, scene);
    ParseShapes(data, scene);
    ParseLights(data, scene);
    ParseCamera(data, scene);
}

std::string SceneParser::path(const std::string& p) const {
    return (p.empty() || p[0] == '/') ? p : baseDir + p;
}

int SceneParser::medium(const std::string& name) const {
    if (name.empty()) return -1;
    auto it = mediumIds.find(name);
    if (it == mediumIds.end()) throw std::runtime_error("json: unknown medium '" + name + "'");
    return it->second;
}

void SceneParser::ParseMedia(const oxr_json::Value& data, Scene& scene) {
    const oxr_json::Value* arr = array_of(data, "medium");
    if (!arr) return;
    for (const oxr_json::Value& m : arr->arr) {
        std::string name = need_string(m, "name", "medium");
        if (m.string("type", "homogeneous") != "homogeneous")
            throw std::runtime_error("json: medium '" + name + "': only homogeneous media exist (Medium.h)");
        Medium md;
        std::string phase = m.string("phase", "isotropic");
        if (phase == "isotropic") md.phase.type = PhaseType::Isotropic;
        else if (phase == "henyey_greenstein" || phase == "hg") md.phase.type = PhaseType::HenyeyGreenstein;
        else throw std::runtime_error("json: medium '" + name + "': unknown phase '" + phase + "'");
        md.phase.g = m.vec3("g", vec3f(0.0f));
        md.sigma_a = m.vec3("sigma_a", md.sigma_a);
        md.sigma_s = m.vec3("sigma_s", md.sigma_s);
        md.scale = (float)m.number("scale", md.scale);
        if (mediumIds.count(name)) throw std::runtime_error("json: medium '" + name + "' defined twice");
        mediumIds[name] = (int)scene.mediums.size();
        scene.AddMedium(md);
    }
}

void SceneParser::ParseMaterials(const oxr_json::Value& data, Scene& scene) {
    const oxr_json::Value* arr = array_of(data, "material");
    if (!arr) return;
    struct Kind { const char* name; MaterialType type; int smooth; };  // smooth: 0 no, 1 near-mirror, 2 Lambert
    static const Kind kinds[] = {
        {"diffuse", Diffuse, 0}, {"rough_diffuse", Diffuse, 0}, {"smooth_diffuse", Diffuse, 2},
        {"conductor", Conductor, 0}, {"rough_conductor", Conductor, 0}, {"smooth_conductor", Conductor, 1},
        {"dielectric", Dielectric, 0}, {"rough_dielectric", Dielectric, 0}, {"smooth_dielectric", Dielectric, 1},
        {"plastic", Plastic, 0}, {"rough_plastic", Plastic, 0}, {"smooth_plastic", Plastic, 1},
        {"metal_workflow", MetalWorkflow, 0}, {"thin_dielectric", ThinDielectric, 0},
        {"clearcoated_conductor", ClearCoatedConductor, 0}, {"diffuse_transmitter", DiffuseTransmitter, 0},
    };
    for (const oxr_json::Value& m : arr->arr) {
        std::string name = need_string(m, "name", "material");
        std::string type = need_string(m, "type", "material");
        const Kind* kind = nullptr;
        for (const Kind& k : kinds)
            if (type == k.name) kind = &k;
        if (!kind) throw std::runtime_error("json: material '" + name + "': unknown type '" + type + "'");
        Material mat;
        mat.type = kind->type;
        if (kind->smooth == 1) mat.roughness = kSmoothRoughness;
        if (kind->smooth == 2) mat.roughness = 0.0f;
        // a parameter: its constant, or an image when p_texture_type is "image"
        auto image = [&](const std::string& p, int& id) {
            std::string tt = m.string(p + "_texture_type", m.get(p + "_path") ? "image" : "constant");
            if (tt == "constant") return false;
            if (tt != "image")
                throw std::runtime_error("json: material '" + name + "': " + p + "_texture_type '" + tt + "'");
            std::string file = path(need_string(m, (p + "_path").c_str(), ("material '" + name + "'").c_str()));
            auto it = textureIds.find(file);
            if (it == textureIds.end()) {
                it = textureIds.emplace(file, (int)scene.textures.size()).first;
                scene.AddTexture(file);
            }
            id = it->second;
            return true;
        };
        if (!image("albedo", mat.albedoTextureID)) mat.albedo = m.vec3("albedo", mat.albedo);
        if (!image("roughness", mat.roughnessTextureID) && kind->smooth == 0)
            mat.roughness = (float)m.number("roughness", mat.roughness);
        if (!image("anisotropy", mat.anisotropyTextureID)) mat.anisotropy = (float)m.number("anisotropy", mat.anisotropy);
        if (!image("specular", mat.specularTextureID)) mat.specular = m.vec3("specular", mat.specular);
        if (!image("metallic", mat.metallicTextureID)) mat.metallic = (float)m.number("metallic", mat.metallic);
        if (!image("coat_roughness_u", mat.coat_roughness_uTextureID))
            mat.coat_roughness_u = (float)m.number("coat_roughness_u", mat.coat_roughness_u);
        if (!image("coat_roughness_v", mat.coat_roughness_vTextureID))
            mat.coat_roughness_v = (float)m.number("coat_roughness_v", mat.coat_roughness_v);
        image("normal", mat.normalTextureID);
        mat.eta = m.vec3("eta", mat.eta);
        mat.k = m.vec3("k", mat.k);
        mat.int_ior = (float)m.number("int_ior", mat.int_ior);
        mat.ext_ior = (float)m.number("ext_ior", mat.ext_ior);
        const oxr_json::Value* nl = m.get("nonlinear");
        if (nl && nl->kind == oxr_json::Value::Bool) mat.nonlinear = nl->b;
        mat.coat_weight = (float)m.number("coat_weight", mat.coat_weight);
        if (materials.count(name)) throw std::runtime_error("json: material '" + name + "' defined twice");
        materials[name] = mat;
    }
}

void SceneParser::ParseShapes(const oxr_json::Value& data, Scene& scene) {
    const oxr_json::Value* arr = array_of(data, "shape");
    if (!arr) return;
    for (const oxr_json::Value& s : arr->arr) {
        if (s.string("type", "mesh") != "mesh") throw std::runtime_error("json: only mesh shapes exist");
        std::string file = path(need_string(s, "mesh_path", "shape"));
        Material mat;
        std::string mname = s.string("material", "");
        if (!mname.empty()) {
            auto it = materials.find(mname);
            if (it == materials.end()) throw std::runtime_error("json: unknown material '" + mname + "'");
            mat = it->second;
        }
        const vec3f t = s.vec3("translate", vec3f(0.0f)), r = s.vec3("rotate", vec3f(0.0f));
        const vec3f sc = s.vec3("scale", vec3f(1.0f));
        if (sc.x == 0.0f || sc.y == 0.0f || sc.z == 0.0f) throw std::runtime_error("json: shape '" + file + "': zero scale");
        TriangleMesh* mesh = new TriangleMesh();
        try {
            mesh->LoadMesh(file, mat, medium(s.string("medium_ext", "")), medium(s.string("medium_int", "")));
        } catch (...) {
            delete mesh;
            throw;
        }
        // p' = translate + R (scale * p); normals by the inverse transpose, R (n / scale)
        for (vec3f& v : mesh->vertex) v = t + rotate_xyz(v * sc, r);
        for (vec3f& n : mesh->normal) n = normalize(rotate_xyz(n / sc, r));
        scene.AddMesh(mesh);
    }
}

void SceneParser::ParseLights(const oxr_json::Value& data, Scene& scene) {
    const oxr_json::Value* arr = array_of(data, "light");
    if (!arr) return;
    for (const oxr_json::Value& l : arr->arr) {
        std::string type = need_string(l, "type", "light");
        if (type == "infinite") {
            HdrTexture* env = new HdrTexture(path(need_string(l, "hdr_path", "infinite light")));
            const float k = (float)l.number("scale", 1.0);
            for (size_t i = 0; i < env->hdr.size(); i++)
                if (i % 4 != 3) env->hdr[i] *= k;
            scene.AddEnv(env);
            continue;
        }
        Light lt;
        if (type == "quad") lt.type = LightType::Quad;
        else if (type == "sphere") lt.type = LightType::Sphere;
        else throw std::runtime_error("json: unknown light type '" + type + "'");
        lt.radiance = l.vec3("radiance", lt.radiance);
        lt.position = l.vec3("position", lt.position);
        lt.u = l.vec3("u", lt.u);
        lt.v = l.vec3("v", lt.v);
        lt.radius = (float)l.number("radius", lt.radius);
        const oxr_json::Value* ds = l.get("double_side");
        if (ds) lt.doubleSide = ds->kind == oxr_json::Value::Bool ? ds->b : ds->num != 0.0;
        lt.medium = medium(l.string("medium", ""));
        scene.AddLight(lt);
    }
}

void SceneParser::ParseCamera(const oxr_json::Value& data, const Scene& scene) {
    const oxr_json::Value* c = data.get("camera");
    if (!c || c->kind != oxr_json::Value::Object) return;
    hasCamera = true;
    camera.from = c->vec3("look_from", vec3f(0.0f, 0.0f, 1.0f));
    camera.at = c->vec3("look_at", scene.meshes.empty() ? vec3f(0.0f) : scene.bounds.center());
    camera.up = c->vec3("up", vec3f(0.0f, 1.0f, 0.0f));
    camera.medium = medium(c->string("medium", ""));
}

void SceneParser::ParseFilter(const oxr_json::Value& data) {
    std::string t = data.string("type", "gaussian");
    if (t == "triangle") filterType = FilterType::Triangle;
    else if (t == "tent") filterType = FilterType::Tent;
    else if (t == "box") filterType = FilterType::Box;
    else filterType = FilterType::Gaussian;
}
