// Scene model with the reference's types and Scene API
// (core/include/Scene.h:11-214): Camera, FilterType, Phase/Medium,
// MaterialType/Material, TriangleMesh, LightType/Light, TextureFile, Texture,
// HdrTexture and Scene::AddMesh/AddTexture/AddLight/AddMedium/AddEnv with
// bounds.  Device handles (cudaTextureObject_t, LUT pointers) are gone: the
// HIP library owns device data; per-mesh data is de-indexed as
// TriangleMesh::LoadMesh does (core/src/Scene.cpp:83-179).
#pragma once
#include <string>
#include <vector>

#include "vec.h"

using namespace gdt;

struct Camera {
    vec3f from;
    vec3f at;
    vec3f up;
    int medium = -1;
};

enum FilterType { Box, Tent, Triangle, Gaussian };

enum PhaseType { Isotropic, HenyeyGreenstein };
struct Phase {
    PhaseType type = PhaseType::Isotropic;
    vec3f g;
};
enum MediumType { Homogeneous };
struct Medium {
    MediumType type = MediumType::Homogeneous;
    Phase phase;
    vec3f sigma_s = vec3f(0.74f, 0.88f, 1.01f);
    vec3f sigma_a = vec3f(0.032f, 0.17f, 0.48f);
    float scale = 500.0f;
};

enum MaterialType {
    Diffuse,
    Conductor,
    Dielectric,
    Plastic,
    MetalWorkflow,
    ThinDielectric,
    ClearCoatedConductor,
    DiffuseTransmitter
};

struct Material {
    MaterialType type = MaterialType::MetalWorkflow;
    vec3f albedo = 1.0f;
    int albedoTextureID = -1;
    float roughness = 0.1f;
    int roughnessTextureID = -1;
    float anisotropy = 0.0f;
    int anisotropyTextureID = -1;
    vec3f eta{0.14282f, 0.37414f, 1.43944f};
    vec3f k{3.97472f, 2.38066f, 1.59981f};
    float int_ior = 1.5f;
    float ext_ior = 1.0f;
    bool nonlinear = true;
    vec3f specular = 1.0f;
    int specularTextureID = -1;
    float metallic = 0.8f;
    int metallicTextureID = -1;
    float coat_weight = 2.0f;
    float coat_roughness_u = 0.1f;
    int coat_roughness_uTextureID = -1;
    float coat_roughness_v = 0.1f;
    int coat_roughness_vTextureID = -1;
    int normalTextureID = -1;
};

struct TriangleMesh {
    TriangleMesh() = default;
    void LoadMesh(const std::string& objFile, const Material& material, int out_med, int in_med);

    std::vector<vec3f> vertex;
    std::vector<vec3f> normal;
    std::vector<vec2f> texcoord;
    std::vector<vec3i> index;
    Material material;
    int in_medium = -1;
    int out_medium = -1;
};

enum LightType { Quad, Sphere };
struct Light {
    LightType type = LightType::Sphere;
    bool doubleSide = false;
    vec3f radiance = 3.0f;
    vec3f position{6.0f, 2.0f, 0.0f};
    vec3f u{0.0f, 0.0f, 6.0f};
    vec3f v{6.0f, 0.0f, 0.0f};
    float radius = 3.0f;
    int medium = -1;
};

struct TextureFile {
    std::string albedoFile = "";
    std::string roughnessFile = "";
    std::string anisotropyFile = "";
    std::string specularFile = "";
    std::string metallicFile = "";
    std::string coat_roughness_uFile = "";
    std::string coat_roughness_vFile = "";
    std::string normalFile = "";
};

// RGBA8 image, rows flipped like stbi_set_flip_vertically_on_load(true).
// Reads binary PPM (P6) and uncompressed 24/32-bit BMP; other formats throw.
struct Texture {
    Texture(const std::string& fileName);
    std::vector<uint8_t> pixel;
    vec2i resolution{-1};
    int comp{4};
};

// RGBA32F environment map, rows flipped (Scene.cpp:197-206).  Reads Radiance
// RGBE (.hdr, flat or RLE scanlines) and PFM.  A missing or unreadable file
// throws (the reference attached an empty env, Scene.cpp:197-206).
struct HdrTexture {
    HdrTexture(const std::string& fileName);
    HdrTexture(int w, int h, std::vector<float> rgba);
    std::vector<float> hdr;
    int width = 0, height = 0, comp = 4;
};

class Scene {
public:
    Scene() = default;
    ~Scene();
    Scene(const Scene&) = delete;
    Scene& operator=(const Scene&) = delete;

    void AddMesh(const std::string& objFile, Material& material, const TextureFile& textureFile, int out_med = -1,
                 int in_med = -1);
    // De-indexed triangle soup straight from memory (9 floats per triangle).
    void AddTriangles(const std::vector<float>& positions, const std::vector<float>& normals, const Material& material,
                      int out_med = -1, int in_med = -1);
    // A mesh loaded by the caller (the scene file's transformed shapes); the
    // scene takes ownership.
    void AddMesh(TriangleMesh* mesh);
    void AddTexture(const std::string& fileName);
    void AddLight(const Light& l);
    void AddMedium(const Medium& m);
    void AddEnv(const std::string& fileName);
    void AddEnv(HdrTexture* env);

    std::vector<TriangleMesh*> meshes;
    std::vector<Texture*> textures;
    std::vector<Light> lights;
    std::vector<Medium> mediums;
    HdrTexture* env = nullptr;
    box3f bounds;
};
