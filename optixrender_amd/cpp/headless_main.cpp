// Headless driver: the viewer-free counterpart of the reference's main.cpp
// (main.cpp:165-266).  Builds a scene through the same Scene / SceneParser /
// Renderer calls main.cpp makes, renders frames and writes the RGBA8 pixels
// (DownloadPixels) as a binary PPM.
//
//   headless [settings.json] [out.ppm] [frames] [obj...]
// A settings file with shapes is a whole scene (SceneParser), camera
// included.  Otherwise, without OBJ arguments the scene is a procedural
// Cornell box.  OXR_HEADLESS_DUMP=<file> writes the parsed scene as JSON
// (meshes after their transforms, materials, lights, media, environment,
// camera) and exits before any device call; with OXR_HEADLESS_DUMP_BRIEF=1
// each mesh lists its triangle count and bounds instead of its vertices, and
// each texture its size and byte sum instead of its texels.  With
// OXR_HEADLESS_DENOISE=1 in the environment the frames are denoised
// (denoiserOn, the viewer's 'D' key); OXR_HEADLESS_PRESENT=1 reads the frame
// back through the streamed, page-locked path (presentOn).  OBJ meshes take their textures from
// OXR_HEADLESS_ALBEDO / _ROUGHNESS / _METALLIC / _NORMAL (PPM files), and
// OXR_HEADLESS_ENV names an environment map (.hdr or .pfm), as main.cpp
// gives its meshes texture files and its scene an HDR (main.cpp:165-213).
// OXR_HEADLESS_MAIN_SCENE=<models folder> builds main.cpp's own scene from
// that folder with main.cpp's calls in main.cpp's order (head, body, base and
// plane meshes with their texture files, the HDR, then test.json's settings
// through LoadFromJson), the settings argument being ignored.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <vector>

#include "Renderer.h"
#include "SceneParser.h"

static void cornell(Scene& scene) {
    auto quad = [](vec3f a, vec3f b, vec3f c, vec3f d, std::vector<float>& p, std::vector<float>& n) {
        vec3f nn = normalize(cross(normalize(b - a), normalize(c - a)));
        for (vec3f v : {a, b, c, a, c, d}) {
            p.insert(p.end(), {v.x, v.y, v.z});
            n.insert(n.end(), {nn.x, nn.y, nn.z});
        }
    };
    const float x0 = -1, x1 = 1, y0 = 0, y1 = 2, z0 = -1, z1 = 3.2f;
    std::vector<float> gp, gn, rp, rn, ggp, ggn;
    quad({x0, y0, z0}, {x0, y0, z1}, {x1, y0, z1}, {x1, y0, z0}, gp, gn);
    quad({x0, y1, z0}, {x1, y1, z0}, {x1, y1, z1}, {x0, y1, z1}, gp, gn);
    quad({x0, y0, z0}, {x1, y0, z0}, {x1, y1, z0}, {x0, y1, z0}, gp, gn);
    quad({x0, y0, z1}, {x0, y1, z1}, {x1, y1, z1}, {x1, y0, z1}, gp, gn);
    quad({x0, y0, z0}, {x0, y1, z0}, {x0, y1, z1}, {x0, y0, z1}, rp, rn);
    quad({x1, y0, z0}, {x1, y0, z1}, {x1, y1, z1}, {x1, y1, z0}, ggp, ggn);
    Material grey, red, green;
    grey.type = red.type = green.type = MaterialType::Diffuse;
    grey.roughness = red.roughness = green.roughness = 0.0f;
    grey.albedo = vec3f(0.7f);
    red.albedo = vec3f(0.7f, 0.1f, 0.1f);
    green.albedo = vec3f(0.1f, 0.7f, 0.1f);
    scene.AddTriangles(gp, gn, grey);
    scene.AddTriangles(rp, rn, red);
    scene.AddTriangles(ggp, ggn, green);
    Light l;
    l.type = LightType::Quad;
    l.doubleSide = true;
    l.radiance = vec3f(15.0f);
    l.position = vec3f(-0.3f, 1.98f, 0.3f);
    l.u = vec3f(0.6f, 0.0f, 0.0f);
    l.v = vec3f(0.0f, 0.0f, -0.6f);
    scene.AddLight(l);
}

// main.cpp:165-246 with the models folder made explicit.
static void main_scene(const std::string& models, Scene& scene, Camera& camera, SceneParser& parser, bool& ok) {
    TextureFile textureFile;
    Material material;
    Medium m;
    textureFile.albedoFile = models + "/01_Head_Base_Color.png";
    textureFile.roughnessFile = models + "/01_Head_Roughness.png";
    textureFile.metallicFile = models + "/01_Head_Metallic.png";
    textureFile.normalFile = models + "/01_Head_Normal_DirectX.png";
    scene.AddMedium(m);
    scene.AddMesh(models + "/head.obj", material, textureFile, -1, -1);
    textureFile.albedoFile = models + "/02_Body_Base_Color.png";
    textureFile.roughnessFile = models + "/02_Body_Roughness.png";
    textureFile.metallicFile = models + "/02_Body_Metallic.png";
    textureFile.normalFile = models + "/02_Body_Normal_DirectX.png";
    Material body;
    body.type = MaterialType::Dielectric;
    TextureFile none;
    scene.AddMesh(models + "/body.obj", body, none, -1, 0);
    textureFile.albedoFile = models + "/03_Base_Base_Color.png";
    textureFile.metallicFile = models + "/03_Base_Metallic.png";
    textureFile.roughnessFile = models + "/03_Base_Roughness.png";
    textureFile.normalFile = models + "/03_Base_Normal_DirectX.png";
    scene.AddMesh(models + "/base.obj", material, textureFile, -1, -1);
    material = Material();
    material.type = MaterialType::Diffuse;
    textureFile.albedoFile = models + "/grid.jpg";
    textureFile.roughnessFile = models + "/rusty_metal_sheet_rough_1k.png";
    textureFile.normalFile = "";
    scene.AddMesh(models + "/plane.obj", material, textureFile, -1, -1);
    scene.AddEnv(models + "/spaichingen_hill_4k.hdr");
    camera = Camera{vec3f(0.2f, 0.2f, 0.2f), scene.bounds.center(), vec3f(0.f, 1.f, 0.f)};
    camera.medium = -1;
    parser.LoadFromJson(models + "/test.json", scene, ok);
}

static void dump_scene(const std::string& file, const Scene& scene, const SceneParser& parser, const Camera& cam,
                       bool brief) {
    std::ofstream o(file);
    if (!o) throw std::runtime_error("cannot write " + file);
    o.precision(9);
    auto v3 = [&](vec3f v) { o << "[" << v.x << "," << v.y << "," << v.z << "]"; };
    o << "{\"width\":" << parser.width << ",\"height\":" << parser.height << ",\"depth\":" << parser.depth
      << ",\"filter\":" << (int)parser.filterType << ",\"camera\":{\"from\":";
    v3(cam.from); o << ",\"at\":"; v3(cam.at); o << ",\"up\":"; v3(cam.up);
    o << ",\"medium\":" << cam.medium << "},\"meshes\":[";
    for (size_t m = 0; m < scene.meshes.size(); m++) {
        const TriangleMesh* me = scene.meshes[m];
        const Material& a = me->material;
        o << (m ? "," : "") << "{\"in_medium\":" << me->in_medium << ",\"out_medium\":" << me->out_medium
          << ",\"material\":{\"type\":" << (int)a.type << ",\"albedo\":";
        v3(a.albedo);
        o << ",\"roughness\":" << a.roughness << ",\"anisotropy\":" << a.anisotropy << ",\"eta\":";
        v3(a.eta); o << ",\"k\":"; v3(a.k); o << ",\"specular\":"; v3(a.specular);
        o << ",\"metallic\":" << a.metallic << ",\"int_ior\":" << a.int_ior << ",\"ext_ior\":" << a.ext_ior
          << ",\"nonlinear\":" << (a.nonlinear ? 1 : 0) << ",\"coat_weight\":" << a.coat_weight
          << ",\"coat_roughness_u\":" << a.coat_roughness_u << ",\"coat_roughness_v\":" << a.coat_roughness_v
          << ",\"tex\":[" << a.albedoTextureID << ","
          << a.roughnessTextureID << "," << a.anisotropyTextureID << "," << a.specularTextureID << ","
          << a.metallicTextureID << "," << a.coat_roughness_uTextureID << "," << a.coat_roughness_vTextureID << ","
          << a.normalTextureID << "]},\"triangles\":" << me->index.size();
        if (brief) {
            box3f b;
            for (const vec3f& v : me->vertex) b.extend(v);
            o << ",\"lower\":"; v3(b.lower); o << ",\"upper\":"; v3(b.upper); o << "}";
            continue;
        }
        o << ",\"vertex\":[";
        for (size_t i = 0; i < me->vertex.size(); i++) { o << (i ? "," : ""); v3(me->vertex[i]); }
        o << "],\"normal\":[";
        for (size_t i = 0; i < me->normal.size(); i++) { o << (i ? "," : ""); v3(me->normal[i]); }
        o << "],\"texcoord\":[";
        for (size_t i = 0; i < me->texcoord.size(); i++)
            o << (i ? "," : "") << "[" << me->texcoord[i].x << "," << me->texcoord[i].y << "]";
        o << "]}";
    }
    o << "],\"textures\":[";
    for (size_t t = 0; t < scene.textures.size(); t++) {
        const Texture* tx = scene.textures[t];
        o << (t ? "," : "") << "{\"width\":" << tx->resolution.x << ",\"height\":" << tx->resolution.y;
        if (brief) {
            uint64_t sum = 0;
            for (uint8_t b : tx->pixel) sum += b;
            o << ",\"bytes\":" << tx->pixel.size() << ",\"byte_sum\":" << sum << "}";
            continue;
        }
        o << ",\"rgba\":[";
        for (size_t i = 0; i < tx->pixel.size(); i++) o << (i ? "," : "") << (int)tx->pixel[i];
        o << "]}";
    }
    o << "],\"lights\":[";
    for (size_t l = 0; l < scene.lights.size(); l++) {
        const Light& L = scene.lights[l];
        o << (l ? "," : "") << "{\"type\":" << (int)L.type << ",\"double_side\":" << (L.doubleSide ? 1 : 0)
          << ",\"radiance\":";
        v3(L.radiance); o << ",\"position\":"; v3(L.position); o << ",\"u\":"; v3(L.u); o << ",\"v\":"; v3(L.v);
        o << ",\"radius\":" << L.radius << ",\"medium\":" << L.medium << "}";
    }
    o << "],\"media\":[";
    for (size_t m = 0; m < scene.mediums.size(); m++) {
        const Medium& M = scene.mediums[m];
        o << (m ? "," : "") << "{\"phase\":" << (int)M.phase.type << ",\"g\":";
        v3(M.phase.g); o << ",\"sigma_a\":"; v3(M.sigma_a); o << ",\"sigma_s\":"; v3(M.sigma_s);
        o << ",\"scale\":" << M.scale << "}";
    }
    o << "],\"env\":";
    if (scene.env) {
        double sum = 0.0;
        for (size_t i = 0; i < scene.env->hdr.size(); i++)
            if (i % 4 != 3) sum += scene.env->hdr[i];
        o << "{\"width\":" << scene.env->width << ",\"height\":" << scene.env->height << ",\"rgb_sum\":" << sum << "}";
    } else {
        o << "null";
    }
    o << "}\n";
}

int main(int ac, char** av) {
    try {
        std::string settings = ac > 1 ? av[1] : "";
        std::string out = ac > 2 ? av[2] : "image.ppm";
        int frames = ac > 3 ? std::atoi(av[3]) : 16;
        Scene scene;
        Camera camera;
        auto env_str = [](const char* k) { const char* v = std::getenv(k); return std::string(v ? v : ""); };
        if (ac > 4) {
            Material m;
            TextureFile t;
            t.albedoFile = env_str("OXR_HEADLESS_ALBEDO");
            t.roughnessFile = env_str("OXR_HEADLESS_ROUGHNESS");
            t.metallicFile = env_str("OXR_HEADLESS_METALLIC");
            t.normalFile = env_str("OXR_HEADLESS_NORMAL");
            for (int i = 4; i < ac; i++) {
                Material mi = m;  // AddMesh writes the texture ids into its material
                scene.AddMesh(av[i], mi, t, -1, -1);
            }
            if (!env_str("OXR_HEADLESS_ENV").empty()) scene.AddEnv(env_str("OXR_HEADLESS_ENV"));
            camera = Camera{vec3f(0.2f, 0.2f, 0.2f), scene.bounds.center(), vec3f(0.f, 1.f, 0.f)};
        } else {
            cornell(scene);
            camera = Camera{vec3f(0.0f, 1.0f, 3.0f), vec3f(0.0f, 1.0f, 0.0f), vec3f(0.0f, 1.0f, 0.0f)};
        }
        camera.medium = -1;
        SceneParser parser;
        bool ok = false;
        if (!env_str("OXR_HEADLESS_MAIN_SCENE").empty()) {
            Scene own;
            main_scene(env_str("OXR_HEADLESS_MAIN_SCENE"), own, camera, parser, ok);
            scene.meshes.swap(own.meshes);
            scene.textures.swap(own.textures);
            scene.lights.swap(own.lights);
            scene.mediums.swap(own.mediums);
            std::swap(scene.env, own.env);
            scene.bounds = own.bounds;
        } else if (!settings.empty()) {
            Scene parsed;
            parser.LoadSceneFromJson(settings, parsed, ok);
            if (ok && !parsed.meshes.empty()) {  // a whole scene: it replaces the default one
                scene.meshes.swap(parsed.meshes);
                scene.textures.swap(parsed.textures);
                scene.lights.swap(parsed.lights);
                scene.mediums.swap(parsed.mediums);
                std::swap(scene.env, parsed.env);
                scene.bounds = parsed.bounds;
                if (parser.hasCamera) camera = parser.camera;
                else camera = Camera{vec3f(0.2f, 0.2f, 0.2f), scene.bounds.center(), vec3f(0.f, 1.f, 0.f)};
            }
        }
        if (!ok) { parser.width = 256; parser.height = 256; }
        const std::string dump = env_str("OXR_HEADLESS_DUMP");
        if (!dump.empty()) {
            dump_scene(dump, scene, parser, camera, env_str("OXR_HEADLESS_DUMP_BRIEF") == "1");
            return 0;
        }
        Renderer renderer(&scene);
        renderer.Resize(vec2i(parser.width, parser.height));
        renderer.SetCamera(camera);
        renderer.progressive = true;
        const char* dn = std::getenv("OXR_HEADLESS_DENOISE");
        renderer.denoiserOn = dn && std::string(dn) == "1";
        renderer.presentOn = env_str("OXR_HEADLESS_PRESENT") == "1";  // streamed readback
        renderer.launchParams.filterType = parser.filterType;
        std::vector<uint32_t> pixels((size_t)parser.width * parser.height);
        for (int f = 0; f < frames; f++) renderer.Render();
        renderer.DownloadPixels(pixels.data());
        oxr_render_stats st = renderer.LastStats();
        FILE* fp = std::fopen(out.c_str(), "wb");
        if (!fp) throw std::runtime_error("cannot write " + out);
        std::fprintf(fp, "P6\n%d %d\n255\n", parser.width, parser.height);
        for (int y = parser.height - 1; y >= 0; y--)
            for (int x = 0; x < parser.width; x++) {
                uint32_t px = pixels[(size_t)y * parser.width + x];
                unsigned char rgb[3] = {(unsigned char)(px & 255), (unsigned char)((px >> 8) & 255),
                                        (unsigned char)((px >> 16) & 255)};
                std::fwrite(rgb, 1, 3, fp);
            }
        std::fclose(fp);
        std::printf("{\"frames\": %d, \"frame_id\": %d, \"width\": %d, \"height\": %d, \"bvh_ms\": %.3f, "
                    "\"last_frame_ms\": %.3f, \"rays\": %llu}\n",
                    frames, st.frame_id, parser.width, parser.height, renderer.BvhBuildMs(), st.render_ms,
                    (unsigned long long)(st.extension_rays + st.shadow_rays));
    } catch (std::runtime_error& e) {
        std::cout << "FATAL ERROR: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
