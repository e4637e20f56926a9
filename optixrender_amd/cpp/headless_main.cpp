// Headless driver: the viewer-free counterpart of the reference's main.cpp
// (main.cpp:165-266).  Builds a scene through the same Scene / SceneParser /
// Renderer calls main.cpp makes, renders frames and writes the RGBA8 pixels
// (DownloadPixels) as a binary PPM.
//
//   headless [settings.json] [out.ppm] [frames] [obj...]
// Without OBJ arguments the scene is a procedural Cornell box.  With
// OXR_HEADLESS_DENOISE=1 in the environment the frames are denoised
// (denoiserOn, the viewer's 'D' key).  OBJ meshes take their textures from
// OXR_HEADLESS_ALBEDO / _ROUGHNESS / _METALLIC / _NORMAL (PPM files), and
// OXR_HEADLESS_ENV names an environment map (.hdr or .pfm), as main.cpp
// gives its meshes texture files and its scene an HDR (main.cpp:165-213).
#include <cstdio>
#include <cstdlib>
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
        if (!settings.empty()) parser.LoadFromJson(settings, scene, ok);
        if (!ok) { parser.width = 256; parser.height = 256; }
        Renderer renderer(&scene);
        renderer.Resize(vec2i(parser.width, parser.height));
        renderer.SetCamera(camera);
        renderer.progressive = true;
        const char* dn = std::getenv("OXR_HEADLESS_DENOISE");
        renderer.denoiserOn = dn && std::string(dn) == "1";
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
