// Renderer over the C ABI.  Construction mirrors Renderer::Renderer
// (core/src/Renderer.cpp:39-93): upload lights, media, meshes, textures and
// the environment, then build the acceleration structure (here an LBVH on
// the device).  One material slot per mesh, as one SBT record per mesh was
// (Renderer.cpp:584-627).
#include "Renderer.h"

#include <unistd.h>

#include <cstdlib>
#include <cstring>
#include <fstream>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace oxr_cpp {
oxr_ctx* default_context() {
    static oxr_ctx* ctx = nullptr;
    static std::once_flag once;
    std::call_once(once, [] {
        if (oxr_create(0, &ctx) != OXR_OK) throw std::runtime_error("oxr_create(0) failed: no HIP device");
    });
    return ctx;
}
}  // namespace oxr_cpp

namespace {

void copy3(float* d, const vec3f& v) { d[0] = v.x; d[1] = v.y; d[2] = v.z; }

oxr_material to_abi(const Material& m, int in_med, int out_med) {
    oxr_material r;
    std::memset(&r, 0, sizeof(r));
    r.type = (int32_t)m.type;
    copy3(r.albedo, m.albedo);
    r.roughness = m.roughness;
    r.anisotropy = m.anisotropy;
    copy3(r.eta, m.eta);
    copy3(r.k, m.k);
    r.int_ior = m.int_ior;
    r.ext_ior = m.ext_ior;
    r.nonlinear = m.nonlinear ? 1 : 0;
    copy3(r.specular, m.specular);
    r.metallic = m.metallic;
    r.coat_weight = m.coat_weight;
    r.coat_roughness_u = m.coat_roughness_u;
    r.coat_roughness_v = m.coat_roughness_v;
    r.in_medium = in_med;
    r.out_medium = out_med;
    const int ids[8] = {m.albedoTextureID, m.roughnessTextureID, m.anisotropyTextureID, m.specularTextureID,
                        m.metallicTextureID, m.coat_roughness_uTextureID, m.coat_roughness_vTextureID,
                        m.normalTextureID};
    for (int i = 0; i < 8; i++) r.tex[i] = ids[i];
    return r;
}

}  // namespace

Renderer::Renderer(const Scene* s) : scene(s) {
    oxr_ctx* ctx = oxr_cpp::default_context();
    std::vector<float> pos, nrm, uvs;
    std::vector<uint32_t> mid;
    std::vector<oxr_material> mats;
    for (size_t m = 0; m < scene->meshes.size(); m++) {
        const TriangleMesh* mesh = scene->meshes[m];
        mats.push_back(to_abi(mesh->material, mesh->in_medium, mesh->out_medium));
        for (const vec3i& t : mesh->index) {
            for (int k : {t.x, t.y, t.z}) {
                const vec3f& p = mesh->vertex[k];
                const vec3f& n = mesh->normal[k];
                const vec2f& t = mesh->texcoord[k];
                pos.insert(pos.end(), {p.x, p.y, p.z});
                nrm.insert(nrm.end(), {n.x, n.y, n.z});
                uvs.insert(uvs.end(), {t.x, t.y});
            }
            mid.push_back((uint32_t)m);
        }
    }
    std::vector<oxr_light> lights;
    for (const Light& l : scene->lights) {
        oxr_light r;
        std::memset(&r, 0, sizeof(r));
        r.type = (int32_t)l.type;
        r.double_side = l.doubleSide ? 1 : 0;
        copy3(r.radiance, l.radiance);
        copy3(r.position, l.position);
        copy3(r.u, l.u);
        copy3(r.v, l.v);
        r.radius = l.radius;
        r.medium = l.medium;
        lights.push_back(r);
    }
    std::vector<oxr_medium> media;
    for (const Medium& m : scene->mediums) {
        oxr_medium r;
        std::memset(&r, 0, sizeof(r));
        r.type = (int32_t)m.type;
        r.phase_type = (int32_t)m.phase.type;
        copy3(r.g, m.phase.g);
        copy3(r.sigma_s, m.sigma_s);
        copy3(r.sigma_a, m.sigma_a);
        r.scale = m.scale;
        media.push_back(r);
    }
    std::vector<oxr_texture> textures;  // CreateTextures (Renderer.cpp:95-144)
    for (const Texture* t : scene->textures) {
        oxr_texture r;
        r.width = t->resolution.x;
        r.height = t->resolution.y;
        r.rgba8 = t->pixel.data();
        textures.push_back(r);
    }
    oxr_scene_desc d;
    std::memset(&d, 0, sizeof(d));
    d.n_tris = (int64_t)mid.size();
    d.positions = pos.data();
    d.normals = nrm.data();
    d.uvs = uvs.data();
    d.n_textures = (int32_t)textures.size();
    d.textures = textures.empty() ? nullptr : textures.data();
    d.mat_ids = mid.data();
    d.n_materials = (int32_t)mats.size();
    d.materials = mats.data();
    d.n_lights = (int32_t)lights.size();
    d.lights = lights.data();
    d.n_media = (int32_t)media.size();
    d.media = media.data();
    if (scene->env && scene->env->width > 0) {  // an unreadable HDR leaves no environment
        d.env_w = scene->env->width;
        d.env_h = scene->env->height;
        d.env_rgba = scene->env->hdr.data();
    }
    OXR_CHECK(oxr_scene_upload(ctx, &d, &sceneId));
    OXR_CHECK(oxr_scene_build_bvh(ctx, sceneId, &buildMs));
    OXR_CHECK(oxr_scene_select(ctx, sceneId));
}

Renderer::~Renderer() {
    if (sceneId) oxr_scene_release(oxr_cpp::default_context(), sceneId);
}

void Renderer::Render() {
    if (launchParams.frame.size.x == 0) return;  // Renderer.cpp:638-640
    oxr_ctx* ctx = oxr_cpp::default_context();
    OXR_CHECK(oxr_scene_select(ctx, sceneId));
    oxr_render_params p;
    std::memset(&p, 0, sizeof(p));
    p.spp = launchParams.numPixelSamples;
    p.max_depth = launchParams.maxBounce;
    p.rr_start = launchParams.rrStart;
    p.progressive = progressive ? 1 : 0;
    p.filter = (int32_t)launchParams.filterType;
    if (denoiserOn && !denoiserLoaded) {
        const char* env = std::getenv("OXR_DENOISER_WEIGHTS");
        std::string path = env ? env : "";
        if (path.empty()) {  // next to the executable
            char exe[4096];
            ssize_t n = readlink("/proc/self/exe", exe, sizeof(exe) - 1);
            std::string dir = n > 0 ? std::string(exe, (size_t)n) : std::string("./");
            path = dir.substr(0, dir.find_last_of('/') + 1) + "denoiser_kpcnn.f32";
        }
        LoadDenoiser(path);
    }
    p.denoise = denoiserOn ? 1 : 0;  // Renderer.cpp:660-785 (our KPCNN in place of the OptiX model)
    p.flags = presentOn ? OXR_FLAG_PRESENT : 0u;
    OXR_CHECK(oxr_render(ctx, &p));
    launchParams.frame.frameID = LastStats().frame_id;
}

void Renderer::LoadDenoiser(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open denoiser weights " + path);
    std::vector<float> w(OXR_DENOISER_PARAMS);
    f.read(reinterpret_cast<char*>(w.data()), (std::streamsize)(w.size() * sizeof(float)));
    if (f.gcount() != (std::streamsize)(w.size() * sizeof(float)))
        throw std::runtime_error("denoiser weights " + path + ": expected " + std::to_string(w.size()) + " floats");
    OXR_CHECK(oxr_denoiser_load(oxr_cpp::default_context(), w.data(), (int64_t)w.size()));
    denoiserLoaded = true;
}

void Renderer::SetCamera(const Camera& camera) {
    lastSetCamera = camera;
    cameraSet = true;
    launchParams.frame.frameID = 0;  // Renderer.cpp:799
    oxr_camera c;
    copy3(c.from, camera.from);
    copy3(c.at, camera.at);
    copy3(c.up, camera.up);
    c.medium = camera.medium;
    c.cos_fovy = 0.66f;  // Renderer.cpp:803
    OXR_CHECK(oxr_set_camera(oxr_cpp::default_context(), &c));
}

void Renderer::Resize(const vec2i& newSize) {
    OXR_CHECK(oxr_resize(oxr_cpp::default_context(), newSize.x, newSize.y));
    launchParams.frame.size = newSize;
    if (cameraSet) SetCamera(lastSetCamera);  // Renderer.cpp:866
}

void Renderer::DownloadPixels(uint32_t h_pixels[]) {
    size_t n = (size_t)launchParams.frame.size.x * launchParams.frame.size.y;
    if (presentOn) {
        std::memcpy(h_pixels, PresentPixels(), 4 * n);
        return;
    }
    OXR_CHECK(oxr_download_rgba8(oxr_cpp::default_context(), h_pixels, n));
}

const uint32_t* Renderer::PresentPixels() {
    const uint32_t* px = nullptr;
    OXR_CHECK(oxr_present(oxr_cpp::default_context(), &px, nullptr, nullptr, nullptr));
    return px;
}

oxr_render_stats Renderer::LastStats() const {
    oxr_render_stats s;
    OXR_CHECK(oxr_get_stats(oxr_cpp::default_context(), &s));
    return s;
}
