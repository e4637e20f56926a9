// Internal host-side state of the C ABI: device scene, context, and the
// helpers shared by api.cpp (renderer) and serve.cpp (request server).
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "common.h"

namespace oxr {

struct SceneDev {
    int64_t n = 0;
    float* positions = nullptr;
    float* normals = nullptr;
    uint32_t* mat_ids = nullptr;
    float* normals_sorted = nullptr;
    uint32_t* mat_sorted = nullptr;
    oxr_material* mats = nullptr;
    int n_mats = 0;
    oxr_light* lights = nullptr;
    int n_lights = 0;
    oxr_medium* media = nullptr;
    int n_media = 0;
    float4* env = nullptr;
    float4* env_cache = nullptr;
    int env_w = 0, env_h = 0;
    BvhBuffers bvh;
    bool built = false;
    float build_ms = 0.0f;
};


// Renderer::SetCamera (Renderer.cpp:796-813) on the host.
struct CamBasis { float pos[3], dir[3], h[3], v[3]; };
CamBasis camera_basis(const oxr_camera& c, int w, int h);
// Scene part of the wavefront parameters.
void fill_scene_params(WaveParams& p, const SceneDev* s, const float* lut_bsdf, const float* lut_albedo);

}  // namespace oxr

struct oxr_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cu = 256;
    std::string err;
    std::map<uint32_t, oxr::SceneDev*> scenes;
    uint32_t next_id = 1;
    uint32_t cur = 0;
    float* lut_bsdf = nullptr;
    float* lut_albedo = nullptr;
    int width = 0, height = 0;
    float4* color = nullptr;
    float4* albedo = nullptr;
    float4* normal = nullptr;
    uint32_t* rgba = nullptr;
    oxr_camera cam;
    int frame_id = 0;
    // wavefront buffers
    int64_t cap = 0;
    float4 *ray_o = nullptr, *ray_d = nullptr, *thr = nullptr, *rad = nullptr, *aov_alb = nullptr, *aov_nrm = nullptr;
    float4 *hits = nullptr, *sh_env = nullptr, *sh_light = nullptr;
    uint32_t *q0 = nullptr, *q1 = nullptr;
    int* ctr = nullptr;
    unsigned long long* trav = nullptr;
    uint32_t* gstack = nullptr;
    int* overflow = nullptr;
    int trace_grid = 0, shade_grid = 0;
    oxr_tuning tuning;
    float* path_radiance = nullptr;
    size_t path_radiance_floats = 0;
    size_t path_radiance_valid = 0;
    oxr_render_stats stats;
    std::vector<hipEvent_t> events;
    oxr::ReqDesc* reqs = nullptr;  // device request table (single renders use entry 0)
    // captured batch graphs, keyed by the launch parameters they bake in
    struct Graph {
        std::vector<uint8_t> key;
        hipGraphExec_t exec = nullptr;
        std::vector<hipEvent_t> events;  // recorded between stages inside the graph
        std::vector<int> stages;
    };
    std::vector<Graph> graphs;
};
