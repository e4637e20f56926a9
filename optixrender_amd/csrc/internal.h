// Internal host-side state of the C ABI: device scene, context, and the
// helpers shared by api.cpp (renderer) and serve.cpp (request server).
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <map>
#include <mutex>
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
    uint32_t class_mask = 1;  // shade classes the scene can produce (bit c for class c, kShadeClasses)
    oxr_light* lights = nullptr;
    int n_lights = 0;
    oxr_medium* media = nullptr;
    int n_media = 0;
    float4* env = nullptr;
    float4* env_cache = nullptr;
    float4* env_pdf = nullptr;  // (env rgb, cache pdf) per texel
    int env_w = 0, env_h = 0;
    float env_lum_sum = 0.0f;  // sequential luminance sum of the environment (host, at upload)
    BvhBuffers bvh;
    float4* cover_boxes = nullptr;  // cover set of the BVH (2 float4 per box, kCoverBoxes at most)
    // material textures (only when a material names one): RGBA8 texels of all
    // images, (first texel, width, height, 0) per image, texture coordinates
    // (6 floats per triangle) in upload and in sorted order
    uint32_t* texels = nullptr;
    int4* tex_info = nullptr;
    float* uvs = nullptr;
    float* uvs_sorted = nullptr;
    bool textured = false;
    int* cover_count = nullptr;
    bool built = false;
    float build_ms = 0.0f;
    // server requests (submitted and not yet let go by the server's worker)
    // that render this scene: oxr_scene_release refuses the scene meanwhile
    std::atomic<int> server_refs{0};
};


// Renderer::SetCamera (Renderer.cpp:796-813) on the host.
struct CamBasis { float pos[3], dir[3], h[3], v[3]; };
CamBasis camera_basis(const oxr_camera& c, int w, int h);
// Scene part of the wavefront parameters.
void fill_scene_params(WaveParams& p, const SceneDev* s, const float* lut_bsdf, const float* lut_albedo);

// Path pool, shard queues, per-sample window outputs and captured iteration
// graphs of one stream (wavefront.cpp).
struct WaveEngine {
    int64_t pool = 0, window = 0;  // allocated slots / samples
    int grid = 0;                  // trace grid the stack spill is sized for
    float4* path = nullptr;  // 4 float4 per slot
    float4 *hits = nullptr, *sh_env = nullptr, *sh_light = nullptr;
    float4 *s_rad = nullptr, *s_alb = nullptr, *s_nrm = nullptr;
    float* s_dep = nullptr;
    uint32_t *q0 = nullptr, *q1 = nullptr, *gstack = nullptr, *cls_queue = nullptr;
    uint32_t *t0 = nullptr, *t1 = nullptr;    // pending-finish lists per parity
    uint32_t *sl0 = nullptr, *sl1 = nullptr;  // shadow-ray slot lists per parity
    uint32_t *g0 = nullptr, *g1 = nullptr;  // sample ids of the new-path regions
    uint32_t* samples = nullptr;            // per-shard lists of the window's samples to trace
    int* ctr = nullptr;
    int* overflow = nullptr;
    unsigned long long* trav = nullptr;
    int* host_counts = nullptr;  // pinned: two copies of the polled counters
    hipEvent_t poll[2] = {nullptr, nullptr};
    // the sky shading runs on `side` beside the path loop (default priority)
    hipStream_t side = nullptr;
    hipEvent_t side_start = nullptr, side_done = nullptr;
    std::vector<hipEvent_t> stage_events;
    struct Graph {
        std::vector<uint8_t> key;
        hipGraphExec_t exec = nullptr;
    };
    std::vector<Graph> graphs;
};

// One window: launch shape, options, and what it measured.
struct WindowRun {
    int trace_grid = 0, shade_grid = 0;
    int sky_grid = 0;  // the sky shading beside the path loop
    int iters_per_poll = 2;
    bool count = false, eager = false, stage_timing = false;
    bool set_req0 = false;  // write frame0/offset/frames into request 0 on the device
    int frame0 = 0, path_frame_offset = 0, frames = 0;
    int64_t total_samples = 0, npix_total = 0;
    double stage_ms[5] = {0, 0, 0, 0, 0};  // regen (with window start), trace, shade (with tail), -, accumulate
    int launches = 0, iterations = 0;
};

int64_t wave_round_pool(int64_t paths);   // to a positive multiple of kShards * 64
int64_t wave_shard_cap(int64_t window);   // list capacity of one shard for a window
int wave_grid(int blocks);                // to a multiple of the blocks that give every shard a wave
size_t wave_bytes(int64_t pool, int64_t window, int trace_grid);
hipError_t wave_reserve(WaveEngine& e, int64_t pool, int64_t window, int trace_grid);
void wave_release(WaveEngine& e);
void wave_drop_graphs(WaveEngine& e);
void wave_bind(const WaveEngine& e, int64_t pool, WaveParams& p);
hipError_t wave_run_window(WaveEngine& e, const WaveParams& p, WindowRun& r, hipStream_t s);

}  // namespace oxr

struct oxr_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cu = 256;
    std::string err;
    std::map<uint32_t, oxr::SceneDev*> scenes;
    std::mutex scenes_mu;  // guards `scenes` against a server's submit on another thread
    uint32_t next_id = 1;
    uint32_t cur = 0;
    float* lut_bsdf = nullptr;
    float* lut_albedo = nullptr;
    int width = 0, height = 0;
    float4* color = nullptr;
    float4* albedo = nullptr;
    float4* normal = nullptr;
    float* depth = nullptr;
    uint32_t* rgba = nullptr;
    oxr_camera cam;
    int frame_id = 0;
    oxr::WaveEngine eng;
    int trace_grid = 0, shade_grid = 0, sky_grid = 0;
    oxr_tuning tuning;
    float* path_radiance = nullptr;
    size_t path_radiance_floats = 0;
    size_t path_radiance_valid = 0;
    oxr_render_stats stats;
    hipEvent_t t0 = nullptr, t1 = nullptr, td0 = nullptr, td1 = nullptr;
    oxr::ReqDesc* reqs = nullptr;  // device request table (single renders use entry 0)
    uint32_t* noalloc_scratch = nullptr;  // this device's copy of the fixed scratch (oxr_trace_rays_noalloc)
    uint8_t* cover = nullptr;      // camera-ray cover mask of single renders
    size_t cover_tiles = 0;        // its capacity
    // denoiser: weights (oxr_denoiser_load) and output of the last denoised render
    uint16_t* dn_lanes = nullptr;
    float* dn_bias = nullptr;
    oxr::DenoiserDev dn{};
    bool dn_loaded = false;
    float4* denoised = nullptr;
    size_t dn_pixels = 0;  // capacity of denoised
    bool denoised_valid = false;
    // streamed readback (OXR_FLAG_PRESENT): two page-locked frames written in
    // turn on copy_stream; present_done[k] marks frame k's copy complete
    hipStream_t copy_stream = nullptr;
    uint32_t* present_host[2] = {nullptr, nullptr};
    hipEvent_t present_done[2] = {nullptr, nullptr};
    int present_frame[2] = {0, 0};
    int present_next = 0;      // buffer the next present render writes
    int present_newest = -1;   // buffer of the newest present render, -1 none
    bool present_busy[2] = {false, false};  // a copy into k was enqueued (rgba must not change before it ends)
};
